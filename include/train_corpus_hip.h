// train_corpus_hip.h - C++ mirror of the reference trainer entry point over
// the C ABI of distger_hip.h.
//
// Replaces
//   int train_corpus_cuda(int argc, char** argv, const vector<vertex_id_t>& degrees,
//                         SyncQueue& corpus_q, int rank, myEdgeContainer* csr);
// (src/word2vec.cu:1429, declared at src/examples/huge_walk.cpp:13).  The walk
// corpus and the graph stay on the device behind `ctx` (the reference passes
// a queue of walk-file names and a host CSR); argv is scanned for the
// trainer's flags the way ArgPos does it (src/word2vec.cu:1418-1428): a flag
// anywhere in argv, its value in the next argument.  Errors are returned as
// DG_ERR_* codes with a message, never by exit().
#ifndef TRAIN_CORPUS_HIP_H
#define TRAIN_CORPUS_HIP_H

#include <cstdint>
#include <string>
#include <vector>

#include "distger_hip.h"

struct dg_train_flags {
  dg_sgns_params p;         // -size -window -negative -batch-size -sample -alpha -min-count -iter ...
  std::string emb_output;   // -emb_output / -eoutput: embedding file ("" = none)
  int binary = 0;           // -binary
  int debug = 0;            // -debug
  int per_file_alpha = 0;   // -alpha-per-file 1: alpha_n = alpha e^{-0.1 n} per round, alpha 0.025
                            // unless -alpha is given (src/lr_scheduler.hpp:23-31, word2vec.cu:1279)
};

// Scan argv for the trainer flags (src/word2vec.cu:1548-1566) into `out`,
// starting from dg_sgns_params_default.  Returns DG_OK, or DG_ERR_ARG with
// `err` set (a flag without a value, a value that does not parse, -cbow 1 or
// -hs 1, which this trainer does not implement).
int dg_parse_train_flags(int argc, char** argv, dg_train_flags* out, std::string* err);

// Train the appended device corpus of `ctx` with the flags of argv, then
// write the embedding file if -emb_output is given.  `degrees` is the
// per-vertex word count the reference builds its vocabulary from (in + out
// degree, huge_walk.cpp:43-46); it must match the device graph's degrees
// (the device builds the same vocabulary from its CSR), else DG_ERR_ARG.
// `rank` is kept for the reference's signature (one rank here).
int train_corpus_hip(int argc, char** argv, const std::vector<uint32_t>& degrees, dg_ctx* ctx,
                     int rank);

#endif
