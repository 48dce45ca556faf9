// dsgl_hip.h - C++ mirror of the reference's CPU trainer entry point over the
// C ABI of distger_hip.h.
//
// Replaces
//   int dsgl(int argc, char** argv, vector<int>* vertex_cn,
//            vector<vertex_id_t>* new_sort, WalkEngine<real_t, uint32_t>* graph);
// (include/dsgl.hpp:1464-1580).  The walk corpus (the reference's
// graph->local_corpus) and the visit counts stay on the device behind `ctx`;
// training runs the device's DSGL form (DG_SGNS_DSGL_FP32: two sentences per
// step, one snapshot, [target, negatives with multiplicity, target2]) over
// the appended corpus, with the vocabulary ranked by visit count
// (dsgl.hpp:314-382).  argv is scanned as ArgPos does it (dsgl.hpp:1429-1440):
// a flag anywhere in argv, its value in the next argument, the defaults
// those of dsgl.hpp:67-75.  `new_sort` is not taken: the reference reads it
// past its end when it is empty (dsgl.hpp:364).  Errors are returned as
// DG_ERR_* codes with a message, never by exit().
#ifndef DSGL_HIP_H
#define DSGL_HIP_H

#include <string>
#include <vector>

#include "distger_hip.h"

struct dg_dsgl_flags {
  dg_sgns_params p;        // -size -window -negative -sample -alpha -iter -min-count -threads -batch-size
  std::string emb_output;  // -eoutput: embedding file ("" = none)
  int binary = 0;          // -binary
  int debug = 2;           // -debug (dsgl.hpp:67)
  // accepted for the reference's command lines and without effect on one GPU:
  // -save-vocab -read-vocab -disk -sync-period -min-sync-words
  // -full-sync-times -message-size -save-emb (vocabulary files, streaming
  // from disk and the multi-process model sync)
  std::vector<std::string> ignored;
};

// Scan argv for DSGL's flags (dsgl.hpp:1496-1535) into `out`: DSGL's defaults
// (alpha 0.1, sample 1e-3, window 5, negative 5, 12 threads, 5 iterations,
// min-count 5) except the size, 128 instead of 100 (the device form's rows
// are 64, 128 or 256 wide), the device DSGL form, the visit-count
// vocabulary, float initialisation (dsgl.hpp:433-457), the linear rate
// (dsgl.hpp:886-897); -threads sets the compute threads the corpus is split
// over; -batch-size is recorded as 2 * window + 1, which DSGL forces
// (dsgl.hpp:832).  Returns DG_OK, or DG_ERR_ARG with `err` set.
int dg_parse_dsgl_flags(int argc, char** argv, dg_dsgl_flags* out, std::string* err);

// Train the appended device corpus of `ctx` with DSGL's flags of argv, then
// write the embedding file if -eoutput is given.  `vertex_cn` is the
// reference's per-vertex visit count; when not null it must equal the
// device's cumulative counts (DG_ERR_ARG otherwise), from which the device
// builds the same vocabulary.
int dsgl_hip(int argc, char** argv, const std::vector<int>* vertex_cn, dg_ctx* ctx);

#endif
