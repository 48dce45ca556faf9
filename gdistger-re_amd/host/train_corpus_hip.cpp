// train_corpus_hip.cpp - the trainer entry point of the reference
// (train_corpus_cuda, src/word2vec.cu:1429-1600) over the C ABI.
#include "../../include/train_corpus_hip.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

// ArgPos (src/word2vec.cu:1418-1428): index of `flag` in argv, -1 if absent,
// -2 if it is the last argument (no value).
int arg_pos(const char* flag, int argc, char** argv) {
  for (int a = 1; a < argc; ++a)
    if (!std::strcmp(flag, argv[a])) return a == argc - 1 ? -2 : a;
  return -1;
}

bool parse_u32(const char* s, uint32_t* v) {
  errno = 0;
  char* end = nullptr;
  const long long x = std::strtoll(s, &end, 10);
  if (errno || end == s || *end || x < 0 || x > 0xffffffffLL) return false;
  *v = (uint32_t)x;
  return true;
}

bool parse_f32(const char* s, float* v) {
  errno = 0;
  char* end = nullptr;
  const double x = std::strtod(s, &end);
  if (errno || end == s || *end) return false;
  *v = (float)x;
  return true;
}

}  // namespace

int dg_parse_train_flags(int argc, char** argv, dg_train_flags* out, std::string* err) {
  dg_sgns_params_default(&out->p);
  struct U32Flag {
    const char* name;
    uint32_t* dst;
  };
  uint32_t cbow = 0, hs = 0, reuse_neg = 0, threads = 0, binary = 0, debug = 0, per_file = 0;
  uint32_t seed = 0, form = out->p.form, deterministic = 0;
  const U32Flag u32s[] = {
      {"-size", &out->p.dim},         {"-window", &out->p.window},
      {"-negative", &out->p.negative}, {"-min-count", &out->p.min_count},
      {"-iter", &out->p.iter},         {"-batch-size", &out->p.batch_size},
      {"-cbow", &cbow},                {"-hs", &hs},
      {"-reuse-neg", &reuse_neg},      {"-threads", &threads},
      {"-binary", &binary},            {"-debug", &debug},
      {"-alpha-per-file", &per_file},  {"-seed", &seed},
      {"-form", &form},                {"-deterministic", &deterministic},
      {"-negative-draws", &out->p.negative_draws},
  };
  for (const auto& f : u32s) {
    const int i = arg_pos(f.name, argc, argv);
    if (i == -2) {
      *err = std::string("Argument missing for ") + f.name;
      return DG_ERR_ARG;
    }
    if (i > 0 && !parse_u32(argv[i + 1], f.dst)) {
      *err = std::string("bad value for ") + f.name + ": " + argv[i + 1];
      return DG_ERR_ARG;
    }
  }
  struct F32Flag {
    const char* name;
    float* dst;
  };
  const F32Flag f32s[] = {{"-alpha", &out->p.alpha}, {"-sample", &out->p.sample}};
  for (const auto& f : f32s) {
    const int i = arg_pos(f.name, argc, argv);
    if (i == -2) {
      *err = std::string("Argument missing for ") + f.name;
      return DG_ERR_ARG;
    }
    if (i > 0 && !parse_f32(argv[i + 1], f.dst)) {
      *err = std::string("bad value for ") + f.name + ": " + argv[i + 1];
      return DG_ERR_ARG;
    }
  }
  for (const char* name : {"-emb_output", "-eoutput"}) {
    const int i = arg_pos(name, argc, argv);
    if (i == -2) {
      *err = std::string("Argument missing for ") + name;
      return DG_ERR_ARG;
    }
    if (i > 0) out->emb_output = argv[i + 1];
  }
  if (cbow != 0) {
    *err = "-cbow 1: only skip-gram is implemented";
    return DG_ERR_ARG;
  }
  if (hs != 0) {
    *err = "-hs 1: only negative sampling is implemented";
    return DG_ERR_ARG;
  }
  if (seed) out->p.seed = seed;
  out->p.form = form;
  out->p.deterministic = deterministic;
  out->binary = (int)binary;
  out->debug = (int)debug;
  out->per_file_alpha = (int)per_file;
  if (per_file) {
    // the GPU trainer's per-file schedule starts at 0.025 whatever -alpha
    // says (src/word2vec.cu:1279); here -alpha, when given, sets the start
    out->p.alpha_schedule = DG_ALPHA_FIXED;
    if (arg_pos("-alpha", argc, argv) < 0) out->p.alpha = 0.025f;
  }
  return DG_OK;
}

int train_corpus_hip(int argc, char** argv, const std::vector<uint32_t>& degrees, dg_ctx* ctx,
                     int rank) {
  (void)rank;
  if (!ctx) return DG_ERR_ARG;
  dg_train_flags fl;
  std::string err;
  int rc = dg_parse_train_flags(argc, argv, &fl, &err);
  if (rc != DG_OK) {
    std::fprintf(stderr, "train_corpus_hip: %s\n", err.c_str());
    return rc;
  }
  uint32_t v_num = 0;
  uint64_t m = 0;
  if ((rc = dg_graph_info(ctx, &v_num, &m)) != DG_OK) return rc;
  if (!degrees.empty()) {
    // the device vocabulary counts in + out degree of its CSR rows
    if (degrees.size() != v_num) {
      std::fprintf(stderr, "train_corpus_hip: %zu degrees for %u vertices\n", degrees.size(), v_num);
      return DG_ERR_ARG;
    }
    std::vector<uint64_t> row_ptr(v_num + 1);
    if ((rc = dg_graph_get_csr(ctx, row_ptr.data(), nullptr, nullptr)) != DG_OK) return rc;
    for (uint32_t v = 0; v < v_num; ++v)
      if ((uint64_t)degrees[v] != 2 * (row_ptr[v + 1] - row_ptr[v])) {
        std::fprintf(stderr, "train_corpus_hip: degree of vertex %u differs from the device graph\n", v);
        return DG_ERR_ARG;
      }
  }
  auto phase = [&](const char* what) {
    if (fl.debug > 0) {
      std::fprintf(stderr, "train_corpus_hip: %s\n", what);
      std::fflush(stderr);
    }
  };
  phase("building the vocabulary and tables");
  if ((rc = dg_sgns_init(ctx, &fl.p)) != DG_OK) return rc;
  phase("training");
  dg_sgns_stats st;
  if ((rc = dg_sgns_train(ctx, DG_CORPUS_TRAIN, &fl.p, 0.0f, &st)) != DG_OK) return rc;
  if (fl.debug > 0)
    std::fprintf(stderr, "train_corpus_hip: %llu words, %llu kept, %.1f ms, alpha %.6f\n",
                 (unsigned long long)st.words, (unsigned long long)st.kept, st.ms, st.alpha_end);
  phase("writing the embeddings");
  if (!fl.emb_output.empty() && (rc = dg_emb_write(ctx, fl.emb_output.c_str(), fl.binary)) != DG_OK)
    return rc;
  return DG_OK;
}
