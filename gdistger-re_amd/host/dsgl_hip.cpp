// dsgl_hip.cpp - the CPU trainer entry point of the reference (dsgl,
// include/dsgl.hpp:1464-1580) over the C ABI, on the device's DSGL form.
#include "../../include/dsgl_hip.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

// ArgPos (dsgl.hpp:1429-1440): index of `flag` in argv, -1 if absent, -2 if
// it is the last argument (no value).
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

int dg_parse_dsgl_flags(int argc, char** argv, dg_dsgl_flags* out, std::string* err) {
  dg_sgns_params_default(&out->p);
  dg_sgns_params& p = out->p;
  // dsgl.hpp:67-75
  p.form = DG_SGNS_DSGL_FP32;
  p.vocab_source = DG_VOCAB_VISITS;
  p.init_float_ops = 1;
  p.alpha_schedule = DG_ALPHA_LINEAR;
  p.dim = 128;
  p.alpha = 0.1f;
  p.sample = 1e-3f;
  p.window = 5;
  p.negative = 5;
  p.workers = 12;
  p.iter = 5;
  p.min_count = 5;
  uint32_t binary = 0, debug = 2, batch = 0;
  struct U32Flag {
    const char* name;
    uint32_t* dst;
  };
  const U32Flag u32s[] = {
      {"-size", &p.dim},       {"-window", &p.window},   {"-negative", &p.negative},
      {"-threads", &p.workers}, {"-iter", &p.iter},       {"-min-count", &p.min_count},
      {"-batch-size", &batch}, {"-binary", &binary},     {"-debug", &debug},
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
  const F32Flag f32s[] = {{"-alpha", &p.alpha}, {"-sample", &p.sample}};
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
  const int oi = arg_pos("-eoutput", argc, argv);
  if (oi == -2) {
    *err = "Argument missing for -eoutput";
    return DG_ERR_ARG;
  }
  if (oi > 0) out->emb_output = argv[oi + 1];
  for (const char* name : {"-save-vocab", "-read-vocab", "-disk", "-sync-period", "-min-sync-words",
                           "-full-sync-times", "-message-size", "-save-emb"}) {
    const int i = arg_pos(name, argc, argv);
    if (i == -2) {
      *err = std::string("Argument missing for ") + name;
      return DG_ERR_ARG;
    }
    if (i > 0) out->ignored.push_back(name);
  }
  if (p.workers == 0) {
    *err = "-threads must be at least 1";
    return DG_ERR_ARG;
  }
  // DSGL sets batch_size = 2 * window + 1 whatever the flag says (dsgl.hpp:832)
  p.batch_size = 2 * p.window + 1;
  (void)batch;
  out->binary = (int)binary;
  out->debug = (int)debug;
  return DG_OK;
}

int dsgl_hip(int argc, char** argv, const std::vector<int>* vertex_cn, dg_ctx* ctx) {
  if (!ctx) return DG_ERR_ARG;
  dg_dsgl_flags fl;
  std::string err;
  int rc = dg_parse_dsgl_flags(argc, argv, &fl, &err);
  if (rc != DG_OK) {
    std::fprintf(stderr, "dsgl_hip: %s\n", err.c_str());
    return rc;
  }
  uint32_t v_num = 0;
  uint64_t m = 0;
  if ((rc = dg_graph_info(ctx, &v_num, &m)) != DG_OK) return rc;
  if (vertex_cn) {
    // the device ranks its vocabulary by its own cumulative visit counts
    if (vertex_cn->size() != v_num) {
      std::fprintf(stderr, "dsgl_hip: %zu visit counts for %u vertices\n", vertex_cn->size(), v_num);
      return DG_ERR_ARG;
    }
    std::vector<uint32_t> freq(v_num);
    if ((rc = dg_vertex_freq(ctx, freq.data())) != DG_OK) return rc;
    for (uint32_t v = 0; v < v_num; ++v)
      if ((int64_t)(*vertex_cn)[v] != (int64_t)freq[v]) {
        std::fprintf(stderr, "dsgl_hip: visit count of vertex %u differs from the device's\n", v);
        return DG_ERR_ARG;
      }
  }
  auto phase = [&](const char* what) {
    if (fl.debug > 2) {
      std::fprintf(stderr, "dsgl_hip: %s\n", what);
      std::fflush(stderr);
    }
  };
  for (const auto& f : fl.ignored)
    if (fl.debug > 0) std::fprintf(stderr, "dsgl_hip: %s has no effect on one GPU\n", f.c_str());
  phase("building the vocabulary and tables");
  if ((rc = dg_sgns_init(ctx, &fl.p)) != DG_OK) return rc;
  phase("training");
  dg_sgns_stats st;
  if ((rc = dg_sgns_train(ctx, DG_CORPUS_TRAIN, &fl.p, 0.0f, &st)) != DG_OK) return rc;
  if (fl.debug > 1)
    std::fprintf(stderr, "dsgl_hip: %llu words, %llu kept, %u compute threads, %.1f ms, alpha %.6f\n",
                 (unsigned long long)st.words, (unsigned long long)st.kept, fl.p.workers, st.ms,
                 st.alpha_end);
  phase("writing the embeddings");
  if (!fl.emb_output.empty() && (rc = dg_emb_write(ctx, fl.emb_output.c_str(), fl.binary)) != DG_OK)
    return rc;
  return DG_OK;
}
