// huge_walk.cpp - command line of the reference's huge_walk
// (src/examples/huge_walk.cpp) on one MI355X: load the graph, run the
// information-oriented walk rounds, write each round's walk file, train SGNS
// on the corpus and write the embeddings.
//
// Walker flags as option_helper.hpp:112-156 names them: -g graph (binary
// edges, include/storage.hpp:15-38), -v vertex count, -w walkers per round,
// -o walk-file prefix ("<o>-0-<iter>.txt", walk.hpp:686), --min_L, --min_R,
// --make-undirected.  Trainer flags as train_corpus_cuda scans them
// (src/word2vec.cu:1548-1566): -size -window -negative -batch-size -sample
// -alpha -min-count -iter -emb_output/-eoutput -binary -debug, plus -form
// (0 plain FP32, 1 batched FP32, 2 batched BF16), -deterministic,
// -negative-draws, -alpha-per-file.  Extra walker flags: -seed, -max-length,
// -max-rounds, -rounds-mode (0: 30 rounds as walk.hpp:740-744, 1: also stop
// at |delta RE| <= 0.001), -cn (0: common-neighbour counts from the edge
// data as the reference reads them, 1: computed on the device).
// -trainer dsgl trains with the CPU trainer's entry point instead
// (dsgl_hip, include/dsgl.hpp:1464): its flags and defaults (-threads for
// the compute threads, -eoutput), the device's DSGL form.
// --print-params prints the parsed parameters and exits without a GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/distger_hip.h"
#include "../../include/dsgl_hip.h"
#include "../../include/train_corpus_hip.h"

namespace {

int arg_pos(const char* flag, int argc, char** argv) {
  for (int a = 1; a < argc; ++a)
    if (!std::strcmp(flag, argv[a])) return a == argc - 1 ? -2 : a;
  return -1;
}

bool has_flag(const char* flag, int argc, char** argv) {
  for (int a = 1; a < argc; ++a)
    if (!std::strcmp(flag, argv[a])) return true;
  return false;
}

// value of a numeric flag, `def` when absent; false on a missing or bad value
bool num_flag(const char* flag, int argc, char** argv, unsigned long long def,
              unsigned long long* out) {
  const int i = arg_pos(flag, argc, argv);
  if (i == -2) {
    std::fprintf(stderr, "Argument missing for %s\n", flag);
    return false;
  }
  if (i < 0) {
    *out = def;
    return true;
  }
  char* end = nullptr;
  *out = std::strtoull(argv[i + 1], &end, 10);
  if (end == argv[i + 1] || *end) {
    std::fprintf(stderr, "bad value for %s: %s\n", flag, argv[i + 1]);
    return false;
  }
  return true;
}

// With -debug: one line per phase on stderr, flushed, so a run that stops
// shows where.
bool g_debug = false;
void phase(const char* what) {
  if (!g_debug) return;
  std::fprintf(stderr, "huge_walk: %s\n", what);
  std::fflush(stderr);
}

int fail(dg_ctx* ctx, int rc, const char* what) {
  std::fprintf(stderr, "huge_walk: %s: %s (%d)\n", what, ctx ? dg_last_error(ctx) : "", rc);
  if (ctx) dg_close(ctx);
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2 || has_flag("-h", argc, argv) || has_flag("--help", argc, argv)) {
    std::printf("usage: huge_walk -g graph.bin -v V [-w walkers] [-o walk_prefix] [--min_L 20] "
                "[--min_R 5] [--make-undirected] [-trainer gpu|dsgl] [-size 128 -window 10 -negative 5 ...] "
                "[-emb_output emb.txt]\n");
    return argc < 2 ? 1 : 0;
  }
  const int gi = arg_pos("-g", argc, argv);
  unsigned long long v_num = 0, walkers = 0, min_L = 0, min_R = 0, seed = 0, max_len = 0;
  unsigned long long max_rounds = 0, rounds_mode = 0, cn = 0;
  if (gi == -2 || !num_flag("-v", argc, argv, 0, &v_num) || !num_flag("-w", argc, argv, 0, &walkers) ||
      !num_flag("--min_L", argc, argv, 20, &min_L) || !num_flag("--min_R", argc, argv, 5, &min_R) ||
      !num_flag("-seed", argc, argv, 1, &seed) || !num_flag("-max-length", argc, argv, 0, &max_len) ||
      !num_flag("-max-rounds", argc, argv, 30, &max_rounds) ||
      !num_flag("-rounds-mode", argc, argv, 0, &rounds_mode) || !num_flag("-cn", argc, argv, 0, &cn))
    return 1;
  const int oi = arg_pos("-o", argc, argv);
  if (oi == -2) {
    std::fprintf(stderr, "Argument missing for -o\n");
    return 1;
  }
  const bool undirected = has_flag("--make-undirected", argc, argv);
  // -trainer gpu (train_corpus_cuda's flags, src/word2vec.cu:1548-1566) or
  // dsgl (the CPU trainer's flags and defaults, include/dsgl.hpp:1496-1535)
  const int ti = arg_pos("-trainer", argc, argv);
  if (ti == -2 || (ti > 0 && std::strcmp(argv[ti + 1], "gpu") && std::strcmp(argv[ti + 1], "dsgl"))) {
    std::fprintf(stderr, "huge_walk: -trainer takes gpu or dsgl\n");
    return 1;
  }
  const bool use_dsgl = ti > 0 && !std::strcmp(argv[ti + 1], "dsgl");
  dg_train_flags tf;
  dg_dsgl_flags df;
  std::string err;
  if (dg_parse_train_flags(argc, argv, &tf, &err) != DG_OK ||
      (use_dsgl && dg_parse_dsgl_flags(argc, argv, &df, &err) != DG_OK)) {
    std::fprintf(stderr, "huge_walk: %s\n", err.c_str());
    return 1;
  }
  if (use_dsgl && tf.per_file_alpha) {
    std::fprintf(stderr, "huge_walk: -alpha-per-file is the GPU trainer's schedule (-trainer gpu)\n");
    return 1;
  }
  if (gi < 0 || v_num == 0) {
    std::fprintf(stderr, "huge_walk: -g and -v are required\n");
    return 1;
  }
  if (walkers == 0) walkers = v_num;
  if (has_flag("--print-params", argc, argv) && use_dsgl) {
    std::printf("{\"trainer\": \"dsgl\", \"graph\": \"%s\", \"v_num\": %llu, \"dim\": %u, \"window\": %u, "
                "\"negative\": %u, \"min_count\": %u, \"iter\": %u, \"batch_size\": %u, "
                "\"sample\": %g, \"alpha\": %g, \"threads\": %u, \"form\": %u, \"vocab_source\": %u, "
                "\"eoutput\": \"%s\", \"binary\": %d, \"debug\": %d, \"ignored\": %zu}\n",
                argv[gi + 1], v_num, df.p.dim, df.p.window, df.p.negative, df.p.min_count, df.p.iter,
                df.p.batch_size, (double)df.p.sample, (double)df.p.alpha, df.p.workers, df.p.form,
                df.p.vocab_source, df.emb_output.c_str(), df.binary, df.debug, df.ignored.size());
    return 0;
  }
  if (has_flag("--print-params", argc, argv)) {
    std::printf("{\"graph\": \"%s\", \"v_num\": %llu, \"walkers\": %llu, \"min_L\": %llu, \"min_R\": %llu, "
                "\"undirected\": %d, \"seed\": %llu, \"max_length\": %llu, \"max_rounds\": %llu, "
                "\"rounds_mode\": %llu, \"cn\": %llu, \"output\": \"%s\", \"dim\": %u, \"window\": %u, "
                "\"negative\": %u, \"min_count\": %u, \"iter\": %u, \"batch_size\": %u, "
                "\"sample\": %g, \"alpha\": %g, \"form\": %u, \"deterministic\": %u, "
                "\"emb_output\": \"%s\", \"binary\": %d, \"alpha_per_file\": %d}\n",
                argv[gi + 1], v_num, walkers, min_L, min_R, (int)undirected, seed, max_len, max_rounds,
                rounds_mode, cn, oi > 0 ? argv[oi + 1] : "", tf.p.dim, tf.p.window, tf.p.negative,
                tf.p.min_count, tf.p.iter, tf.p.batch_size, (double)tf.p.sample, (double)tf.p.alpha,
                tf.p.form, tf.p.deterministic, tf.emb_output.c_str(), tf.binary, tf.per_file_alpha);
    return 0;
  }

  g_debug = tf.debug > 0;
  phase("opening the device");
  dg_ctx* ctx = nullptr;
  int rc = dg_open(0, &ctx);
  if (rc != DG_OK) return fail(nullptr, rc, "dg_open");
  phase(cn ? "loading the graph (common neighbours computed)" : "loading the graph");
  rc = dg_graph_load_bin(ctx, argv[gi + 1], (uint32_t)v_num, undirected ? 1 : 0,
                         cn ? DG_CN_COMPUTE : DG_CN_FROM_DATA);
  if (rc != DG_OK) return fail(ctx, rc, "load_graph");
  phase("graph loaded");

  dg_walk_params wp;
  dg_walk_params_default(&wp, (uint32_t)v_num);
  wp.seed = seed;
  wp.walkers_per_round = (uint32_t)walkers;
  wp.min_length = (uint32_t)min_L;
  wp.init_round = (uint32_t)min_R;
  wp.max_length = (uint32_t)max_len;
  wp.max_rounds = (uint32_t)max_rounds;
  wp.rounds_mode = (uint32_t)rounds_mode;
  if ((rc = dg_walk_reset(ctx)) != DG_OK || (rc = dg_train_corpus_clear(ctx)) != DG_OK)
    return fail(ctx, rc, "reset");
  if (tf.per_file_alpha) {
    phase("initialising the trainer");
    if ((rc = dg_sgns_init(ctx, &tf.p)) != DG_OK) return fail(ctx, rc, "sgns_init");
  }

  for (uint32_t iter = 1; iter <= wp.max_rounds; ++iter) {
    const uint32_t n = iter == 1 ? wp.walkers_per_round * wp.init_round : wp.walkers_per_round;
    dg_round_stats rs;
    std::memset(&rs, 0, sizeof(rs));
    // a round whose relative entropy is undefined (vertices never visited, as
    // with few walks) still has its corpus; the reference would divide by zero
    const int round_rc = dg_walk_round(ctx, &wp, iter, 0, n, &rs);
    if (round_rc != DG_OK && round_rc != DG_ERR_RE_UNDEFINED) return fail(ctx, round_rc, "walk round");
    if (tf.debug > 0)
      std::fprintf(stderr, "round %u: %u walks, %llu vertices, RE %.6f, %.2f ms\n", iter, rs.walkers,
                   (unsigned long long)rs.tokens, rs.re, rs.walk_ms);
    if (oi > 0) {
      const std::string path = std::string(argv[oi + 1]) + "-0-" + std::to_string(iter) + ".txt";
      phase("writing the walk file");
      if ((rc = dg_corpus_write_text(ctx, path.c_str(), 1)) != DG_OK) return fail(ctx, rc, "write walks");
    }
    if (tf.per_file_alpha) {
      // one walk file, one learning rate: ExponentialDecayLR(alpha, 0.1), its
      // step advanced once per file (src/lr_scheduler.hpp:23-31,
      // src/word2vec.cu:1279-1290)
      const float alpha = dg_lr_value(DG_LR_EXPONENTIAL, tf.p.alpha, 0.1f, 0.0f, (int)(iter - 1));
      dg_sgns_stats st;
      if ((rc = dg_sgns_train(ctx, DG_CORPUS_ROUND, &tf.p, alpha, &st)) != DG_OK)
        return fail(ctx, rc, "train");
    } else if ((rc = dg_corpus_append(ctx)) != DG_OK) {
      return fail(ctx, rc, "corpus");
    }
    if (round_rc == DG_OK && rs.stop) break;
  }

  if (tf.per_file_alpha) {
    phase("writing the embeddings");
    if (!tf.emb_output.empty() && (rc = dg_emb_write(ctx, tf.emb_output.c_str(), tf.binary)) != DG_OK)
      return fail(ctx, rc, "write embeddings");
  } else {
    // vertex degrees as huge_walk hands them to the trainer (in + out degree)
    uint32_t vn = 0;
    uint64_t m = 0;
    dg_graph_info(ctx, &vn, &m);
    std::vector<uint64_t> row_ptr(vn + 1);
    dg_graph_get_csr(ctx, row_ptr.data(), nullptr, nullptr);
    std::vector<uint32_t> degrees(vn);
    for (uint32_t v = 0; v < vn; ++v) degrees[v] = (uint32_t)(2 * (row_ptr[v + 1] - row_ptr[v]));
    phase("training on the walks");
    if (use_dsgl) {
      if ((rc = dsgl_hip(argc, argv, nullptr, ctx)) != DG_OK) return fail(ctx, rc, "train");
    } else if ((rc = train_corpus_hip(argc, argv, degrees, ctx, 0)) != DG_OK) {
      return fail(ctx, rc, "train");
    }
  }
  phase("closing the device");
  dg_close(ctx);
  phase("done");
  return 0;
}
