/*
 * dsgl_oracle.c - CPU restatement of DSGL, the reference's CPU trainer
 * (include/dsgl.hpp:783-1427), for one process.
 *
 * TEST INFRASTRUCTURE ONLY (see dg_oracle.h): the FP64 form checks the
 * device's DSGL trainer; the FP32 multi-threaded form is the CPU baseline the
 * benchmark times.
 *
 * What is the reference's, what is chosen here:
 *   - kept: the split of the corpus (with one separator per walk) into equal
 *     word ranges, one per compute thread (dsgl.hpp:814-825), each starting
 *     wherever its range starts; two sentences trained in lockstep, one
 *     window shrink for both; a batch of max(L1, L2) * K negatives shared by
 *     the positions until two sentences have finished, read round-robin from
 *     position neg_count % batch (dsgl.hpp:1052-1083,1135-1137); the outputs
 *     [target, distinct negatives with multiplicity, target2], a negative
 *     equal to the target dropped (dsgl.hpp:1138-1159); label 1 only for
 *     (target, first-sentence contexts) and (target2, second-sentence
 *     contexts), every gradient from one snapshot of the rows
 *     (dsgl.hpp:1241-1310, the -D USE_MKL branch the reference builds with,
 *     CMakeLists.txt:41); sentence-local input rows and batch-local negative
 *     rows, the two targets updated in the table directly; the linear rate
 *     every 10000 counted words; the sentences in hand dropped when a range
 *     runs out; subsampling by sqrt(r) + r with r = sample * train_words / cn.
 *   - chosen (deviations, each for a reason):
 *       * draws: Philox streams instead of one LCG per thread - subsampling
 *         per corpus token as the other trainers draw it, the window shrink
 *         per (worker, step), the negatives per (worker, batch) - so a
 *         device run and this restatement draw the same numbers;
 *       * write-backs add (live - snapshot) to the table instead of
 *         overwriting it with the live copy, so concurrent workers' updates
 *         (and a target's update to a row the negative batch also holds)
 *         are not lost (the survey's ledger: dsgl.hpp:1352-1356,1401-1404);
 *       * subsampling reads the word's own count (the reference indexes the
 *         vertex-indexed counts with the rank, dsgl.hpp:923,988 vs :344);
 *       * the rate: word_count_actual * num_procs becomes the worker's own
 *         counted words times the number of workers (the threads advance in
 *         step, and a sum over racing threads is not reproducible);
 *       * the negative cursor and batch state restart with every epoch.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../gdistger-re_amd/csrc/dg_philox.h"
#include "dg_sgns_oracle.h"

#define DSGL_MAX_EXP 6
#define DSGL_MAX_SENTENCE 1000  /* MAX_SENTENCE_LENGTH, dsgl.hpp:23 */
#define DSGL_MAX_OUT 18         /* target, up to 16 negatives, target2 */

typedef struct {
  uint64_t cur, end;  /* virtual positions */
  uint64_t walk;      /* walk holding cur */
} dsgl_cursor;

/* Word j of the Philox stream (counter (j / 4, worker, c2, tag), key). */
static uint32_t dsgl_raw(uint64_t key, uint32_t worker, uint32_t c2, uint32_t tag, uint32_t j) {
  const dg_u32x4 o = dg_philox4x32_10(j >> 2, worker, c2, tag, (uint32_t)key, (uint32_t)(key >> 32));
  const uint32_t r = j & 3u;
  return r == 0 ? o.x0 : (r == 1 ? o.x1 : (r == 2 ? o.x2 : o.x3));
}

/* A negative from a draw: table[x % table_size], 0 -> x % (vocab - 1) + 1
 * (dsgl.hpp:1060-1064). */
static int32_t dsgl_negative(const dgo_dsgl_args* a, uint32_t x) {
  int32_t smp = a->table[x % a->table_size];
  if (!smp) smp = (int32_t)(x % (a->vocab - 1) + 1);
  return smp;
}

/* Kept: the token's draw (corpus token index ti, as subsample_*_kernel and
 * dgo_subsample draw it) is at most the word's keep limit. */
static int dsgl_keep(const dgo_dsgl_args* a, uint64_t ti, uint32_t r, uint64_t key) {
  const dg_u32x4 o = dg_philox4x32_10(0, (uint32_t)ti, (uint32_t)(ti >> 32), DG_TAG_SAMPLE,
                                      (uint32_t)key, (uint32_t)(key >> 32));
  return (o.x0 & 0xFFFFu) <= a->keep[r];
}

#define REAL double
#define FN(name) name##_f64
#include "dsgl_worker.inc"
#undef REAL
#undef FN

#define REAL float
#define FN(name) name##_f32
#include "dsgl_worker.inc"
#undef REAL
#undef FN

/* DSGL subsampling (dsgl.hpp:922-926): ran = sqrtf(r) + r with
 * r = sample * train_words / cn in float; a token is dropped when
 * ran < draw / 65536, so it is kept iff draw <= floor(ran * 65536). */
void dgo_dsgl_keep_limits(const uint64_t* cn_of_rank, uint32_t vocab, float sample,
                          uint64_t train_words, uint32_t* keep) {
  for (uint32_t w = 0; w < vocab; ++w) {
    if (!(sample > 0) || cn_of_rank[w] == 0) {
      keep[w] = 65536;
      continue;
    }
    const float ratio = (sample * (float)train_words) / (float)cn_of_rank[w];
    const float ran = sqrtf(ratio) + ratio;
    const float lim = ran * 65536.0f;
    keep[w] = lim >= 65536.0f ? 65536u : (uint32_t)lim;
  }
}

/* DSGL's unigram table (dsgl.hpp:121-147): pow(count, 0.75f) summed in
 * double, d1 held in float, the index clamped before its term is added. */
void dgo_dsgl_unigram_table(const uint64_t* cn_of_rank, uint32_t vocab, uint32_t table_size,
                            int32_t* table) {
  const float power = 0.75f;
  double pw = 0.;
  for (uint32_t i = 0; i < vocab; ++i) pw += pow((double)cn_of_rank[i], (double)power);
  int i = 0;
  float d1 = (float)(pow((double)cn_of_rank[0], (double)power) / pw);
  for (uint32_t a = 0; a < table_size; ++a) {
    table[a] = i;
    if (a / (float)table_size > d1) {
      i++;
      if (i >= (int)vocab) i = (int)vocab - 1;
      d1 += (float)(pow((double)cn_of_rank[i], (double)power) / pw);
    }
  }
}

static int check_args(const dgo_dsgl_args* a) {
  if (!a || !a->tokens || !a->offsets || a->n_walks == 0 || a->workers == 0 || a->window == 0 ||
      a->vocab < 2 || a->negative + 2 > DSGL_MAX_OUT || a->dim == 0)
    return -1;
  return 0;
}

/* FP64: iter epochs, in each the workers one after another (the order a
 * single device wavefront runs them in deterministic mode). */
int dgo_dsgl_train_f64(const dgo_dsgl_args* a, float* syn0, float* syn1) {
  if (check_args(a)) return -1;
  const size_t n = (size_t)a->vocab * a->dim;
  double* s0 = (double*)malloc(sizeof(double) * n);
  double* s1 = (double*)malloc(sizeof(double) * n);
  uint64_t* actual = (uint64_t*)calloc(a->workers, sizeof(uint64_t));
  float* alpha = (float*)malloc(sizeof(float) * a->workers);
  dsgl_ws_f64 ws;
  int rc = (s0 && s1 && actual && alpha) ? ws_alloc_f64(&ws, a->dim, a->negative, a->window) : -1;
  if (rc == 0) {
    for (size_t i = 0; i < n; ++i) {
      s0[i] = syn0[i];
      s1[i] = syn1[i];
    }
    for (uint32_t t = 0; t < a->workers; ++t) alpha[t] = a->alpha0;
    for (uint32_t ep = 0; ep < a->iter; ++ep)
      for (uint32_t t = 0; t < a->workers; ++t)
        dsgl_worker_epoch_f64(a, t, a->epoch0 + ep, s0, s1, &actual[t], &alpha[t], &ws);
    for (size_t i = 0; i < n; ++i) {
      syn0[i] = (float)s0[i];
      syn1[i] = (float)s1[i];
    }
    ws_free_f64(&ws);
  }
  free(s0);
  free(s1);
  free(actual);
  free(alpha);
  return rc;
}

/* FP32 on `threads` OpenMP threads, worker t on thread t % threads, racing on
 * the shared tables as the reference's threads do.  Returns 0, or -1. */
int dgo_dsgl_train_f32(const dgo_dsgl_args* a, float* syn0, float* syn1, int threads) {
  if (check_args(a) || threads < 1) return -1;
  uint64_t* actual = (uint64_t*)calloc(a->workers, sizeof(uint64_t));
  float* alpha = (float*)malloc(sizeof(float) * a->workers);
  if (!actual || !alpha) {
    free(actual);
    free(alpha);
    return -1;
  }
  for (uint32_t t = 0; t < a->workers; ++t) alpha[t] = a->alpha0;
  int bad = 0;
  for (uint32_t ep = 0; ep < a->iter; ++ep) {
#pragma omp parallel num_threads(threads) reduction(| : bad)
    {
      dsgl_ws_f32 ws;
      if (ws_alloc_f32(&ws, a->dim, a->negative, a->window) != 0) {
        bad |= 1;
      } else {
#pragma omp for schedule(static, 1)
        for (uint32_t t = 0; t < a->workers; ++t)
          dsgl_worker_epoch_f32(a, t, a->epoch0 + ep, syn0, syn1, &actual[t], &alpha[t], &ws);
        ws_free_f32(&ws);
      }
    }
  }
  free(actual);
  free(alpha);
  return bad ? -1 : 0;
}
