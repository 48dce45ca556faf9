/*
 * sgns_oracle.c - CPU restatement of the skip-gram negative-sampling trainer
 * (word2vec.cu path and the batched shared-negative form).
 *
 * TEST INFRASTRUCTURE ONLY (see dg_oracle.h).  Single-threaded, sequential
 * order; float arithmetic kept in the same operation order as the kernels it
 * checks, so deterministic GPU modes can be compared bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../gdistger-re_amd/csrc/dg_philox.h"
#include "dg_sgns_oracle.h"

#define EXP_TABLE_SIZE 1000
#define MAX_EXP 6

/* ------------------------------------------------------------------------
 * S1: vocabulary.  src/word2vec.cu:791-809 (words = vertex ids, counts =
 * degrees) and :632-659 (SortVocab: descending count, words below min_count
 * dropped except index 0).  Ties are broken by vertex id (qsort is unstable).
 * ---------------------------------------------------------------------- */
static const uint64_t* g_sort_cn;
static int cmp_vocab(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  if (g_sort_cn[x] != g_sort_cn[y]) return g_sort_cn[x] > g_sort_cn[y] ? -1 : 1;
  return (x > y) - (x < y);
}

int dgo_vocab(uint32_t v_num, const uint64_t* cn, uint32_t min_count, uint32_t* vertex_of_rank,
              uint32_t* rank_of_vertex, uint64_t* cn_of_rank, uint32_t* vocab_size,
              uint64_t* train_words) {
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (v_num ? v_num : 1));
  if (!order) return -1;
  for (uint32_t i = 0; i < v_num; ++i) order[i] = i;
  g_sort_cn = cn;
  qsort(order, v_num, sizeof(uint32_t), cmp_vocab);
  uint32_t n = 0;
  uint64_t tw = 0;
  for (uint32_t i = 0; i < v_num; ++i) rank_of_vertex[i] = 0xffffffffu;
  for (uint32_t a = 0; a < v_num; ++a) {
    uint32_t v = order[a];
    if (cn[v] < min_count && a != 0) continue;
    vertex_of_rank[n] = v;
    cn_of_rank[n] = cn[v];
    rank_of_vertex[v] = n;
    tw += cn[v];
    ++n;
  }
  free(order);
  *vocab_size = n;
  *train_words = tw;
  return 0;
}

/* S2: src/word2vec.cu:868-873 - LCG from 1, syn0 = ((r & 0xFFFF)/65536f - 0.5)/d
 * evaluated in double and rounded once (float_ops = 0), or the float
 * expression of include/dsgl.hpp:452-456 (float_ops = 1). */
void dgo_init_syn0(uint32_t vocab, uint32_t dim, int float_ops, float* syn0) {
  unsigned long long next_random = 1;
  for (uint64_t a = 0; a < vocab; ++a)
    for (uint64_t b = 0; b < dim; ++b) {
      next_random = next_random * (unsigned long long)25214903917ULL + 11;
      if (float_ops)
        syn0[a * dim + b] = (((next_random & 0xFFFF) / 65536.f) - 0.5f) / (float)dim;
      else
        syn0[a * dim + b] = (float)((((double)((float)(next_random & 0xFFFF) / (float)65536)) - 0.5) /
                                    (double)(long long)dim);
    }
}

/* S3: src/word2vec.cu:537-552 unigram table ~ count^0.75 (double d1).  The
 * reference reads vocab[vocab_size] when the index runs past the end; here
 * that increment adds nothing. */
void dgo_unigram_table(const uint64_t* cn_of_rank, uint32_t vocab, uint32_t table_size,
                       int32_t* table) {
  double train_words_pow = 0;
  for (uint32_t a = 0; a < vocab; ++a) train_words_pow += pow((double)cn_of_rank[a], 0.75);
  uint32_t i = 0;
  double d1 = pow((double)cn_of_rank[0], 0.75) / train_words_pow;
  for (uint32_t a = 0; a < table_size; ++a) {
    table[a] = (int32_t)i;
    if (a / (double)table_size > d1) {
      i++;
      if (i < vocab) d1 += pow((double)cn_of_rank[i], 0.75) / train_words_pow;
    }
    if (i >= vocab) i = vocab - 1;
  }
}

/* src/word2vec.cu:1569-1577: sigmoid table, exp in double of a float argument. */
void dgo_exp_table(float* t) {
  for (int i = 0; i < EXP_TABLE_SIZE; ++i) {
    t[i] = (float)exp((double)((i / (float)EXP_TABLE_SIZE * 2 - 1) * MAX_EXP));
    t[i] = t[i] / (t[i] + 1);
  }
  t[EXP_TABLE_SIZE] = t[EXP_TABLE_SIZE - 1];
}

/* Subsampling keep limit per word (src/word2vec.cu:1015-1020 with float
 * arithmetic): the token is dropped when ran < (r & 0xFFFF)/65536, i.e. kept
 * iff (r & 0xFFFF) <= floor(ran * 65536); 65536 means always kept. */
void dgo_keep_limits(const uint64_t* cn_of_rank, uint32_t vocab, float sample, uint64_t train_words,
                     uint32_t* keep) {
  for (uint32_t w = 0; w < vocab; ++w) {
    if (!(sample > 0)) {
      keep[w] = 65536;
      continue;
    }
    float st = sample * (float)train_words;
    float c = (float)cn_of_rank[w];
    float ran = (sqrtf(c / st) + 1) * st / c;
    float lim = ran * 65536.0f;
    keep[w] = lim >= 65536.0f ? 65536u : (uint32_t)lim;
  }
}

/* Subsampled sentences: each walk of the corpus becomes one sentence (tokens
 * mapped to vocabulary ranks; vertices outside the vocabulary skipped, as
 * SearchVocab == -1 at src/word2vec.cu:1012).  The draw for corpus token i
 * in epoch e is word 0 of Philox(ctr = (0, i lo, i hi, SAMPLE)) under the key
 * dg_epoch_key(seed, e). */
uint64_t dgo_subsample(const uint32_t* tokens, const uint64_t* offsets, uint64_t n_walks,
                       const uint32_t* rank_of_vertex, const uint32_t* keep, uint64_t seed,
                       uint32_t epoch, int32_t* sen, uint64_t* sen_off) {
  uint64_t n = 0;
  const uint64_t key = dg_epoch_key(seed, epoch);
  sen_off[0] = 0;
  for (uint64_t w = 0; w < n_walks; ++w) {
    for (uint64_t i = offsets[w]; i < offsets[w + 1]; ++i) {
      uint32_t r = rank_of_vertex[tokens[i]];
      if (r == 0xffffffffu) continue;
      dg_u32x4 o = dg_philox4x32_10(0, (uint32_t)i, (uint32_t)(i >> 32), DG_TAG_SAMPLE,
                                    (uint32_t)key, (uint32_t)(key >> 32));
      if ((o.x0 & 0xFFFFu) > keep[r]) continue;
      sen[n++] = (int32_t)r;
    }
    sen_off[w + 1] = n;
  }
  return n;
}

static float table_sigmoid_grad(float f, int label, float alpha, const float* expt) {
  /* src/word2vec.cu:384-389 */
  if (f > MAX_EXP) return (label - 1) * alpha;
  if (f < -MAX_EXP) return (label - 0) * alpha;
  return (label - expt[(int)((f + MAX_EXP) * (EXP_TABLE_SIZE / MAX_EXP / 2))]) * alpha;
}

/* Dot product in the order of the reference block reduction
 * (src/word2vec.cu:373-381 for d = 2 * FSIZE; generalised to d = 64 * k):
 * s[l] = sum over j of x[l + 64 j] * y[l + 64 j] taken in order j = 0, 1, ...
 * (pairs folded as f[x] += f[x + FSIZE] per halving), then the 64 partials
 * folded with offsets 32, 16, 8, 4, 2, 1. */
static float tree_dot(const float* x, const float* y, uint32_t dim) {
  float s[64];
  uint32_t per = dim / 64;
  for (uint32_t l = 0; l < 64; ++l) {
    /* per-lane values p[j] = x[l + 64 j] * y[l + 64 j]; halving folds */
    float p[16];
    for (uint32_t j = 0; j < per; ++j) p[j] = x[l + 64 * j] * y[l + 64 * j];
    uint32_t n = per;
    while (n > 1) {
      uint32_t h = n / 2;
      for (uint32_t j = 0; j < h; ++j) p[j] = p[j] + p[j + h];
      n = h;
    }
    s[l] = p[0];
  }
  for (uint32_t off = 32; off >= 1; off /= 2)
    for (uint32_t l = 0; l < off; ++l) s[l] = s[l] + s[l + off];
  return s[0];
}

/* ------------------------------------------------------------------------
 * S5: skip_gram_kernel (src/word2vec.cu:305-399), negative sampling only,
 * sentences of one batch in order (block index = sentence index in batch).
 * ---------------------------------------------------------------------- */
void dgo_sg_plain_batch(const int32_t* sen, const uint64_t* sen_off, uint64_t n_sen,
                        uint32_t window, uint32_t negative, const int32_t* table,
                        uint32_t table_size, uint32_t vocab, float alpha, const float* expt,
                        uint32_t dim, float* syn0, float* syn1) {
  float* neu1e = (float*)malloc(sizeof(float) * dim);
  for (uint64_t s = 0; s < n_sen; ++s) {
    unsigned long next_random = (unsigned long)s;
    for (uint64_t pos = sen_off[s]; pos < sen_off[s + 1]; ++pos) {
      int word = sen[pos];
      next_random = next_random * (unsigned long)2514903917UL + 11;
      int b = (int)(next_random % window);
      for (int a = b; a < (int)window * 2 + 1 - b; a++) {
        if (a == (int)window) continue;
        int64_t c = (int64_t)pos - (int64_t)window + a;
        if (c < (int64_t)sen_off[s] || c >= (int64_t)sen_off[s + 1]) continue;
        int last_word = sen[c];
        float* x = syn0 + (uint64_t)last_word * dim;
        for (uint32_t i = 0; i < dim; ++i) neu1e[i] = 0;
        for (uint32_t d = 0; d < negative + 1; d++) {
          int target, label;
          if (d == 0) {
            target = word;
            label = 1;
          } else {
            next_random = next_random * (unsigned long)25214903917UL + 11;
            target = table[(next_random >> 16) % table_size];
            if (target == 0) target = (int)(next_random % (vocab - 1) + 1);
            if (target == word) continue;
            label = 0;
          }
          float* y = syn1 + (uint64_t)target * dim;
          float f = tree_dot(x, y, dim);
          float g = table_sigmoid_grad(f, label, alpha, expt);
          for (uint32_t i = 0; i < dim; ++i) {
            neu1e[i] += g * y[i];
            y[i] += g * x[i];
          }
        }
        for (uint32_t i = 0; i < dim; ++i) x[i] += neu1e[i];
      }
    }
  }
  free(neu1e);
}

/* ------------------------------------------------------------------------
 * S6: __sgNegReuse (src/word2vec.cu:231-303): a sentence's `negative`
 * negatives (negs[s * negative + k], drawn per sentence as the host loop at
 * :1057-1062 does) are shared by all its positions; every position trains
 * the full window (a from 0 to 2 window, no shrink); per (context, target)
 * pair the same dot product, table sigmoid and updates as skip_gram_kernel,
 * a negative equal to the centre word skipped (:263).  Sentences in order,
 * the outputs of a pair in order (the reference runs them in parallel
 * warps, which differs only when a sentence draws the same negative twice).
 * ---------------------------------------------------------------------- */
void dgo_sg_reuse_batch(const int32_t* sen, const uint64_t* sen_off, uint64_t n_sen,
                        uint32_t window, uint32_t negative, const int32_t* negs, float alpha,
                        const float* expt, uint32_t dim, float* syn0, float* syn1) {
  float* neu1e = (float*)malloc(sizeof(float) * dim);
  for (uint64_t s = 0; s < n_sen; ++s) {
    for (uint64_t pos = sen_off[s]; pos < sen_off[s + 1]; ++pos) {
      int word = sen[pos];
      for (int a = 0; a < (int)window * 2 + 1; a++) {
        if (a == (int)window) continue;
        int64_t c = (int64_t)pos - (int64_t)window + a;
        if (c < (int64_t)sen_off[s] || c >= (int64_t)sen_off[s + 1]) continue;
        float* x = syn0 + (uint64_t)sen[c] * dim;
        for (uint32_t i = 0; i < dim; ++i) neu1e[i] = 0;
        for (uint32_t d = 0; d < negative + 1; d++) {
          int target = d == 0 ? word : negs[s * negative + d - 1];
          int label = d == 0 ? 1 : 0;
          if (d > 0 && target == word) continue;
          float* y = syn1 + (uint64_t)target * dim;
          float f = tree_dot(x, y, dim);
          float g = table_sigmoid_grad(f, label, alpha, expt);
          for (uint32_t i = 0; i < dim; ++i) {
            neu1e[i] += g * y[i];
            y[i] += g * x[i];
          }
        }
        for (uint32_t i = 0; i < dim; ++i) x[i] += neu1e[i];
      }
    }
  }
  free(neu1e);
}

/* The negative-reuse form's negatives of sentence s of an epoch: word k of
 * Philox(counter (k / 4, s, 0, REUSE), key dg_epoch_key(seed, epoch)) as x,
 * table[x % table_size], 0 -> x % (vocab - 1) + 1 (word2vec.cu:1057-1062). */
void dgo_reuse_negs(uint64_t n_sen, uint32_t negative, uint64_t seed, uint32_t epoch,
                    const int32_t* table, uint32_t table_size, uint32_t vocab, int32_t* out) {
  const uint64_t key = dg_epoch_key(seed, epoch);
  for (uint64_t s = 0; s < n_sen; ++s) {
    dg_stream st;
    dg_stream_init(&st, key, (uint32_t)s, 0u, DG_TAG_SGNS_REUSE);
    for (uint32_t k = 0; k < negative; ++k) {
      uint32_t x = dg_stream_next(&st);
      int32_t v = table[x % table_size];
      out[s * negative + k] = v == 0 ? (int32_t)(x % (vocab - 1) + 1) : v;
    }
  }
}

/* Reference batch boundaries (src/word2vec.cu:999-1041): at most 15000
 * sentences, and a batch closes once it holds >= (15000 - 1) * 20 tokens. */
uint64_t dgo_plain_batches(const uint64_t* sen_off, uint64_t n_sen, uint64_t* batch_first,
                           uint64_t cap) {
  uint64_t nb = 0, s = 0;
  while (s < n_sen && nb < cap) {
    batch_first[nb++] = s;
    uint64_t cnt = 0;
    uint64_t base = sen_off[s];
    while (s < n_sen && cnt < 15000) {
      ++cnt;
      ++s;
      if (sen_off[s] - base >= (15000 - 1) * 20) break;
    }
  }
  if (nb < cap) batch_first[nb] = n_sen;
  return nb;
}

/* ------------------------------------------------------------------------
 * Batched form (shared negatives): skip_gram_kernel's order (word2vec.cu:
 * 305-399) - sentences in order, positions in order, contexts of the shrunk
 * window in order, the output row updated after every (context, output)
 * pair and the context row after all outputs of its pair - except that each
 * position draws its K negatives once, shared by all its contexts (the
 * per-position negatives of dsgl.hpp:1100-1159).  K = `draws` (`negative`
 * when 0), each negative's gradient weighted by wneg = negative / K (a float,
 * 1 when K = negative).  Tables are held in double for the whole call and
 * rounded to float at the end: the FP64 reference the device forms follow.
 *
 * Draws: Philox stream (ctr = (blk, first_sen + s, q, SGNS), key
 * dg_epoch_key(seed, epoch)) of position q of sentence s, first_sen + s being
 * the sentence's index in the epoch's corpus: word 0 gives b = word % window,
 * word k (1..K) negative k: table[word % table_size], 0 -> word % (vocab-1) + 1.
 * ---------------------------------------------------------------------- */
static double sig_d(double f, const float* expt, int exact) {
  if (exact) return 1.0 / (1.0 + exp(-f));
  if (f > MAX_EXP) return 1.0;
  if (f < -MAX_EXP) return 0.0;
  return (double)expt[(int)((float)(f + MAX_EXP) * (EXP_TABLE_SIZE / MAX_EXP / 2))];
}

void dgo_position_draws(uint64_t seed, uint64_t gs, uint32_t q, uint32_t epoch, uint32_t window,
                        uint32_t draws, const int32_t* table, uint32_t table_size, uint32_t vocab,
                        uint32_t* b, int32_t* negs) {
  dg_stream st;
  dg_stream_init(&st, dg_epoch_key(seed, epoch), (uint32_t)gs, q, DG_TAG_SGNS);
  *b = dg_stream_next(&st) % window;
  for (uint32_t k = 0; k < draws; ++k) {
    uint32_t x = dg_stream_next(&st);
    int32_t n = table[x % table_size];
    if (n == 0) n = (int32_t)(x % (vocab - 1) + 1);
    negs[k] = n;
  }
}

void dgo_sg_shared(const int32_t* sen, const uint64_t* sen_off, uint64_t first_sen, uint64_t n_sen,
                   uint32_t window, uint32_t negative, uint32_t draws, const int32_t* table,
                   uint32_t table_size, uint32_t vocab, float alpha, const float* expt,
                   int exact_sigmoid, uint64_t seed, uint32_t epoch, uint32_t dim, float* syn0,
                   float* syn1) {
  const uint32_t K = draws ? draws : negative;
  const float wneg = (float)negative / (float)K;
  uint64_t n_rows = 0;
  for (uint64_t i = 0; i < sen_off[n_sen]; ++i)
    if ((uint64_t)sen[i] + 1 > n_rows) n_rows = (uint64_t)sen[i] + 1;
  for (uint32_t i = 0; i < table_size; ++i)
    if ((uint64_t)table[i] + 1 > n_rows) n_rows = (uint64_t)table[i] + 1;
  if (vocab > n_rows) n_rows = vocab;
  double* x0 = (double*)malloc(sizeof(double) * n_rows * dim);
  double* y1 = (double*)malloc(sizeof(double) * n_rows * dim);
  for (uint64_t i = 0; i < n_rows * dim; ++i) {
    x0[i] = (double)syn0[i];
    y1[i] = (double)syn1[i];
  }
  int32_t* negs = (int32_t*)malloc(sizeof(int32_t) * (K ? K : 1));
  double* neu1e = (double*)malloc(sizeof(double) * dim);
  for (uint64_t s = 0; s < n_sen; ++s) {
    const int32_t* t = sen + sen_off[s];
    const int L = (int)(sen_off[s + 1] - sen_off[s]);
    for (int q = 0; q < L; ++q) {
      uint32_t b;
      dgo_position_draws(seed, first_sen + s, (uint32_t)q, epoch, window, K, table, table_size,
                         vocab, &b, negs);
      const int w = (int)window - (int)b;
      for (int c = q - w; c <= q + w; ++c) {
        if (c < 0 || c >= L || c == q) continue;
        double* x = x0 + (uint64_t)t[c] * dim;
        for (uint32_t d = 0; d < dim; ++d) neu1e[d] = 0.0;
        for (uint32_t n = 0; n <= K; ++n) {
          const int32_t row = n == 0 ? t[q] : negs[n - 1];
          if (n > 0 && row == t[q]) continue;
          double* y = y1 + (uint64_t)row * dim;
          double f = 0.0;
          for (uint32_t d = 0; d < dim; ++d) f += x[d] * y[d];
          double g = ((n == 0 ? 1.0 : 0.0) - sig_d(f, expt, exact_sigmoid)) * alpha;
          if (n > 0) g *= wneg;
          for (uint32_t d = 0; d < dim; ++d) {
            neu1e[d] += g * y[d];
            y[d] += g * x[d];
          }
        }
        for (uint32_t d = 0; d < dim; ++d) x[d] += neu1e[d];
      }
    }
  }
  for (uint64_t i = 0; i < n_rows * dim; ++i) {
    syn0[i] = (float)x0[i];
    syn1[i] = (float)y1[i];
  }
  free(neu1e);
  free(negs);
  free(x0);
  free(y1);
}

/* ------------------------------------------------------------------------
 * The same order as dgo_sg_shared, evaluated the way the batched device
 * forms evaluate it: per position, from the rows as the earlier positions
 * left them (x = syn0 rows of the contexts, y = syn1 rows of the target and
 * negatives), the gradients of the contexts c_1 < c_2 < ... in order,
 *   s_i[n] = x_i.y_n + sum_{j<i} g_j[n] x_i.x_j
 *                    + sum_{j<i, v_j = v_i} sum_n' g_j[n'] y_n'.y_n
 *   g_i[n] = (label_n - sig(s_i[n])) alpha w_n
 * (v_i the vertex of context i), then the updates
 *   syn0[v] += sum_{i: v_i = v} [ sum_n g_i[n] y_n + sum_{j<i} (g_i . g_j) x_j ]
 *   syn1[y_n] += sum_i g_i[n] x_i + sum_{i, j<i, v_j = v_i} sum_n' g_i[n] g_j[n'] y_n'.
 * These are the sequential updates of dgo_sg_shared up to terms of second
 * order in the updates (an output row or a context row changed by an earlier
 * pair of the same position shifts a later logit or update in proportion to
 * that change; the change of a change is dropped), and two negatives of one
 * position on the same row do not see each other's updates.  FP64.
 * ---------------------------------------------------------------------- */
void dgo_sg_shared_first_order(const int32_t* sen, const uint64_t* sen_off, uint64_t first_sen,
                               uint64_t n_sen, uint32_t window, uint32_t negative, uint32_t draws,
                               const int32_t* table, uint32_t table_size, uint32_t vocab,
                               float alpha, const float* expt, int exact_sigmoid, uint64_t seed,
                               uint32_t epoch, uint32_t dim, float* syn0, float* syn1) {
  const uint32_t K = draws ? draws : negative, K1 = K + 1;
  const float wneg = (float)negative / (float)K;
  const uint64_t n_rows = vocab;
  double* x0 = (double*)malloc(sizeof(double) * n_rows * dim);
  double* y1 = (double*)malloc(sizeof(double) * n_rows * dim);
  for (uint64_t i = 0; i < n_rows * dim; ++i) {
    x0[i] = (double)syn0[i];
    y1[i] = (double)syn1[i];
  }
  const int MS = 2 * (int)window;
  int32_t* negs = (int32_t*)malloc(sizeof(int32_t) * (K ? K : 1));
  int32_t* rows = (int32_t*)malloc(sizeof(int32_t) * K1);
  int* live = (int*)malloc(sizeof(int) * K1);
  int* cv = (int*)malloc(sizeof(int) * (MS + 1));
  double* X = (double*)malloc(sizeof(double) * (MS + 1) * dim);   /* context rows by slot */
  double* Y = (double*)malloc(sizeof(double) * K1 * dim);
  double* g = (double*)malloc(sizeof(double) * (MS + 1) * K1);
  double* dX = (double*)malloc(sizeof(double) * (MS + 1) * dim);
  double* dY = (double*)malloc(sizeof(double) * K1 * dim);
  for (uint64_t s = 0; s < n_sen; ++s) {
    const int32_t* t = sen + sen_off[s];
    const int L = (int)(sen_off[s + 1] - sen_off[s]);
    if (L < 2) continue;
    for (int q = 0; q < L; ++q) {
      uint32_t b;
      dgo_position_draws(seed, first_sen + s, (uint32_t)q, epoch, window, K, table, table_size,
                         vocab, &b, negs);
      const int w = (int)window - (int)b;
      int m = 0;
      for (int c = q - w; c <= q + w; ++c)
        if (c >= 0 && c < L && c != q) cv[m++] = t[c];
      for (uint32_t n = 0; n < K1; ++n) {
        rows[n] = n == 0 ? t[q] : negs[n - 1];
        live[n] = n == 0 || rows[n] != t[q];
        memcpy(Y + (size_t)n * dim, y1 + (uint64_t)rows[n] * dim, sizeof(double) * dim);
      }
      for (int i = 0; i < m; ++i) memcpy(X + (size_t)i * dim, x0 + (uint64_t)cv[i] * dim, sizeof(double) * dim);
      for (int i = 0; i < m; ++i)
        for (uint32_t n = 0; n < K1; ++n) {
          double gg = 0.0;
          if (live[n]) {
            double sv = 0.0;
            for (uint32_t d = 0; d < dim; ++d) sv += X[(size_t)i * dim + d] * Y[(size_t)n * dim + d];
            for (int j = 0; j < i; ++j) {
              double xx = 0.0;
              for (uint32_t d = 0; d < dim; ++d) xx += X[(size_t)i * dim + d] * X[(size_t)j * dim + d];
              sv += g[j * K1 + n] * xx;
              if (cv[j] == cv[i])
                for (uint32_t n2 = 0; n2 < K1; ++n2) {
                  double yy = 0.0;
                  for (uint32_t d = 0; d < dim; ++d) yy += Y[(size_t)n2 * dim + d] * Y[(size_t)n * dim + d];
                  sv += g[j * K1 + n2] * yy;
                }
            }
            gg = ((n == 0 ? 1.0 : 0.0) - sig_d(sv, expt, exact_sigmoid)) * alpha;
            if (n > 0) gg *= wneg;
          }
          g[i * K1 + n] = gg;
        }
      for (int i = 0; i < m; ++i)
        for (uint32_t d = 0; d < dim; ++d) {
          double v = 0.0;
          for (uint32_t n = 0; n < K1; ++n) v += g[i * K1 + n] * Y[(size_t)n * dim + d];
          for (int j = 0; j < i; ++j) {
            double gij = 0.0;
            for (uint32_t n = 0; n < K1; ++n) gij += g[i * K1 + n] * g[j * K1 + n];
            v += gij * X[(size_t)j * dim + d];
          }
          dX[(size_t)i * dim + d] = v;
        }
      for (uint32_t n = 0; n < K1; ++n)
        for (uint32_t d = 0; d < dim; ++d) {
          double v = 0.0;
          for (int i = 0; i < m; ++i) {
            v += g[i * K1 + n] * X[(size_t)i * dim + d];
            for (int j = 0; j < i; ++j)
              if (cv[j] == cv[i])
                for (uint32_t n2 = 0; n2 < K1; ++n2) v += g[i * K1 + n] * g[j * K1 + n2] * Y[(size_t)n2 * dim + d];
          }
          dY[(size_t)n * dim + d] = v;
        }
      for (int i = 0; i < m; ++i)
        for (uint32_t d = 0; d < dim; ++d) x0[(uint64_t)cv[i] * dim + d] += dX[(size_t)i * dim + d];
      for (uint32_t n = 0; n < K1; ++n)
        if (live[n])
          for (uint32_t d = 0; d < dim; ++d) y1[(uint64_t)rows[n] * dim + d] += dY[(size_t)n * dim + d];
    }
  }
  for (uint64_t i = 0; i < n_rows * dim; ++i) {
    syn0[i] = (float)x0[i];
    syn1[i] = (float)y1[i];
  }
  free(dY);
  free(dX);
  free(g);
  free(Y);
  free(X);
  free(cv);
  free(live);
  free(rows);
  free(negs);
  free(x0);
  free(y1);
}

/* ------------------------------------------------------------------------
 * Differentiable tuple interface (FP64).  A tuple is (center, context,
 * neg_1..neg_K); x = syn0[context], y_0 = syn1[center], y_k = syn1[neg_k].
 *   z_j = x . y_j,  L = sum_t [ softplus(-z_0) + sum_k softplus(z_k) ]
 * which is -log sig(z_0) - sum_k log sig(-z_k).  A negative equal to the
 * tuple's centre is skipped, as skip_gram_kernel skips it
 * (src/word2vec.cu:366): its logit is reported, its loss term and gradient
 * are 0.  bf16 != 0 rounds every row element to BF16 (nearest even) first,
 * the operands of the device's BF16 form.
 * ---------------------------------------------------------------------- */
static double tup_elem(const float* t, uint64_t i, int bf16) {
  float f = t[i];
  if (bf16) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    memcpy(&f, &u, 4);
  }
  return (double)f;
}

static double softplus(double a) { return a > 0 ? a + log1p(exp(-a)) : log1p(exp(a)); }

static uint32_t tup_target(const uint32_t* tp, uint32_t j) { return j == 0 ? tp[0] : tp[1 + j]; }

static int tup_live(const uint32_t* tp, uint32_t j) { return j == 0 || tp[1 + j] != tp[0]; }

static double tup_dot(const uint32_t* tp, uint32_t j, const float* syn0, const float* syn1,
                      uint32_t dim, int bf16) {
  const uint64_t xr = (uint64_t)tp[1] * dim, yr = (uint64_t)tup_target(tp, j) * dim;
  double z = 0.0;
  for (uint32_t d = 0; d < dim; ++d) z += tup_elem(syn0, xr + d, bf16) * tup_elem(syn1, yr + d, bf16);
  return z;
}

double dgo_tuple_forward(const uint32_t* tuples, uint64_t n, uint32_t K, const float* syn0,
                         const float* syn1, uint32_t dim, int bf16, double* logits) {
  const uint32_t w = 2 + K;
  double loss = 0.0;
  for (uint64_t t = 0; t < n; ++t) {
    const uint32_t* tp = tuples + t * w;
    for (uint32_t j = 0; j < 1 + K; ++j) {
      const double z = tup_dot(tp, j, syn0, syn1, dim, bf16);
      if (logits) logits[t * (1 + K) + j] = z;
      if (tup_live(tp, j)) loss += softplus(j == 0 ? -z : z);
    }
  }
  return loss;
}

/* dL/dz_j = sig(z_j) - label_j (0 for a skipped negative), or the caller's
 * upstream gradient dlogits (also 0 for a skipped negative). */
static double tup_dz(const uint32_t* tp, uint64_t t, uint32_t K, uint32_t j, double z,
                     const double* dlogits) {
  if (!tup_live(tp, j)) return 0.0;
  if (dlogits) return dlogits[t * (1 + K) + j];
  return 1.0 / (1.0 + exp(-z)) - (j == 0 ? 1.0 : 0.0);
}

void dgo_tuple_grad_rows(const uint32_t* tuples, uint64_t n, uint32_t K, const float* syn0,
                         const float* syn1, uint32_t dim, uint32_t vocab, int bf16,
                         const double* dlogits, int table, const uint32_t* rows, uint64_t n_rows,
                         double* out) {
  const uint32_t w = 2 + K;
  int64_t* at = (int64_t*)malloc(sizeof(int64_t) * vocab);
  for (uint32_t r = 0; r < vocab; ++r) at[r] = -1;
  for (uint64_t i = 0; i < n_rows; ++i) at[rows[i]] = (int64_t)i;
  memset(out, 0, sizeof(double) * n_rows * dim);
  for (uint64_t t = 0; t < n; ++t) {
    const uint32_t* tp = tuples + t * w;
    const uint64_t xr = (uint64_t)tp[1] * dim;
    for (uint32_t j = 0; j < 1 + K; ++j) {
      const uint32_t yv = tup_target(tp, j);
      const int64_t dst = table == 0 ? at[tp[1]] : at[yv];
      if (dst < 0) continue;
      const double gz = tup_dz(tp, t, K, j, tup_dot(tp, j, syn0, syn1, dim, bf16), dlogits);
      double* o = out + (uint64_t)dst * dim;
      const float* src = table == 0 ? syn1 : syn0;
      const uint64_t sr = table == 0 ? (uint64_t)yv * dim : xr;
      for (uint32_t d = 0; d < dim; ++d) o[d] += gz * tup_elem(src, sr + d, bf16);
    }
  }
  free(at);
}

/* Central difference (L(w + h e) - L(w - h e)) / 2h of the FP64 loss (no BF16
 * rounding) for coordinate (table, row, col).  Only tuples that read the row
 * change, so only they are evaluated. */
double dgo_tuple_fd(const uint32_t* tuples, uint64_t n, uint32_t K, const float* syn0,
                    const float* syn1, uint32_t dim, int table, uint32_t row, uint32_t col,
                    double h) {
  const uint32_t w = 2 + K;
  double lp = 0.0, lm = 0.0;
  for (uint64_t t = 0; t < n; ++t) {
    const uint32_t* tp = tuples + t * w;
    const int x_hit = table == 0 && tp[1] == row;
    int any = x_hit;
    for (uint32_t j = 0; j < 1 + K && !any; ++j) any = table == 1 && tup_target(tp, j) == row;
    if (!any) continue;
    for (uint32_t j = 0; j < 1 + K; ++j) {
      if (!tup_live(tp, j)) continue;
      const uint32_t yv = tup_target(tp, j);
      const int y_hit = table == 1 && yv == row;
      double zp = 0.0, zm = 0.0;
      for (uint32_t d = 0; d < dim; ++d) {
        const double xv = syn0[(uint64_t)tp[1] * dim + d], yw = syn1[(uint64_t)yv * dim + d];
        const double e = d == col ? h : 0.0;
        zp += (xv + (x_hit ? e : 0.0)) * (yw + (y_hit ? e : 0.0));
        zm += (xv - (x_hit ? e : 0.0)) * (yw - (y_hit ? e : 0.0));
      }
      lp += softplus(j == 0 ? -zp : zp);
      lm += softplus(j == 0 ? -zm : zm);
    }
  }
  return (lp - lm) / (2.0 * h);
}

/* The reference's in-place update over the tuples in order, with the exact
 * sigmoid (skip_gram_kernel, src/word2vec.cu:359-396): per tuple, for each
 * live target j in order, g = (label - sig(z)) * lr with z from the current
 * rows; neu1e += g * y_j; y_j += g * x; then x += neu1e. */
void dgo_tuple_inplace(const uint32_t* tuples, uint64_t n, uint32_t K, double lr, uint32_t dim,
                       double* syn0, double* syn1) {
  const uint32_t w = 2 + K;
  double* neu1e = (double*)malloc(sizeof(double) * dim);
  for (uint64_t t = 0; t < n; ++t) {
    const uint32_t* tp = tuples + t * w;
    double* x = syn0 + (uint64_t)tp[1] * dim;
    memset(neu1e, 0, sizeof(double) * dim);
    for (uint32_t j = 0; j < 1 + K; ++j) {
      if (!tup_live(tp, j)) continue;
      double* y = syn1 + (uint64_t)tup_target(tp, j) * dim;
      double z = 0.0;
      for (uint32_t d = 0; d < dim; ++d) z += x[d] * y[d];
      const double g = ((j == 0 ? 1.0 : 0.0) - 1.0 / (1.0 + exp(-z))) * lr;
      for (uint32_t d = 0; d < dim; ++d) neu1e[d] += g * y[d];
      for (uint32_t d = 0; d < dim; ++d) y[d] += g * x[d];
    }
    for (uint32_t d = 0; d < dim; ++d) x[d] += neu1e[d];
  }
  free(neu1e);
}
