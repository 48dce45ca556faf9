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
 * in epoch e is word 0 of Philox(ctr = (0, i lo, i hi ^ e<<24, SAMPLE)). */
uint64_t dgo_subsample(const uint32_t* tokens, const uint64_t* offsets, uint64_t n_walks,
                       const uint32_t* rank_of_vertex, const uint32_t* keep, uint64_t seed,
                       uint32_t epoch, int32_t* sen, uint64_t* sen_off) {
  uint64_t n = 0;
  sen_off[0] = 0;
  for (uint64_t w = 0; w < n_walks; ++w) {
    for (uint64_t i = offsets[w]; i < offsets[w + 1]; ++i) {
      uint32_t r = rank_of_vertex[tokens[i]];
      if (r == 0xffffffffu) continue;
      dg_u32x4 o = dg_philox4x32_10(0, (uint32_t)i, (uint32_t)(i >> 32) ^ (epoch << 24),
                                    DG_TAG_SAMPLE, (uint32_t)seed, (uint32_t)(seed >> 32));
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
 * Batched form: tiles of T consecutive positions of a sentence; each
 * position draws its own K negatives, shared by the contexts of its window
 * (the per-position minibatch of dsgl.hpp:1100-1159).  Within a position the
 * contexts are taken in order and each (context, output) pair sees the
 * output row as updated by the earlier contexts of that position, as
 * skip_gram_kernel updates it (word2vec.cu:360-396); everything else of a
 * launch of `minibatch` tiles is taken from the tables as they were at its
 * start, and the row sums are added once.  Arithmetic in double, rows stored
 * as float: the FP64 reference the FP32 and BF16 device forms are measured
 * against.
 *
 * K = `draws` negatives per position (`negative` when 0), each weighted by
 * wneg = negative / K (a float; 1 when K = negative).  T = `tile` when
 * non-zero, else 8.
 * Tile (s, j): positions p0 = T j .. p1 - 1 = min(p0 + T, L) - 1 of sentence
 * s (L >= 2).  Philox stream (ctr = (blk, gs, j | epoch << 16, SGNS)), gs =
 * first_sen + s: word q < np gives b_{p0+q} = word % window; word np + q K + k
 * gives position q's negative k: table[word % table_size] (0 -> word %
 * (vocab-1) + 1).  Contexts c in [max(0, p0-W), min(L, p1+W)).  Output
 * column n = q (K+1) + r of position p = p0 + q: r = 0 its target t_p (label
 * 1), r > 0 negative r-1 (label 0, skipped when equal to t_p).  For each
 * column, with y~ = y_n and the contexts c of win(p) in ascending order:
 *   g[c][n] = (label - sig(x_c . y~)) alpha (wneg for negatives);  y~ += g x_c
 *   dX[c] = sum_n g[c][n] y_n ;  dY[n] = sum_c g[c][n] x_c   (staged rows)
 * win(p) = {c != p : |c - p| <= W - b_p}.  Row contributions per tile in the
 * order contexts (ascending c), then columns (ascending n).
 * ---------------------------------------------------------------------- */
static double sig_d(double f, const float* expt, int exact) {
  if (exact) return 1.0 / (1.0 + exp(-f));
  if (f > MAX_EXP) return 1.0;
  if (f < -MAX_EXP) return 0.0;
  return (double)expt[(int)((float)(f + MAX_EXP) * (EXP_TABLE_SIZE / MAX_EXP / 2))];
}

typedef struct {
  uint32_t row;
  uint32_t table; /* 0 = input (syn0), 1 = output (syn1) */
  uint64_t order;
  double* vec;
} contrib_t;

static int cmp_contrib(const void* a, const void* b) {
  const contrib_t* x = (const contrib_t*)a;
  const contrib_t* y = (const contrib_t*)b;
  if (x->table != y->table) return x->table < y->table ? -1 : 1;
  if (x->row != y->row) return x->row < y->row ? -1 : 1;
  return (x->order > y->order) - (x->order < y->order);
}

uint32_t dgo_tile_positions(uint32_t draws, uint32_t tile) {
  (void)draws;
  return tile ? tile : 8;
}

uint64_t dgo_tile_count(const uint64_t* sen_off, uint64_t n_sen, uint32_t tile) {
  uint64_t n = 0, T = tile;
  for (uint64_t s = 0; s < n_sen; ++s) {
    uint64_t L = sen_off[s + 1] - sen_off[s];
    if (L >= 2) n += (L + T - 1) / T;
  }
  return n;
}

void dgo_sg_batched(const int32_t* sen, const uint64_t* sen_off, uint64_t first_sen,
                    uint64_t n_sen, uint64_t minibatch, uint32_t window, uint32_t negative,
                    const int32_t* table, uint32_t table_size, uint32_t vocab, float alpha,
                    const float* expt, int exact_sigmoid, uint64_t seed, uint32_t epoch,
                    uint32_t dim, uint32_t tile, uint32_t draws, float* syn0, float* syn1) {
  const int K = (int)(draws ? draws : negative), W = (int)window, K1 = K + 1;
  const float wneg = (float)negative / (float)K;
  const int T = (int)dgo_tile_positions((uint32_t)K, tile);
  uint64_t n_tiles = dgo_tile_count(sen_off, n_sen, (uint32_t)T);
  uint64_t* ts = (uint64_t*)malloc(sizeof(uint64_t) * (n_tiles ? n_tiles : 1));
  uint32_t* tj = (uint32_t*)malloc(sizeof(uint32_t) * (n_tiles ? n_tiles : 1));
  uint64_t nt = 0;
  for (uint64_t s = 0; s < n_sen; ++s) {
    uint64_t L = sen_off[s + 1] - sen_off[s];
    if (L < 2) continue;
    for (uint32_t j = 0; (uint64_t)j * T < L; ++j) {
      ts[nt] = s;
      tj[nt] = j;
      ++nt;
    }
  }
  int32_t* out = (int32_t*)malloc(sizeof(int32_t) * (size_t)T * K1);
  int* b = (int*)malloc(sizeof(int) * (size_t)T);
  double* ytil = (double*)malloc(sizeof(double) * dim);
  for (uint64_t t0 = 0; t0 < n_tiles; t0 += minibatch) {
    uint64_t t1 = t0 + minibatch < n_tiles ? t0 + minibatch : n_tiles;
    uint64_t cap = (t1 - t0) * ((uint64_t)T * K1 + (uint64_t)T + 2 * (uint64_t)W);
    contrib_t* cs = (contrib_t*)malloc(sizeof(contrib_t) * (cap ? cap : 1));
    double* vecs = (double*)calloc(cap ? cap * dim : 1, sizeof(double));
    uint64_t ns = 0;
    for (uint64_t ti = t0; ti < t1; ++ti) {
      uint64_t s = ts[ti];
      const int32_t* t = sen + sen_off[s];
      int L = (int)(sen_off[s + 1] - sen_off[s]);
      int p0 = (int)tj[ti] * T, p1 = p0 + T < L ? p0 + T : L;
      int np = p1 - p0;
      dg_stream st;
      dg_stream_init(&st, seed, (uint32_t)(first_sen + s), tj[ti] | (epoch << 16), DG_TAG_SGNS);
      for (int q = 0; q < np; ++q) b[q] = (int)(dg_stream_next(&st) % window);
      for (int q = 0; q < np; ++q) out[q * K1] = t[p0 + q];
      for (int q = 0; q < np; ++q)
        for (int k = 0; k < K; ++k) {
          uint32_t x = dg_stream_next(&st);
          int32_t n = table[x % table_size];
          if (n == 0) n = (int32_t)(x % (vocab - 1) + 1);
          out[q * K1 + 1 + k] = n;
        }
      int c_lo = p0 - W > 0 ? p0 - W : 0, c_hi = p1 + W < L ? p1 + W : L;
      int nc = c_hi - c_lo;
      int no = np * K1;
      double* g = (double*)calloc((size_t)nc * no + 1, sizeof(double)); /* [c][n] */
      for (int n = 0; n < no; ++n) {
        int q = n / K1, r = n % K1, p = p0 + q, w = W - b[q];
        int32_t row = out[n];
        if (r > 0 && row == t[p]) continue;
        double lab = r == 0 ? 1.0 : 0.0;
        for (uint32_t d = 0; d < dim; ++d) ytil[d] = (double)syn1[(uint64_t)row * dim + d];
        for (int c = p - w > c_lo ? p - w : c_lo; c <= p + w && c < c_hi; ++c) {
          if (c == p) continue;
          const float* x = syn0 + (uint64_t)t[c] * dim;
          double f = 0.0;
          for (uint32_t d = 0; d < dim; ++d) f += (double)x[d] * ytil[d];
          double gg = (lab - sig_d(f, expt, exact_sigmoid)) * alpha;
          if (r > 0) gg *= wneg;
          g[(size_t)(c - c_lo) * no + n] = gg;
          for (uint32_t d = 0; d < dim; ++d) ytil[d] += gg * (double)x[d];
        }
      }
      for (int c = 0; c < nc; ++c) {
        double* v = vecs + ns * dim;
        for (int n = 0; n < no; ++n) {
          double gg = g[(size_t)c * no + n];
          if (gg == 0.0) continue;
          for (uint32_t d = 0; d < dim; ++d) v[d] += gg * syn1[(uint64_t)out[n] * dim + d];
        }
        cs[ns].row = (uint32_t)t[c_lo + c];
        cs[ns].table = 0;
        cs[ns].order = ns;
        cs[ns].vec = v;
        ++ns;
      }
      for (int n = 0; n < no; ++n) {
        double* v = vecs + ns * dim;
        for (int c = 0; c < nc; ++c) {
          double gg = g[(size_t)c * no + n];
          if (gg != 0.0)
            for (uint32_t d = 0; d < dim; ++d) v[d] += gg * syn0[(uint64_t)t[c_lo + c] * dim + d];
        }
        cs[ns].row = (uint32_t)out[n];
        cs[ns].table = 1;
        cs[ns].order = ns;
        cs[ns].vec = v;
        ++ns;
      }
      free(g);
    }
    qsort(cs, ns, sizeof(contrib_t), cmp_contrib);
    double* acc = (double*)malloc(sizeof(double) * dim);
    uint64_t i = 0;
    while (i < ns) {
      uint64_t j = i;
      for (uint32_t d = 0; d < dim; ++d) acc[d] = 0.0;
      while (j < ns && cs[j].row == cs[i].row && cs[j].table == cs[i].table) {
        for (uint32_t d = 0; d < dim; ++d) acc[d] += cs[j].vec[d];
        ++j;
      }
      float* w = (cs[i].table == 0 ? syn0 : syn1) + (uint64_t)cs[i].row * dim;
      for (uint32_t d = 0; d < dim; ++d) w[d] = (float)((double)w[d] + acc[d]);
      i = j;
    }
    free(acc);
    free(cs);
    free(vecs);
  }
  free(ytil);
  free(b);
  free(out);
  free(ts);
  free(tj);
}

/* ------------------------------------------------------------------------
 * Tuple loss for the differentiation checks (FP64):
 *   L = sum_t [ -log sig(x_ctx . y_center) - sum_k log sig(-x_ctx . y_negk) ]
 * with x = input rows (syn0), y = output rows (syn1), tuples laid out as
 * (center, context, neg_1..neg_K).
 * ---------------------------------------------------------------------- */
double dgo_tuple_loss(const uint32_t* tuples, uint64_t n, uint32_t K, const double* syn0,
                      const double* syn1, uint32_t dim) {
  double loss = 0.0;
  uint32_t w = 2 + K;
  for (uint64_t t = 0; t < n; ++t) {
    const double* x = syn0 + (uint64_t)tuples[t * w + 1] * dim;
    for (uint32_t j = 0; j < 1 + K; ++j) {
      uint32_t r = j == 0 ? tuples[t * w] : tuples[t * w + 1 + j];
      const double* y = syn1 + (uint64_t)r * dim;
      double z = 0.0;
      for (uint32_t d = 0; d < dim; ++d) z += x[d] * y[d];
      /* -log sig(s z) = log(1 + exp(-s z)), s = +1 for the centre, -1 for negatives */
      double a = j == 0 ? -z : z;
      loss += a > 0 ? a + log1p(exp(-a)) : log1p(exp(a));
    }
  }
  return loss;
}
