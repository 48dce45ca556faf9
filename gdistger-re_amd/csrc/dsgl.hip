// dsgl.hip - DSGL, the reference's CPU trainer (include/dsgl.hpp:783-1427), on
// gfx950: one wavefront is one DSGL compute thread.
//
// Semantics (the same as the oracle's restatement, oracle/dsgl_oracle.c):
//   * the corpus, with one separator after every walk (local_corpus,
//     include/path.hpp:96,110), is split into `workers` equal word ranges
//     (dsgl.hpp:814-825); a worker reads sentences from its range: every
//     position read counts as a word, a separator or the range's last
//     position closes a sentence, subsampling drops tokens by
//     sqrt(r) + r (r = sample * train_words / cn) with a Philox draw per
//     corpus token, sentences of one token are skipped;
//   * two sentences advance in lockstep, one window shrink per step for both
//     (dsgl.hpp:1102-1134); their distinct input rows are copied into
//     sentence buffers (here in LDS) and go back to syn0 when the sentence
//     ends;
//   * a batch of max(L1, L2) * K negatives is drawn whenever two sentences
//     have finished (dsgl.hpp:1052-1083); its distinct output rows are
//     copied into a batch buffer (here per-wavefront scratch in HBM) and go
//     back to syn1 with the batch; each step reads K of them from position
//     neg_count % batch;
//   * per step the outputs are [target, distinct negatives with their
//     multiplicity, target2] (a negative equal to the target dropped,
//     dsgl.hpp:1138-1159), the contexts those of both windows; every
//     gradient comes from one snapshot: G = Y X^T, g = (label - sig) * alpha
//     * multiplicity with label 1 for (target, first-sentence contexts) and
//     (target2, second-sentence contexts), dY = G X, dX = G^T Y
//     (dsgl.hpp:1241-1310, the MKL branch); dX goes into the sentence
//     buffers, dY into the batch buffer and, for the two targets, into syn1
//     directly;
//   * the learning rate steps every 10000 counted words (dsgl.hpp:882-894),
//     the worker's own count times the number of workers standing in for
//     the shared counter.
// Chosen here: write-backs add live - snapshot (atomics) instead of
// overwriting the rows, so concurrent wavefronts' updates compose; Philox
// draws; a context that occurs several times in one window is one pool entry
// weighted by its count (the same sum in another order).
//
// Per step on MFMA (v_mfma_f32_16x16x4f32): the outputs are the 16 rows of
// the M dimension (target, at most 14 distinct negatives, target2), the distinct contexts are the N (or
// K) dimension in tiles of 16, the embedding dims the K (or N) dimension.
#include <algorithm>
#include <cstring>
#include <vector>

#include "dg_internal.h"
#include "dg_philox.h"

namespace {

constexpr int kDsglMaxSentence = 1000;  // MAX_SENTENCE_LENGTH (dsgl.hpp:23)
constexpr int kDsglMaxExp = 6;
constexpr int kDsglExpTable = 1000;
constexpr int kEntCap = 64;        // distinct contexts of one step (2 windows x 2W <= 60)

// G's row stride for window W: the step's distinct contexts are at most 4 W,
// in tiles of 16, plus 4 floats against bank conflicts.
__host__ __device__ inline uint32_t dsgl_gl(uint32_t window) { return 16u * ((4u * window + 15u) / 16u) + 4u; }

// Floats of the G buffer: yrows rows of gl, and at least the 64 ints the
// step's context scan and the negative de-duplication stage there (G is dead
// while they run).
__host__ __device__ inline size_t dsgl_g_floats(uint32_t yrows, uint32_t gl) {
  const size_t n = (size_t)yrows * gl;
  return n < 64 ? 64 : n;
}

typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

__device__ __forceinline__ f32x4_t mma4(float4 a, float4 b, f32x4_t c) {
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b.z, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b.w, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A load that reads the level-2 cache (values this wavefront wrote earlier
// in the launch with other lanes, or added by atomics).
__device__ __forceinline__ float load_l2(const float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until every memory operation this wavefront has issued is complete:
// its stores and atomics have reached the level-2 cache of its XCD, where
// the level-2 loads above read them.  The rows a wavefront reads back this
// way are its own scratch rows (and, in the deterministic order, every row),
// so an agent-scope release, which also writes the XCD's level-2 cache back
// to memory, is not needed; it made each step 38% slower.
__device__ __forceinline__ void drain() { __builtin_amdgcn_s_waitcnt(0); }

// Rows whose loads are issued together before any of them is used: the
// kernel runs at one or two wavefronts per SIMD (its LDS sets that), so a
// row at a time would expose one memory round trip per row.
constexpr int kRowBatch = 8;

__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// word j of the Philox stream (counter (j / 4, worker, c2, tag), key)
__device__ __forceinline__ uint32_t draw_word(uint64_t key, uint32_t worker, uint32_t c2,
                                              uint32_t tag, uint32_t j) {
  const dg_u32x4 o = dg_philox4x32_10(j >> 2, worker, c2, tag, (uint32_t)key, (uint32_t)(key >> 32));
  const uint32_t r = j & 3u;
  return r == 0 ? o.x0 : (r == 1 ? o.x1 : (r == 2 ? o.x2 : o.x3));
}

struct DsglArgs {
  const uint32_t* tok;
  const uint64_t* off;
  uint64_t n_walks;
  uint64_t words_num;      // tokens + one separator per walk
  const uint32_t* rank;
  uint32_t rank_size;
  const uint32_t* keep;
  const int32_t* table;
  uint32_t table_size;
  uint32_t vocab;
  const float* expt;
  float* syn0;
  float* syn1;
  uint32_t window;
  uint32_t K;
  uint32_t workers;
  uint32_t lcap;           // sentence tokens one buffer holds
  uint32_t vcap;           // distinct words one buffer holds (<= lcap)
  uint32_t ncap;           // negative draws of one batch at most (lcap * K)
  uint32_t yrows;          // output rows held: K + 2
  uint32_t gl;             // G row stride (dsgl_gl of the window)
  float alpha0;
  float total;             // (float)(iter * train_words + 1)
  uint64_t key;            // Philox key of the epoch
  int exact;
  int deterministic;
  float* scratch;          // per wavefront: sentence snapshots, batch snapshot and live rows
  uint64_t scratch_stride; // floats per wavefront
  unsigned long long* actual;  // per worker: counted words, carried across epochs
  float* alpha;            // per worker: learning rate, carried across epochs
  unsigned long long* kept;    // tokens kept by subsampling
  uint32_t* err;           // set when a sentence holds more than vcap distinct words
};

// LDS of one wavefront (dynamic; sizes follow lcap, vcap, ncap, yrows, gl
// and DIM).  Output rows past yrows, and their G rows, are read as the zero
// row Z.
struct DsglLds {
  float* V;       // [2][vcap][LDV] live sentence rows
  float* Y;       // [yrows][LDV] the step's output rows
  float* G;       // [yrows][gl]
  float* Z;       // [LDV] zeros
  int* sen;       // [2][lcap]
  int* slot;      // [2][lcap] word index of each position
  int* word;      // [2][vcap] distinct words
  int* nword;     // [ncap] negative draws of the batch
  int* nslot;     // [ncap] batch-buffer index of each draw
  int* nid;       // [ncap] distinct negatives
  int* ent_row;   // [kEntCap] row of each distinct context (s * vcap + word index)
  int* out_id;    // [16]
  int* out_meta;  // [16]
  int* out_slot;  // [16]
  uint8_t* ent_cnt;  // [kEntCap] occurrences of each distinct context (<= 32)
  int* tmp;       // [64], in G's storage
};

template <int DIM>
__host__ __device__ constexpr int dsgl_ldv() { return DIM + 4; }

template <int DIM>
__host__ __device__ inline size_t dsgl_lds_bytes(uint32_t lcap, uint32_t vcap, uint32_t ncap,
                                                 uint32_t yrows, uint32_t gl) {
  constexpr int LDV = dsgl_ldv<DIM>();
  size_t f = (size_t)2 * vcap * LDV + (size_t)yrows * LDV + dsgl_g_floats(yrows, gl) + LDV;
  size_t i = (size_t)2 * 2 * lcap + 2 * (size_t)vcap + 3 * (size_t)ncap + kEntCap + 3 * 16;
  return ((f * 4 + i * 4 + kEntCap) + 15) & ~(size_t)15;
}

template <int DIM>
__device__ DsglLds dsgl_carve(char* base, uint32_t lcap, uint32_t vcap, uint32_t ncap, uint32_t yrows,
                              uint32_t gl) {
  constexpr int LDV = dsgl_ldv<DIM>();
  DsglLds L;
  float* f = reinterpret_cast<float*>(base);
  L.V = f;
  f += (size_t)2 * vcap * LDV;
  L.Y = f;
  f += (size_t)yrows * LDV;
  L.G = f;
  L.tmp = reinterpret_cast<int*>(f);
  f += dsgl_g_floats(yrows, gl);
  L.Z = f;
  f += LDV;
  int* p = reinterpret_cast<int*>(f);
  L.sen = p;
  p += 2 * lcap;
  L.slot = p;
  p += 2 * lcap;
  L.word = p;
  p += 2 * vcap;
  L.nword = p;
  p += ncap;
  L.nslot = p;
  p += ncap;
  L.nid = p;
  p += ncap;
  L.ent_row = p;
  p += kEntCap;
  L.out_id = p;
  p += 16;
  L.out_meta = p;
  p += 16;
  L.out_slot = p;
  p += 16;
  L.ent_cnt = reinterpret_cast<uint8_t*>(p);
  return L;
}

__device__ __forceinline__ float dsgl_sig(float f, const float* expt, int exact) {
  if (exact) return (float)(1.0 / (1.0 + exp(-(double)f)));
  if (f > (float)kDsglMaxExp) return 1.0f;
  if (f < -(float)kDsglMaxExp) return 0.0f;
  return expt[(int)((f + (float)kDsglMaxExp) * (float)(kDsglExpTable / kDsglMaxExp / 2))];
}

// For lane i = i0 + lane of a list v[0, n): the first j < i with v[j] ==
// v[i], or -1.  Earlier entries are compared 64 at a time from a register
// (v_readlane), not by a loop of dependent LDS reads per lane.
__device__ __forceinline__ int first_earlier(const int* v, int n, int i0, int i) {
  const int lane = (int)(threadIdx.x & 63u);
  const int w = i < n ? v[i] : 0;
  int first_of = -1;
  for (int c0 = 0; c0 <= i0; c0 += 64) {
    const int cj = c0 + lane < n ? v[c0 + lane] : 0;
    const int kn = n - c0 < 64 ? n - c0 : 64;
    for (int k = 0; k < kn; ++k) {
      const int vk = __builtin_amdgcn_readlane(cj, k);
      if (first_of < 0 && c0 + k < i && vk == w) first_of = c0 + k;
    }
  }
  return i < n ? first_of : -1;
}

// Two wavefronts per SIMD up to dim 128 (at most 256 registers): the LDS of
// a wavefront (27 KB at dim 128 on 22-token walks) fits six per CU.
template <int DIM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(DIM <= 128 ? 2 : 1)))
void dsgl_kernel(DsglArgs a) {
  extern __shared__ __attribute__((aligned(16))) char dsgl_smem[];
  constexpr int LDV = dsgl_ldv<DIM>();
  constexpr int NT = DIM / 16;  // dim tiles / k chunks
  constexpr int PER = DIM / 64; // dims per lane in row copies
  const int lane = threadIdx.x & 63;
  const int l16 = lane & 15, g4 = lane >> 4;
  DsglLds L = dsgl_carve<DIM>(dsgl_smem, a.lcap, a.vcap, a.ncap, a.yrows, a.gl);
  for (int i = lane; i < LDV; i += 64) L.Z[i] = 0.0f;
  const uint64_t gw = blockIdx.x;
  uint64_t nw = gridDim.x;
  if (a.deterministic) {
    nw = 1;
    if (gw != 0) return;
  }
  float* snap = a.scratch + gw * a.scratch_stride;        // [2][lcap][DIM]
  float* nsnap = snap + (size_t)2 * a.lcap * DIM;          // [ncap][DIM]
  float* nlive = nsnap + (size_t)a.ncap * DIM;             // [ncap][DIM]
  const int W = (int)a.window, K = (int)a.K;
  const uint32_t lcap = a.lcap, vcap = a.vcap;
  const int yr = (int)a.yrows, gl = (int)a.gl;
  unsigned long long kept_total = 0;

  for (uint64_t t = gw; t < a.workers; t += nw) {
    // ---- the worker's word range (dsgl.hpp:814-825)
    const uint64_t q = a.words_num / a.workers;
    const uint64_t start = q * t;
    const uint64_t end = start + (t == a.workers - 1 ? a.words_num - (uint64_t)(a.workers - 1) * q : q);
    uint64_t cur = start;
    uint64_t walk;
    {
      uint64_t lo = 0, hi = a.n_walks;  // last walk w with off[w] + w <= start
      while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (a.off[mid] + mid <= start) lo = mid; else hi = mid;
      }
      walk = uni64(lo);
    }
    uint64_t wc = 0, last_wc = 0;
    unsigned long long actual = a.actual[t];
    float alpha = a.alpha[t];
    int len[2] = {0, 0}, pos[2] = {0, 0}, nwords[2] = {0, 0};
    int tsn = 0, change_neg = 1, nbn = 0, nneg = 0;
    uint32_t neg_count = 0, step = 0, batch = 0;

    while (true) {
      if (a.deterministic) __threadfence();
      // ---- the learning rate (dsgl.hpp:882-894)
      if (wc - last_wc > 10000) {
        actual += wc - last_wc;
        last_wc = wc;
        float al = a.alpha0 * (1.0f - (float)(actual * (unsigned long long)a.workers) / a.total);
        if (al < a.alpha0 * 0.0001f) al = a.alpha0 * 0.0001f;
        alpha = al;
      }
      // ---- sentences (dsgl.hpp:899-1030)
      for (int s = 0; s < 2; ++s) {
        while (len[s] == 0) {
          if (cur >= end) break;
          int n = 0;
          while (true) {
            uint64_t sep = uni64(a.off[walk + 1] + walk);
            while (sep < cur) {
              ++walk;
              sep = uni64(a.off[walk + 1] + walk);
            }
            // the sentence closes at the first separator or at the range's
            // last position, whichever comes first (never kept)
            const uint64_t stop = sep < end - 1 ? sep : end - 1;
            const uint64_t span = stop - cur;  // positions before the stop
            const uint64_t p = cur + (uint64_t)lane;
            bool kp = false;
            int r = -1;
            if ((uint64_t)lane < span) {
              const uint64_t ti = p - walk;
              const uint32_t v = a.tok[ti];
              const uint32_t rr = v < a.rank_size ? a.rank[v] : 0xffffffffu;
              if (rr != 0xffffffffu) {
                const dg_u32x4 o = dg_philox4x32_10(0, (uint32_t)ti, (uint32_t)(ti >> 32), DG_TAG_SAMPLE,
                                                    (uint32_t)a.key, (uint32_t)(a.key >> 32));
                kp = (o.x0 & 0xFFFFu) <= a.keep[rr];
                r = (int)rr;
              }
            }
            const uint64_t km = __ballot(kp);
            const int nk = __popcll(km);
            const int room = kDsglMaxSentence - n;
            if (nk >= room) {
              // the sentence reaches MAX_SENTENCE_LENGTH at the room-th kept token
              uint64_t m = km;
              for (int i = 1; i < room; ++i) m &= m - 1;
              const int last = __builtin_ctzll(m);
              if (kp && lane <= last) L.sen[s * lcap + n + (int)lanes_below(km)] = r;
              n += room;
              wc += (uint64_t)last + 1;
              cur += (uint64_t)last + 1;
              break;
            }
            if (kp) L.sen[s * lcap + n + (int)lanes_below(km)] = r;
            n += nk;
            if (span < 64) {
              wc += span + 1;
              cur += span + 1;
              break;
            }
            wc += 64;
            cur += 64;
          }
          kept_total += (unsigned long long)n;
          wave_sync();
          if (n <= 1) continue;
          len[s] = n;
          pos[s] = 0;
          // distinct words in order of first appearance, and each position's word
          // (64 positions at a time: a position's first occurrence takes the
          // next word index, a repeat the index of its first occurrence)
          int nwd = 0;
          const int* sen_s = L.sen + s * lcap;
          for (int a0 = 0; a0 < n; a0 += 64) {
            const int i = a0 + lane;
            const int first_of = first_earlier(sen_s, n, a0, i);
            const bool first = i < n && first_of < 0;
            const uint64_t fm = __ballot(first);
            if (first) {
              int widx = nwd + (int)lanes_below(fm);
              if (widx >= (int)vcap) {
                // more distinct words than the launch planned for: flagged
                // (the host fails the call), never written past the buffer
                *a.err = 1u;
                widx = 0;
              } else {
                L.word[s * vcap + widx] = sen_s[i];
              }
              L.slot[s * lcap + i] = widx;
            }
            nwd += __popcll(fm);
            wave_sync();
            if (i < n && !first) L.slot[s * lcap + i] = L.slot[s * lcap + first_of];
            wave_sync();
          }
          if (nwd > (int)vcap) nwd = (int)vcap;
          nwords[s] = nwd;
          // live rows in LDS, snapshots in scratch (dsgl.hpp:942-945)
          for (int u0 = 0; u0 < nwd; u0 += kRowBatch) {
            float v[kRowBatch][PER];
#pragma unroll
            for (int i = 0; i < kRowBatch; ++i) {
              if (u0 + i >= nwd) continue;
              const float* src = a.syn0 + (uint64_t)L.word[s * vcap + u0 + i] * DIM;
#pragma unroll
              for (int j = 0; j < PER; ++j) v[i][j] = src[lane + 64 * j];
            }
#pragma unroll
            for (int i = 0; i < kRowBatch; ++i) {
              const int u = u0 + i;
              if (u >= nwd) continue;
#pragma unroll
              for (int j = 0; j < PER; ++j) {
                L.V[(s * vcap + u) * LDV + lane + 64 * j] = v[i][j];
                snap[((size_t)s * lcap + u) * DIM + lane + 64 * j] = v[i][j];
              }
            }
          }
          wave_sync();
        }
      }
      if (cur >= end) {
        // the range is exhausted: sentences in hand and the batch are dropped
        actual += wc - last_wc;
        break;
      }
      // ---- a new negative batch (dsgl.hpp:1052-1083)
      if (change_neg) {
        nbn = len[0] > len[1] ? len[0] : len[1];
        const int nd = nbn * K;
        for (int i0 = 0; i0 < nd; i0 += 64) {
          const int i = i0 + lane;
          if (i < nd) {
            const uint32_t x = draw_word(a.key, (uint32_t)t, batch, DG_TAG_DSGL_NEG, (uint32_t)i);
            int smp = a.table[x % a.table_size];
            if (!smp) smp = (int)(x % (a.vocab - 1) + 1);
            L.nword[i] = smp;
          }
        }
        wave_sync();
        // distinct draws in order of first appearance
        int nd_first = 0;
        for (int i0 = 0; i0 < nd; i0 += 64) {
          const int i = i0 + lane;
          const int first_of = first_earlier(L.nword, nd, i0, i);
          const uint64_t fm = __ballot(i < nd && first_of < 0);
          if (i < nd && first_of < 0) {
            const int k = nd_first + (int)lanes_below(fm);
            L.nid[k] = L.nword[i];
            L.nslot[i] = k;
          }
          nd_first += __popcll(fm);
          wave_sync();
          if (i < nd && first_of >= 0) L.tmp[lane] = first_of;
          wave_sync();
          if (i < nd && first_of >= 0) L.nslot[i] = L.nslot[L.tmp[lane]];
          wave_sync();
        }
        nneg = nd_first;
        for (int k0 = 0; k0 < nneg; k0 += kRowBatch) {
          float v[kRowBatch][PER];
#pragma unroll
          for (int i = 0; i < kRowBatch; ++i) {
            if (k0 + i >= nneg) continue;
            const float* src = a.syn1 + (uint64_t)L.nid[k0 + i] * DIM;
#pragma unroll
            for (int j = 0; j < PER; ++j) v[i][j] = src[lane + 64 * j];
          }
#pragma unroll
          for (int i = 0; i < kRowBatch; ++i) {
            const size_t k = (size_t)(k0 + i);
            if (k0 + i >= nneg) continue;
#pragma unroll
            for (int j = 0; j < PER; ++j) {
              nsnap[k * DIM + lane + 64 * j] = v[i][j];
              nlive[k * DIM + lane + 64 * j] = v[i][j];
            }
          }
        }
        drain();
        ++batch;
        change_neg = 0;
      }
      // ---- one step: both positions (dsgl.hpp:1085-1356)
      const int target = uni(L.sen[pos[0]]);
      const int target2 = uni(L.sen[lcap + pos[1]]);
      // outputs: [target, distinct negatives with multiplicity, target2]
      ++neg_count;
      const int nst = (int)(neg_count % (uint32_t)nbn);
      int smp = -1, meta = 0, sl = 0;
      bool nfirst = false;
      if (lane < K) {
        smp = L.nword[nst * K + lane];
        sl = L.nslot[nst * K + lane];
        nfirst = smp != target;
        for (int k = 0; k < K; ++k) {
          const int o = L.nword[nst * K + k];
          if (o == smp) {
            ++meta;
            if (k < lane) nfirst = false;
          }
        }
      }
      const uint64_t om = __ballot(nfirst);
      const int O = 2 + __popcll(om);
      if (nfirst) {
        const int o = 1 + (int)lanes_below(om);
        L.out_id[o] = smp;
        L.out_meta[o] = meta;
        L.out_slot[o] = sl;
      }
      if (lane == 0) {
        L.out_id[0] = target;
        L.out_meta[0] = 1;
        L.out_id[O - 1] = target2;
        L.out_meta[O - 1] = 1;
      }
      wave_sync();
      // output rows: the targets from syn1, the negatives from the batch
      // buffer, from the level-2 cache (the rows this wavefront's atomics
      // moved); the first eight rows' loads are in flight while the contexts
      // are found
      float yv[8][PER];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
#pragma unroll
        for (int j = 0; j < PER; ++j) yv[r][j] = 0.0f;
        if (r < O) {
          const bool tab = (r == 0 || r == O - 1);
          const float* src = tab ? a.syn1 + (uint64_t)L.out_id[r] * DIM
                                 : nlive + (size_t)L.out_slot[r] * DIM;
#pragma unroll
          for (int j = 0; j < PER; ++j) yv[r][j] = load_l2(src + lane + 64 * j);
        }
      }
      const int b = (int)(draw_word(a.key, (uint32_t)t, step, DG_TAG_DSGL_WIN, 0) % (uint32_t)W);
      ++step;
      // contexts: lanes 0..31 the first window, 32..63 the second
      const int hs = lane >> 5, li = lane & 31;
      const int span = 2 * (W - b);
      int my = -1;
      {
        const int i = b + li + (b + li >= W ? 1 : 0);
        const int c = pos[hs] - W + i;
        if (li < span && c >= 0 && c < len[hs]) my = L.slot[hs * lcap + c];
      }
      L.tmp[lane] = my;
      wave_sync();
      bool first = my >= 0;
      int cnt = 0;
      for (int j = 0; j < 32; ++j) {
        const int o = L.tmp[hs * 32 + j];
        cnt += (my >= 0 && o == my) ? 1 : 0;
        if (j < li && o == my) first = false;
      }
      const uint64_t em = __ballot(first);
      const int E = __popcll(em);
      const int E1 = __popcll(em & 0xffffffffull);
      if (first) {
        const int e = (int)lanes_below(em);
        L.ent_row[e] = hs * (int)vcap + my;
        L.ent_cnt[e] = (uint8_t)cnt;
      }
#pragma unroll
      for (int r = 0; r < 8; ++r)
        if (r < yr)  // rows O .. yr - 1 zero
#pragma unroll
          for (int j = 0; j < PER; ++j) L.Y[r * LDV + lane + 64 * j] = yv[r][j];
      if (yr > 8) {
        // more than six negatives: rows 8 .. 15
        float yw[8][PER];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
#pragma unroll
          for (int j = 0; j < PER; ++j) yw[r][j] = 0.0f;
          if (8 + r < O) {
            const bool tab = 8 + r == O - 1;
            const float* src = tab ? a.syn1 + (uint64_t)L.out_id[8 + r] * DIM
                                   : nlive + (size_t)L.out_slot[8 + r] * DIM;
#pragma unroll
            for (int j = 0; j < PER; ++j) yw[r][j] = load_l2(src + lane + 64 * j);
          }
        }
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (8 + r < yr)
#pragma unroll
            for (int j = 0; j < PER; ++j) L.Y[(8 + r) * LDV + lane + 64 * j] = yw[r][j];
      }
      wave_sync();
      // LDS offsets (from V) of the context rows: xg[et] the row lane l16
      // supplies to G's tile et, xo[ec][x] the row 16 ec + 4 g4 + x (the k
      // index of dY's tile ec, the row of dX's output); the zero row past E.
      // The tiles' products are unrolled to kET (E <= kEntCap) so that their
      // MFMA chains and LDS reads interleave.
      constexpr int kET = kEntCap / 16;
      const int ET = (E + 15) >> 4;
      const int zoff = (int)(L.Z - L.V);
      int xg[kET], xo[kET][4];
#pragma unroll
      for (int et = 0; et < kET; ++et) {
        const int e = 16 * et + l16;
        xg[et] = e < E ? L.ent_row[e] * LDV : zoff;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          const int e4 = 16 * et + 4 * g4 + x;
          xo[et][x] = e4 < E ? L.ent_row[e4] * LDV : zoff;
        }
      }
      // G = Y X^T and g (dsgl.hpp:1241-1273)
      {
        f32x4_t acc[kET];
#pragma unroll
        for (int et = 0; et < kET; ++et) acc[et] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        const float* yrow = l16 < yr ? L.Y + l16 * LDV : L.Z;
#pragma unroll
        for (int kc = 0; kc < NT; ++kc) {
          const float4 ya = *reinterpret_cast<const float4*>(yrow + 16 * kc + 4 * g4);
#pragma unroll
          for (int et = 0; et < kET; ++et) {
            if (et >= ET) break;
            const float4 xb = *reinterpret_cast<const float4*>(L.V + xg[et] + 16 * kc + 4 * g4);
            acc[et] = mma4(ya, xb, acc[et]);
          }
        }
#pragma unroll
        for (int et = 0; et < kET; ++et) {
          if (et >= ET) break;
          const int e = 16 * et + l16;
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            const int i = 4 * g4 + x;
            float g = 0.0f;
            if (i < O && e < E) {
              const bool s1 = e < E1;
              const float label = ((i == 0 && s1) || (i == O - 1 && !s1)) ? 1.0f : 0.0f;
              g = (label - dsgl_sig(acc[et][x], a.expt, a.exact)) * alpha * (float)L.out_meta[i] *
                  (float)L.ent_cnt[e];
            }
            if (i < yr) L.G[i * gl + e] = g;
          }
        }
      }
      wave_sync();
      // dY = G X: lane holds dY[4 g4 + x][16 tt + l16], added to the row of
      // output 4 g4 + x (the targets in syn1, the negatives in the batch buffer)
      {
        float* dst[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          const int i = 4 * g4 + x;
          dst[x] = i >= O ? nullptr
                          : (i == 0 ? a.syn1 + (uint64_t)target * DIM
                                    : (i == O - 1 ? a.syn1 + (uint64_t)target2 * DIM
                                                  : nlive + (size_t)L.out_slot[i] * DIM));
        }
        f32x4_t acc[NT];
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) acc[tt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        const float* grow = l16 < yr ? L.G + l16 * gl : L.Z;
#pragma unroll
        for (int ec = 0; ec < kET; ++ec) {
          if (ec >= ET) break;
          const float4 ga = *reinterpret_cast<const float4*>(grow + 16 * ec + 4 * g4);
#pragma unroll
          for (int tt = 0; tt < NT; ++tt) {
            float xv[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) xv[x] = L.V[xo[ec][x] + 16 * tt + l16];
            acc[tt] = mma4(ga, make_float4(xv[0], xv[1], xv[2], xv[3]), acc[tt]);
          }
        }
#pragma unroll
        for (int tt = 0; tt < NT; ++tt)
#pragma unroll
          for (int x = 0; x < 4; ++x)
            if (dst[x]) atomicAdd(dst[x] + 16 * tt + l16, acc[tt][x]);
      }
      // dX = G^T Y into the sentence buffers: lane holds dX[16 et + 4 g4 + x][16 tt + l16]
      {
        // this lane's k rows 4 g4 .. 4 g4 + 3, rows past yr read from the zero row
        const float* yk[4];
        const float* gk[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          const int i = 4 * g4 + x;
          yk[x] = i < yr ? L.Y + i * LDV : L.Z;
          gk[x] = i < yr ? L.G + i * gl : L.Z;
        }
        float4 yb[NT];
#pragma unroll
        for (int tt = 0; tt < NT; ++tt)
          yb[tt] = make_float4(yk[0][16 * tt + l16], yk[1][16 * tt + l16], yk[2][16 * tt + l16], yk[3][16 * tt + l16]);
#pragma unroll
        for (int et = 0; et < kET; ++et) {
          if (et >= ET) break;
          const float4 a4 = make_float4(gk[0][16 * et + l16], gk[1][16 * et + l16], gk[2][16 * et + l16],
                                        gk[3][16 * et + l16]);
          f32x4_t acc[NT];
#pragma unroll
          for (int tt = 0; tt < NT; ++tt) acc[tt] = mma4(a4, yb[tt], f32x4_t{0.f, 0.f, 0.f, 0.f});
          wave_sync();
#pragma unroll
          for (int tt = 0; tt < NT; ++tt)
#pragma unroll
            for (int x = 0; x < 4; ++x)
              if (16 * et + 4 * g4 + x < E) L.V[xo[et][x] + 16 * tt + l16] += acc[tt][x];
        }
      }
      drain();
      wave_sync();
      // ---- finished sentences and batch back to the tables (dsgl.hpp:1358-1404)
      for (int s = 0; s < 2; ++s) {
        if (++pos[s] < len[s]) continue;
        for (int u0 = 0; u0 < nwords[s]; u0 += kRowBatch) {
          float sv[kRowBatch][PER];
#pragma unroll
          for (int i = 0; i < kRowBatch; ++i) {
            if (u0 + i >= nwords[s]) continue;
#pragma unroll
            for (int j = 0; j < PER; ++j)
              sv[i][j] = load_l2(snap + ((size_t)s * lcap + u0 + i) * DIM + lane + 64 * j);
          }
#pragma unroll
          for (int i = 0; i < kRowBatch; ++i) {
            const int u = u0 + i;
            if (u >= nwords[s]) continue;
            float* dst = a.syn0 + (uint64_t)L.word[s * vcap + u] * DIM;
#pragma unroll
            for (int j = 0; j < PER; ++j) {
              const int d = lane + 64 * j;
              atomicAdd(dst + d, L.V[(s * vcap + u) * LDV + d] - sv[i][j]);
            }
          }
        }
        len[s] = 0;
        ++tsn;
      }
      if (tsn >= 2) {
        for (int k0 = 0; k0 < nneg; k0 += kRowBatch) {
          float lv[kRowBatch][PER], sv[kRowBatch][PER];
#pragma unroll
          for (int i = 0; i < kRowBatch; ++i) {
            if (k0 + i >= nneg) continue;
            const size_t k = (size_t)(k0 + i);
#pragma unroll
            for (int j = 0; j < PER; ++j) {
              lv[i][j] = load_l2(nlive + k * DIM + lane + 64 * j);
              sv[i][j] = load_l2(nsnap + k * DIM + lane + 64 * j);
            }
          }
#pragma unroll
          for (int i = 0; i < kRowBatch; ++i) {
            if (k0 + i >= nneg) continue;
            float* dst = a.syn1 + (uint64_t)L.nid[k0 + i] * DIM;
#pragma unroll
            for (int j = 0; j < PER; ++j) atomicAdd(dst + lane + 64 * j, lv[i][j] - sv[i][j]);
          }
        }
        change_neg = 1;
        tsn = 0;
      }
      wave_sync();
    }
    if (lane == 0) {
      a.actual[t] = actual;
      a.alpha[t] = alpha;
    }
  }
  if (lane == 0 && kept_total) atomicAdd(a.kept, kept_total);
}

// The most distinct vertices any walk holds (out[0], by atomicMax): a
// sentence is a run of one walk's kept tokens, so its distinct words, which
// the sentence buffers hold, are at most this.  One wavefront per walk, 64
// positions at a time; position p counts when no earlier position of its
// walk holds the same vertex.
__global__ __launch_bounds__(256) void walk_distinct_kernel(const uint32_t* tok, const uint64_t* off,
                                                            uint64_t n_walks, uint32_t* out) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
  uint32_t best = 0;
  for (uint64_t w = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < n_walks; w += nw) {
    const uint64_t b = uni64(off[w]);
    const uint64_t L = uni64(off[w + 1]) - b;
    uint32_t cnt = 0;
    for (uint64_t a0 = 0; a0 < L; a0 += 64) {
      const bool va = a0 + (uint64_t)lane < L;
      const uint32_t x = va ? tok[b + a0 + lane] : 0u;
      bool first = va;
      for (uint64_t b0 = 0; b0 <= a0; b0 += 64) {
        const uint32_t y = b0 == a0 ? x : tok[b + b0 + lane];  // b0 < a0: a whole earlier chunk
        const int jn = (int)(L - b0 < 64 ? L - b0 : 64);
        for (int j = 0; j < jn; ++j) {
          const uint32_t yj = (uint32_t)__shfl((int)y, j, 64);
          if (b0 + (uint64_t)j < a0 + (uint64_t)lane && yj == x) first = false;
        }
      }
      cnt += (uint32_t)__popcll(__ballot(first));
    }
    best = cnt > best ? cnt : best;
  }
  if (lane == 0 && best) atomicMax(out, best);
}

__global__ void dsgl_flags_init_kernel(uint32_t* flags) {
  if (threadIdx.x < 2) flags[threadIdx.x] = 0u;
}

__global__ void dsgl_state_init_kernel(unsigned long long* actual, float* alpha, uint32_t n,
                                       float alpha0, unsigned long long* kept) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    actual[i] = 0ull;
    alpha[i] = alpha0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *kept = 0ull;
}

template <int DIM>
int launch_dims(dg_ctx* ctx, const DsglArgs& a, size_t lds, unsigned grid) {
  auto k = dsgl_kernel<DIM>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return dg_fail(ctx, DG_ERR_HIP, "DSGL kernel LDS attribute: %s", hipGetErrorString(e));
  k<<<grid, 64, lds, ctx->stream>>>(a);
  e = hipGetLastError();
  if (e != hipSuccess) return dg_fail(ctx, DG_ERR_HIP, "DSGL kernel launch: %s", hipGetErrorString(e));
  return DG_OK;
}

template <int DIM>
int plan_dims(dg_ctx* ctx, uint32_t lcap, uint32_t vcap, uint32_t ncap, uint32_t yrows, uint32_t gl,
              size_t* lds, unsigned* per_cu) {
  *lds = dsgl_lds_bytes<DIM>(lcap, vcap, ncap, yrows, gl);
  if (*lds > 160 * 1024)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED,
                   "DSGL form: sentences of %u tokens at dim %d need %zu bytes of LDS per wavefront "
                   "(160 KB available)", lcap, DIM, *lds);
  int n = 0;
  hipError_t e = hipFuncSetAttribute((const void*)dsgl_kernel<DIM>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)*lds);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dsgl_kernel<DIM>, 64, *lds);
  if (e != hipSuccess) return dg_fail(ctx, DG_ERR_HIP, "DSGL occupancy: %s", hipGetErrorString(e));
  *per_cu = n < 1 ? 1u : (unsigned)n;
  return DG_OK;
}

}  // namespace

struct dg_dsgl_state {
  float* scratch = nullptr;
  size_t scratch_cap = 0;
  unsigned long long* actual = nullptr;
  size_t actual_cap = 0;
  float* alpha = nullptr;
  size_t alpha_cap = 0;
  unsigned long long* kept = nullptr;
  uint32_t* flags = nullptr;  // [0] the most distinct vertices of a walk, [1] the kernel's overflow flag
};

void dg_dsgl_free(dg_ctx* ctx) {
  dg_dsgl_state* s = ctx->dsgl;
  if (!s) return;
  if (s->scratch) hipFree(s->scratch);
  if (s->actual) hipFree(s->actual);
  if (s->alpha) hipFree(s->alpha);
  if (s->kept) hipFree(s->kept);
  if (s->flags) hipFree(s->flags);
  delete s;
  ctx->dsgl = nullptr;
}

int dg_dsgl_train(dg_ctx* ctx, const dg_dsgl_call& c, uint64_t* kept, double* ms, float* alpha_end) {
  if (c.dim != 64 && c.dim != 128 && c.dim != 256)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "DSGL form: dim %u (64, 128 or 256)", c.dim);
  if (!ctx->dsgl) ctx->dsgl = new dg_dsgl_state();
  dg_dsgl_state* st = ctx->dsgl;
  const uint32_t lcap = c.max_walk < (uint64_t)kDsglMaxSentence ? (uint32_t)(c.max_walk ? c.max_walk : 1)
                                                                 : (uint32_t)kDsglMaxSentence;
  const uint32_t ncap = lcap * (c.K ? c.K : 1);
  // the sentence buffers hold a walk's distinct vertices, fewer than its
  // positions when walks revisit vertices (12.3 of 22 on average for the
  // 2.2M-vertex Chung-Lu walks, 18 at most): the smaller LDS footprint fits
  // more wavefronts per CU
  if (!st->flags) DG_HIP(ctx, hipMalloc((void**)&st->flags, 2 * sizeof(uint32_t)));
  size_t tk;
  DG_CHECK(dg_time_begin(ctx, "sgns_dsgl", &tk));
  dsgl_flags_init_kernel<<<1, 64, 0, ctx->stream>>>(st->flags);
  {
    const uint64_t waves = std::min<uint64_t>(c.n_walks, (uint64_t)ctx->num_cus * 32);
    walk_distinct_kernel<<<(unsigned)((waves + 3) / 4), 256, 0, ctx->stream>>>(c.tok, c.off, c.n_walks,
                                                                              st->flags);
  }
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tk));
  uint32_t most = 0;
  DG_HIP(ctx, hipMemcpyAsync(&most, st->flags, sizeof(most), hipMemcpyDeviceToHost, ctx->stream));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const uint32_t vcap = std::max<uint32_t>(1u, std::min<uint32_t>(lcap, most));
  const uint32_t yrows = c.K + 2, gl = dsgl_gl(c.window);
  size_t lds = 0;
  unsigned per_cu = 1;
  switch (c.dim) {
    case 64: DG_CHECK(plan_dims<64>(ctx, lcap, vcap, ncap, yrows, gl, &lds, &per_cu)); break;
    case 128: DG_CHECK(plan_dims<128>(ctx, lcap, vcap, ncap, yrows, gl, &lds, &per_cu)); break;
    default: DG_CHECK(plan_dims<256>(ctx, lcap, vcap, ncap, yrows, gl, &lds, &per_cu)); break;
  }
  const uint64_t resident = (uint64_t)per_cu * (uint64_t)ctx->num_cus;
  // default: one worker per resident wavefront, at most five per CU.  Every
  // worker has the same share of the corpus, so the launch lasts as long as
  // its slowest wavefront; at dim 128 on the 2.2M-vertex round six per CU
  // (two SIMDs holding two) ran 33% slower than five, four 17% slower.
  const uint64_t dflt = std::min<uint64_t>(std::min<uint64_t>(per_cu, 5u) * (uint64_t)ctx->num_cus, 1u << 20);
  uint32_t workers = c.workers ? c.workers : (uint32_t)dflt;
  if ((uint64_t)workers > c.n_walks + c.n_tok) workers = (uint32_t)std::max<uint64_t>(1, (c.n_walks + c.n_tok) / 64);
  const unsigned grid = c.deterministic ? 1u : (unsigned)std::min<uint64_t>(workers, resident);
  const uint64_t stride = ((uint64_t)2 * lcap + 2 * (uint64_t)ncap) * c.dim;
  DG_CHECK(dg_reserve(ctx, &st->scratch, &st->scratch_cap, (size_t)(grid * stride)));
  DG_CHECK(dg_reserve(ctx, &st->actual, &st->actual_cap, (size_t)workers));
  DG_CHECK(dg_reserve(ctx, &st->alpha, &st->alpha_cap, (size_t)workers));
  if (!st->kept) DG_HIP(ctx, hipMalloc((void**)&st->kept, sizeof(unsigned long long)));
  dsgl_state_init_kernel<<<(workers + 255) / 256, 256, 0, ctx->stream>>>(st->actual, st->alpha, workers,
                                                                          c.alpha0, st->kept);
  DG_HIP(ctx, hipGetLastError());
  DsglArgs a;
  a.tok = c.tok;
  a.off = c.off;
  a.n_walks = c.n_walks;
  a.words_num = c.n_tok + c.n_walks;
  a.rank = c.rank;
  a.rank_size = c.rank_size;
  a.keep = c.keep;
  a.table = c.table;
  a.table_size = c.table_size;
  a.vocab = c.vocab;
  a.expt = c.expt;
  a.syn0 = c.syn0;
  a.syn1 = c.syn1;
  a.window = c.window;
  a.K = c.K;
  a.workers = workers;
  a.lcap = lcap;
  a.vcap = vcap;
  a.ncap = ncap;
  a.yrows = yrows;
  a.gl = gl;
  a.err = st->flags + 1;
  a.alpha0 = c.alpha0;
  a.total = (float)((uint64_t)c.iter * c.train_words + 1);
  a.exact = c.exact;
  a.deterministic = c.deterministic;
  a.scratch = st->scratch;
  a.scratch_stride = stride;
  a.actual = st->actual;
  a.alpha = st->alpha;
  a.kept = st->kept;
  double total_ms = 0.0;
  for (uint32_t ep = 0; ep < c.iter; ++ep) {
    a.key = dg_epoch_key(c.seed, c.epoch0 + ep);
    DG_CHECK(dg_time_begin(ctx, "sgns_dsgl", &tk));
    switch (c.dim) {
      case 64: DG_CHECK(launch_dims<64>(ctx, a, lds, grid)); break;
      case 128: DG_CHECK(launch_dims<128>(ctx, a, lds, grid)); break;
      default: DG_CHECK(launch_dims<256>(ctx, a, lds, grid)); break;
    }
    DG_CHECK(dg_time_end(ctx, tk));
  }
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& pe : ctx->pending) {
    float m = 0.f;
    if (pe.name == "sgns_dsgl") {
      DG_HIP(ctx, hipEventElapsedTime(&m, pe.start, pe.stop));
      total_ms += m;
    }
  }
  uint32_t over = 0;
  DG_HIP(ctx, hipMemcpy(&over, st->flags + 1, sizeof(over), hipMemcpyDeviceToHost));
  if (over)
    return dg_fail(ctx, DG_ERR_STATE,
                   "DSGL form: a sentence held more than %u distinct words (the corpus changed during the call?)",
                   vcap);
  unsigned long long k = 0;
  DG_HIP(ctx, hipMemcpy(&k, st->kept, sizeof(k), hipMemcpyDeviceToHost));
  *kept = k;
  *ms = total_ms;
  DG_HIP(ctx, hipMemcpy(alpha_end, st->alpha + (workers - 1), sizeof(float), hipMemcpyDeviceToHost));
  return DG_OK;
}
