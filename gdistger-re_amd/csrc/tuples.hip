// tuples.hip - the differentiable SGNS tuple interface on gfx950: an explicit
// forward (logits, loss) and a hand-written adjoint whose row gradients are
// summed by a sorted-key segmented reduction, so that the result does not
// depend on the schedule (no Hogwild races).
//
// A tuple is (center, context, neg_1..neg_K) in table rows; x = syn0[context],
// y_0 = syn1[center], y_k = syn1[neg_k]; z_j = x . y_j and
//   L = sum_t [ softplus(-z_0) + sum_k softplus(z_k) ].
// skip_gram_kernel (src/word2vec.cu:359-396) takes one SGD step on the same
// pairs in place: g_j = (label_j - sig(z_j)) * alpha, y_j += g_j x,
// x += sum_j g_j y_j, i.e. W -= alpha * dL/dW for one tuple at a time; a
// negative equal to the centre is skipped (src/word2vec.cu:366), here too.
//
// Pipeline of a backward call (n tuples, N = n (2 + K) gradient entries):
//   1. tuple_fwd_kernel<bwd>: a group of lanes per tuple gathers x and the
//      1+K y rows, reduces the 1+K dots together (a butterfly that halves
//      the value set at each step), and writes dL/dz, the context row's
//      gradient sum_j dz_j y_j (a staged row per tuple) and N (key, entry)
//      pairs: key = table bit | row.  It also clears the sort's counters.
//   2. stable LSD radix sort of the pairs by key, 8 bits per pass: one
//      histogram launch, then one launch per pass whose tiles find their
//      offsets by look-back.  Entries of one row become a segment, in tuple
//      order.
//   3. tuple_seg_kernel: the sorted entries are cut into blocks of 64; one
//      wavefront sums each block's runs of equal keys in order.  A run that
//      starts and ends inside its block is a finished row gradient, stored
//      at the run's first sorted position; a run that crosses a block edge
//      leaves its partial sum, and tuple_fix_kernel adds a crossing row's
//      partials in block order.  Hot rows (Zipf heads) are thus split over
//      many wavefronts with a fixed summation order.
//   4. apply (optional, in tuple_fix_kernel): W -= lr * grad per row, after
//      every gradient input was read from the old rows; the BF16 shadow rows
//      follow.
// BF16 form: x and y are read from BF16 copies of the tables (half the
// gather bytes), products and sums in FP32, updates to the FP32 rows.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "dg_internal.h"

namespace {

constexpr int kSegBlock = 64;  // sorted entries per wavefront in the segmented sum
constexpr int kMaxTupleNeg = 15;

__device__ __forceinline__ uint16_t to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// D FP32 values as BF16 at p (p aligned to 2 D bytes): pairs as 32-bit stores.
template <int D>
__device__ __forceinline__ void store_bf16(uint16_t* p, const float (&v)[D]) {
  if constexpr (D % 2 == 0) {
#pragma unroll
    for (int q = 0; q < D; q += 2)
      *reinterpret_cast<uint32_t*>(p + q) = (uint32_t)to_bf16(v[q]) | ((uint32_t)to_bf16(v[q + 1]) << 16);
  } else {
#pragma unroll
    for (int q = 0; q < D; ++q) p[q] = to_bf16(v[q]);
  }
}

// D consecutive elements of a row starting at element `first`, as floats.
template <typename E, int D>
struct RowIO;

template <int D>
struct RowIO<float, D> {
  __device__ static void load(const float* p, float (&v)[D]) {
    if constexpr (D == 1) {
      v[0] = p[0];
    } else if constexpr (D == 2) {
      const float2 q = *reinterpret_cast<const float2*>(p);
      v[0] = q.x;
      v[1] = q.y;
    } else {
      static_assert(D % 4 == 0, "D");
#pragma unroll
      for (int i = 0; i < D; i += 4) {
        const float4 q = *reinterpret_cast<const float4*>(p + i);
        v[i] = q.x;
        v[i + 1] = q.y;
        v[i + 2] = q.z;
        v[i + 3] = q.w;
      }
    }
  }
};

template <int D>
struct RowIO<uint16_t, D> {
  __device__ static void load(const uint16_t* p, float (&v)[D]) {
    if constexpr (D == 1) {
      v[0] = __uint_as_float((uint32_t)p[0] << 16);
    } else if constexpr (D == 2) {
      const uint32_t q = *reinterpret_cast<const uint32_t*>(p);
      v[0] = __uint_as_float(q << 16);
      v[1] = __uint_as_float(q & 0xffff0000u);
    } else if constexpr (D % 8 == 0) {
#pragma unroll
      for (int i = 0; i < D; i += 8) {
        const uint4 q = *reinterpret_cast<const uint4*>(p + i);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          v[i + 2 * k] = __uint_as_float(w[k] << 16);
          v[i + 2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
        }
      }
    } else {
      static_assert(D % 4 == 0, "D");
#pragma unroll
      for (int i = 0; i < D; i += 4) {
        const uint2 q = *reinterpret_cast<const uint2*>(p + i);
        v[i] = __uint_as_float(q.x << 16);
        v[i + 1] = __uint_as_float(q.x & 0xffff0000u);
        v[i + 2] = __uint_as_float(q.y << 16);
        v[i + 3] = __uint_as_float(q.y & 0xffff0000u);
      }
    }
  }
};

template <int D>
__device__ __forceinline__ void store_row(float* p, const float (&v)[D]) {
  if constexpr (D == 1) {
    p[0] = v[0];
  } else if constexpr (D == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  } else {
#pragma unroll
    for (int i = 0; i < D; i += 4)
      *reinterpret_cast<float4*>(p + i) = make_float4(v[i], v[i + 1], v[i + 2], v[i + 3]);
  }
}

// Sum NV per-lane values over each group of G lanes (G = 32 or 16) at once.
// Each butterfly step exchanges half of the remaining values with the lane
// `off` away, so after log2(NV) steps every lane holds one value index,
// idx = (lane % G) / (G / NV), summed over the lanes that differ in the
// exchanged bits; the remaining steps finish the sum.  Every lane of a group
// then holds its group's full sum of value idx.
template <int NV, int G>
__device__ __forceinline__ float group_multi_sum(float (&p)[NV], uint32_t lane) {
  static_assert(NV <= G, "values per group");
#pragma unroll
  for (int h = NV / 2, off = G / 2; h >= 1; h >>= 1, off >>= 1) {
    const bool up = (lane & (uint32_t)off) != 0;
#pragma unroll
    for (int i = 0; i < h; ++i) {
      const float keep = up ? p[i + h] : p[i];
      const float send = up ? p[i] : p[i + h];
      p[i] = keep + __shfl_xor(send, off, 64);
    }
  }
  float v = p[0];
#pragma unroll
  for (int off = G / 2 / NV; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ float softplus_f(float a) {
  return a > 0.0f ? a + log1pf(expf(-a)) : log1pf(expf(a));
}

struct TupleArgs {
  const uint32_t* tuples;  // n x (2 + K)
  uint64_t n;
  uint32_t K;
  const void* xt;          // syn0 (float or BF16 shadow)
  const void* yt;          // syn1
  float* logits;           // forward (and the fused step's backward): n x (1 + K)
  double* wloss;           // forward: one loss partial per wavefront of the grid; null in a
                           // plain backward
  const float* dlogits;    // backward: upstream dL/dz, or null for the loss's own
  float* dz;               // backward: n x (1 + K)
  float* stage;            // backward: n x dim, context-row gradient per tuple
  uint32_t* keys;          // backward: N
  uint32_t* vals;          // backward: N
  uint32_t row_bits;
  uint32_t* clear;         // backward: words the sort's launches expect zero
  uint64_t n_clear;
};

// 64 / G tuples per wavefront, one per group of G lanes (D = dim / G
// consecutive elements per lane): G = 32 makes an FP32 row one 16-byte load
// per lane at dim 128, G = 16 does the same for a BF16 row (four tuples per
// wavefront, twice the rows in flight).  The next group of tuples' ids are
// loaded before the current one is reduced.
// OUT: the logits and the wavefront's loss partial (the forward, and the
// backward of the fused step).  BWD: dL/dz, the staged
// context-row gradient and the sort pairs (entries [0, n) for syn0 rows,
// [n, N) for syn1 rows, in tuple order).
// Register target: three wavefronts per SIMD (168 VGPRs).  The FP32 backward
// with the forward's outputs (the fused step) otherwise takes 172 and runs
// two, 54 us instead of 43 at 65,536 tuples.
template <typename E, int DIM, int NV, bool BWD, int G, bool OUT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void tuple_fwd_kernel(TupleArgs a) {
  constexpr int T = 64 / G;                 // tuples per wavefront
  constexpr int D = DIM / G;
  constexpr int kStride = G / NV;           // lanes per value index after the sum
  constexpr bool kKeepY = NV * D <= 64;     // else the stage pass reloads y rows
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t hl = lane & (uint32_t)(G - 1), hb = lane & ~(uint32_t)(G - 1);  // lane within the group, group base
  const uint32_t grp = lane / G;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const E* xt = static_cast<const E*>(a.xt);
  const E* yt = static_cast<const E*>(a.yt);
  const uint32_t K1 = a.K + 1, W = a.K + 2;
  const uint32_t idx = hl / kStride;  // value index this lane ends with
  const uint64_t groups = (a.n + T - 1) / T;
  double wsum = 0.0;  // forward: this wavefront's loss, its tuples in processing order
  if constexpr (BWD) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n_clear;
         i += (uint64_t)gridDim.x * blockDim.x)
      a.clear[i] = 0u;
  }
  uint64_t pr = wave;
  uint32_t id = 0;
  if (pr < groups) {
    const uint64_t t = (uint64_t)T * pr + grp;
    id = (t < a.n && hl < W) ? a.tuples[t * W + hl] : 0u;
  }
  for (; pr < groups; pr += nwaves) {
    const uint64_t t = (uint64_t)T * pr + grp;
    const bool tv = t < a.n;  // this group has a tuple (the last wavefront may be short)
    const uint32_t center = __shfl(id, (int)hb, 64);
    const uint32_t ctx = __shfl(id, (int)hb + 1, 64);
    const uint32_t my_target = __shfl(id, (int)hb + (idx == 0 ? 0 : (int)(1 + idx) & (G - 1)), 64);
    const bool my_live = tv && idx < K1 && (idx == 0 || my_target != center);
    float x[D];
    if (tv) RowIO<E, D>::load(xt + (uint64_t)ctx * DIM + hl * D, x);
    float y[kKeepY ? NV : 1][D];
    float p[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      p[j] = 0.0f;
      if ((uint32_t)j < K1) {
        const uint32_t r = __shfl(id, (int)hb + (j == 0 ? 0 : 1 + j), 64);
        float yj[D];
        if (tv) {
          RowIO<E, D>::load(yt + (uint64_t)r * DIM + hl * D, yj);
#pragma unroll
          for (int q = 0; q < D; ++q) p[j] += x[q] * yj[q];
        }
        if constexpr (kKeepY) {
#pragma unroll
          for (int q = 0; q < D; ++q) y[j][q] = tv ? yj[q] : 0.0f;
        }
      }
    }
    // ids of the next group of tuples, in flight during the sums below
    const uint64_t pn = pr + nwaves;
    uint32_t id_next = 0;
    if (pn < groups) {
      const uint64_t tn = (uint64_t)T * pn + grp;
      id_next = (tn < a.n && hl < W) ? a.tuples[tn * W + hl] : 0u;
    }
    const float z = group_multi_sum<NV, G>(p, lane);
    const bool writer = tv && (hl % kStride) == 0 && idx < K1;
    if constexpr (OUT) {  // the forward's outputs (also from the backward of the fused step)
      if (writer) a.logits[t * K1 + idx] = z;
      // loss: one lane per live value index, summed in index order
      const float l = (writer && my_live) ? softplus_f(idx == 0 ? -z : z) : 0.0f;
      float acc = 0.0f;
#pragma unroll
      for (int j = 0; j < NV; ++j) acc += __shfl(l, (int)hb + j * kStride, 64);
#pragma unroll
      for (int q = 0; q < T; ++q) {  // the step's tuples in group order, on every lane
        const float lq = __shfl(acc, q * G, 64);
        if ((uint64_t)T * pr + q < a.n) wsum += (double)lq;
      }
    }
    if constexpr (BWD) {
      float dz = 0.0f;
      if (my_live) {
        if (a.dlogits)
          dz = a.dlogits[t * K1 + idx];
        else
          dz = 1.0f / (1.0f + expf(-z)) - (idx == 0 ? 1.0f : 0.0f);
      }
      if (writer) {
        a.dz[t * K1 + idx] = dz;
        a.keys[a.n + t * K1 + idx] = (1u << a.row_bits) | my_target;
        a.vals[a.n + t * K1 + idx] = (uint32_t)(t * K1 + idx);
      }
      float c[D];
#pragma unroll
      for (int q = 0; q < D; ++q) c[q] = 0.0f;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        if ((uint32_t)j < K1) {
          const float gj = __shfl(dz, (int)hb + j * kStride, 64);
          if constexpr (kKeepY) {
#pragma unroll
            for (int q = 0; q < D; ++q) c[q] += gj * y[j][q];
          } else {
            const uint32_t r = __shfl(id, (int)hb + (j == 0 ? 0 : 1 + j), 64);
            if (tv) {
              float yj[D];
              RowIO<E, D>::load(yt + (uint64_t)r * DIM + hl * D, yj);
#pragma unroll
              for (int q = 0; q < D; ++q) c[q] += gj * yj[q];
            }
          }
        }
      }
      if (tv) {
        store_row<D>(a.stage + t * DIM + hl * D, c);
        if (hl == 0) {
          a.keys[t] = ctx;
          a.vals[t] = (uint32_t)t;
        }
      }
    }
    id = id_next;
  }
  if constexpr (OUT) {
    if (lane == 0) a.wloss[wave] = wsum;
  }
}

// Deterministic loss: the forward's wavefronts each sum their tuples in a
// fixed order; one block of kLossThreads adds the wavefront partials, thread
// i those at i, i + kLossThreads, ... in order (eight loads in flight), then
// a fixed tree.  For a given tuple count the grid, and so the order, is
// fixed on a device.
constexpr int kLossThreads = 1024;

__global__ __launch_bounds__(kLossThreads) void loss_sum_kernel(const double* wloss, uint32_t count, double* out) {
  __shared__ double part[kLossThreads];
  double s = 0.0;
  for (uint32_t i0 = threadIdx.x; i0 < count; i0 += 8 * kLossThreads) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t i = i0 + (uint32_t)u * kLossThreads;
      v[u] = i < count ? wloss[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = kLossThreads / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

// ---- stable LSD radix sort of (key, value) pairs, 8 bits per pass ------------
// One histogram launch counts the digits of every pass; each pass is then one
// launch.  Tiles are numbered by workgroup id when every tile of the pass is
// resident at once (no more tiles than CUs), otherwise they take consecutive
// ids from a ticket counter in the order they start; either way a tile's
// predecessors have started before it.  A tile ranks its keys by digit, then
// publishes, per digit, its count in its tile word (flag | 30-bit count, one
// 32-bit word) and adds it to the word of its group of kGroup tiles (tiles
// added << 32 | count sum, one 64-bit atomic), so relaxed device-scope
// accesses order both.  Its prefix
// over earlier tiles is the sum of the earlier groups' words (each read once
// all its tiles have added) and of the earlier tiles of its own group: a few
// independent loads instead of a chain through every tile, and a tile only
// waits for tiles that started before it, so every wait ends.  Each wait is
// bounded by a poll limit and each store by the output size: a limit hit
// sets the error word the host reads after the call.
constexpr int kSortThreads = 1024;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortIpt = 4;                          // keys per lane per tile
constexpr int kSortTile = kSortThreads * kSortIpt;   // 4,096 keys
constexpr int kGroup = 16;                           // tiles per group word
constexpr int kHistChunk = 2048;                     // keys per histogram workgroup
constexpr int kMaxPasses = 4;
constexpr uint32_t kCountMask = (1u << 30) - 1u;
constexpr uint32_t kFlagReady = 1u << 30;
constexpr uint32_t kSpinLimit = 1u << 22;            // polls of one word (seconds, never reached)

__device__ __forceinline__ uint32_t lb_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t lb_load(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_add(uint64_t* p, uint64_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A failed look-back wait or output bound: the code goes to the host word
// (read after the call) and to a device word (the update kernel then leaves
// the tables unchanged).
struct SortErr {
  uint32_t* host;
  uint32_t* dev;
};

__device__ __forceinline__ void sort_fail(const SortErr& e, uint32_t code) {
  *e.host = code;
  *e.dev = code;
}

// Look-back words of one column (a digit): tile words tw[t * stride] and
// group words gw[g * stride].
struct LbColumn {
  uint32_t* tw;
  uint64_t* gw;
  uint32_t stride;
};

// Publish a tile's count in its tile word and its group's word.
__device__ __forceinline__ void lb_publish(const LbColumn& c, uint32_t tile, uint32_t count) {
  lb_store(c.tw + (uint64_t)tile * c.stride, kFlagReady | count);
  lb_add(c.gw + (uint64_t)(tile / kGroup) * c.stride, (1ull << 32) | count);
}

// Part of a tile's exclusive prefix: items j = first, first + step, ... of
// the list [earlier groups..., earlier tiles of the own group...], each read
// once ready (all loads of a round in flight together).
template <int R>
__device__ uint32_t lb_prefix_part(const LbColumn& c, uint32_t tile, uint32_t first, uint32_t step,
                                   const SortErr& err) {
  const uint32_t g = tile / kGroup, t0 = g * kGroup;
  const uint32_t m = g + (tile - t0);
  uint32_t part = 0;
  for (uint32_t j0 = first; j0 < m; j0 += step * R) {
    uint64_t w[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const uint32_t j = j0 + step * u;
      w[u] = j >= m ? 0ull : j < g ? lb_load(c.gw + (uint64_t)j * c.stride)
                                   : (uint64_t)lb_load(c.tw + (uint64_t)(t0 + j - g) * c.stride);
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const uint32_t j = j0 + step * u;
      if (j >= m) continue;
      uint32_t polls = 0;
      while (j < g ? (w[u] >> 32) != (uint64_t)kGroup : (w[u] & kFlagReady) == 0ull) {
        if (++polls > kSpinLimit) {
          sort_fail(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        w[u] = j < g ? lb_load(c.gw + (uint64_t)j * c.stride)
                     : (uint64_t)lb_load(c.tw + (uint64_t)(t0 + j - g) * c.stride);
      }
      part += j < g ? (uint32_t)w[u] : (uint32_t)w[u] & kCountMask;
    }
  }
  return part;
}

// Exclusive scan of one value per thread over the workgroup (at most 16
// wavefronts).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= (uint32_t)o) inc += u;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wave; ++w) before += wsum[w];
  __syncthreads();
  return before + inc - v;
}

// Digit counts of every pass over the whole input, in kHistCopies copies
// ghist[copy][pass][digit] (workgroup w adds to copy w % kHistCopies, which
// spreads the adds to the hot digits of skewed keys), cleared by the launch
// that wrote the keys: counted per workgroup in LDS, then added.
constexpr int kHistCopies = 16;

__global__ __launch_bounds__(256) void radix_hist_kernel(const uint32_t* keys, uint64_t n, int passes,
                                                         uint32_t* ghist) {
  __shared__ uint32_t h[kMaxPasses][256];
  for (int p = 0; p < kMaxPasses; ++p) h[p][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t lo = (uint64_t)blockIdx.x * kHistChunk;
#pragma unroll
  for (int i = 0; i < kHistChunk / 256; ++i) {
    const uint64_t k = lo + (uint64_t)i * 256 + threadIdx.x;
    if (k < n) {
      const uint32_t key = keys[k];
      for (int p = 0; p < passes; ++p) atomicAdd(&h[p][(key >> (8 * p)) & 255u], 1u);
    }
  }
  __syncthreads();
  uint32_t* g = ghist + (uint64_t)(blockIdx.x % kHistCopies) * kMaxPasses * 256;
  for (int p = 0; p < passes; ++p)
    if (h[p][threadIdx.x]) atomicAdd(&g[p * 256 + threadIdx.x], h[p][threadIdx.x]);
}

struct SortPass {
  const uint32_t* keys;
  const uint32_t* vals;
  uint64_t n;
  int pass;
  const uint32_t* ghist;  // [copy][pass][digit] over the whole input
  uint32_t* tile_words;   // this pass's [tile][digit]
  uint64_t* group_words;  // this pass's [group][digit]
  uint32_t* ticket;       // null: tile = workgroup id (every tile is resident at once)
  SortErr err;
  uint32_t* okeys;
  uint32_t* ovals;
};

// One pass: wavefront w of a tile ranks keys [w * 256, (w + 1) * 256) of the
// tile, 64 at a time in order; lanes with the same digit are ranked by a
// ballot match and the wavefront's running digit counts live in LDS (one
// wavefront reads and writes its own counters, in order).  Per digit: the
// offsets of the wavefronts within the tile, the tile's count, its prefix
// over earlier tiles (four threads per digit) and the digit's global base.
// The tile is then laid out in LDS in output order, so that each digit's
// keys leave as one contiguous run.
__global__ __launch_bounds__(kSortThreads) void radix_pass_kernel(SortPass a) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t wcnt[kSortWaves][256];
  __shared__ uint32_t dbase[256], lbase[256];
  __shared__ uint32_t wsum[kSortWaves];
  __shared__ uint32_t skey[kSortTile], sval[kSortTile];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_tile = a.ticket ? atomicAdd(a.ticket, 1u) : blockIdx.x;
  // keys with digit d in the whole input, summed below (threads d < 256)
  uint32_t gv[kHistCopies];
  if (threadIdx.x < 256u) {
#pragma unroll
    for (int r = 0; r < kHistCopies; ++r) gv[r] = a.ghist[((uint64_t)r * kMaxPasses + a.pass) * 256 + threadIdx.x];
  }
  for (uint32_t i = threadIdx.x; i < (uint32_t)(kSortWaves * 256); i += kSortThreads) (&wcnt[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const int shift = 8 * a.pass;
  const uint64_t tbase = (uint64_t)tile * kSortTile;
  const uint32_t tlen = (uint32_t)(a.n - tbase < (uint64_t)kSortTile ? a.n - tbase : kSortTile);
  const uint32_t wofs = wave * 64 * kSortIpt;  // this wavefront's first key in the tile
  uint32_t kk[kSortIpt], vv[kSortIpt], pos[kSortIpt];
#pragma unroll
  for (int i = 0; i < kSortIpt; ++i) {
    const uint32_t k = wofs + i * 64 + lane;
    kk[i] = k < tlen ? a.keys[tbase + k] : 0u;
    vv[i] = k < tlen ? a.vals[tbase + k] : 0u;
  }
  const uint64_t lanes_lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int i = 0; i < kSortIpt; ++i) {
    const bool valid = wofs + i * 64 + lane < tlen;
    const uint32_t dig = (kk[i] >> shift) & 255u;
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot(((dig >> b) & 1u) != 0u);
      same &= ((dig >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t before = wcnt[wave][dig];
    const uint32_t rank = (uint32_t)__popcll(same & lanes_lt);
    if (valid && rank == 0u) wcnt[wave][dig] = before + (uint32_t)__popcll(same);
    pos[i] = (dig << 16) | (before + rank);
  }
  __syncthreads();
  // threads d < 256: the wavefront offsets and the count of digit d, published
  uint32_t gcount = 0, tcount = 0;
  if (threadIdx.x < 256u) {
    const uint32_t d = threadIdx.x;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
      const uint32_t c = wcnt[w][d];
      wcnt[w][d] = tcount;
      tcount += c;
    }
    lb_publish(LbColumn{a.tile_words + d, a.group_words + d, 256}, tile, tcount);
#pragma unroll
    for (int r = 0; r < kHistCopies; ++r) gcount += gv[r];
  }
  // global base and offset within the tile of each digit: scans over the digits
  const uint32_t gbase = block_excl_scan(gcount, wsum);
  const uint32_t tb = block_excl_scan(tcount, wsum);
  if (threadIdx.x < 256u) {
    dbase[threadIdx.x] = gbase;
    lbase[threadIdx.x] = tb;
  }
  __syncthreads();
  // the tile in output order
#pragma unroll
  for (int i = 0; i < kSortIpt; ++i) {
    if (wofs + i * 64 + lane < tlen) {
      const uint32_t dig = pos[i] >> 16;
      const uint32_t slot = lbase[dig] + wcnt[wave][dig] + (pos[i] & 0xffffu);
      skey[slot] = kk[i];
      sval[slot] = vv[i];
    }
  }
  // prefix of digit d over earlier tiles: threads 4d .. 4d + 3
  {
    const uint32_t d = threadIdx.x >> 2, q = threadIdx.x & 3u;
    uint32_t part = lb_prefix_part<4>(LbColumn{a.tile_words + d, a.group_words + d, 256}, tile, q, 4, a.err);
    part += __shfl_xor(part, 1, 64);
    part += __shfl_xor(part, 2, 64);
    if (q == 0) dbase[d] += part;
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < tlen; j += kSortThreads) {
    const uint32_t key = skey[j];
    const uint32_t dig = (key >> shift) & 255u;
    const uint64_t dst = (uint64_t)dbase[dig] + (j - lbase[dig]);
    if (dst < a.n) {
      a.okeys[dst] = key;
      a.ovals[dst] = sval[j];
    } else {
      sort_fail(a.err, 2u);
    }
  }
}

struct SegArgs {
  const uint32_t* keys;   // sorted, N
  const uint32_t* vals;   // entry ids in sorted order
  uint64_t N;
  uint64_t n;             // tuples (entries [0, n) are syn0 entries)
  uint32_t K;
  uint32_t row_bits;
  const uint32_t* tuples;
  const float* dz;
  const float* stage;
  const void* xt;         // syn0 rows for the syn1 entries (float or BF16)
  float* grad;            // at the sorted position of each segment's first entry: its row gradient
  float* carry;           // per block: sum of its last run when that run continues
  float* headp;           // per block: sum of its first run when that run started earlier
  uint32_t* fin;          // per block: 1 if headp finishes a row
  const uint32_t* sort_err;  // non-zero: the sort failed, the tables are left unchanged
  // apply (W -= lr * grad), or syn0 == null for gradients only
  float lr;
  float* syn0;
  float* syn1;
  uint16_t* bf0;          // BF16 shadow rows that follow the FP32 rows, or null
  uint16_t* bf1;
};

__device__ __forceinline__ uint32_t lane_u32(uint32_t v, uint32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ float lane_f32(float v, uint32_t l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}

constexpr int kRowsInFlight = 8;

// W[row of key] -= lr * g for R rows at a time: keys and gradient
// positions (sorted entry index) come from lanes of `key` and `hp` selected
// by `mask` (one wavefront, D = dim / 64 elements per lane).
template <int DIM, int R>
__device__ void apply_rows(const SegArgs& a, uint64_t mask, uint32_t key, uint32_t hp) {
  constexpr int D = DIM / 64;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t rmask = (1u << a.row_bits) - 1u;
  while (mask) {
    uint32_t e[R];
    int m = 0;
#pragma unroll
    for (int u = 0; u < R; ++u) {
      e[u] = 64u;
      if (mask) {
        e[u] = (uint32_t)__builtin_ctzll(mask);
        mask &= mask - 1ull;
        m = u + 1;
      }
    }
    float g[R][D], w[R][D];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      if (u < m) {
        const uint32_t k = lane_u32(key, e[u]);
        float* W = ((k >> a.row_bits) ? a.syn1 : a.syn0) + (uint64_t)(k & rmask) * DIM + lane * D;
        RowIO<float, D>::load(a.grad + (uint64_t)lane_u32(hp, e[u]) * DIM + lane * D, g[u]);
        RowIO<float, D>::load(W, w[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
      if (u < m) {
        const uint32_t k = lane_u32(key, e[u]);
        const bool t1 = (k >> a.row_bits) != 0u;
        const uint64_t off = (uint64_t)(k & rmask) * DIM + lane * D;
#pragma unroll
        for (int q = 0; q < D; ++q) w[u][q] = w[u][q] - a.lr * g[u][q];
        store_row<D>((t1 ? a.syn1 : a.syn0) + off, w[u]);
        uint16_t* sh = t1 ? a.bf1 : a.bf0;
        if (sh) store_bf16<D>(sh + off, w[u]);
      }
    }
  }
}

// W[row of key k] -= lr * g, the gradient g held by the wavefront.
template <int DIM>
__device__ __forceinline__ void apply_one(const SegArgs& a, uint32_t k, const float (&g)[DIM / 64]) {
  constexpr int D = DIM / 64;
  const uint32_t lane = threadIdx.x & 63u;
  const bool t1 = (k >> a.row_bits) != 0u;
  const uint64_t off = (uint64_t)(k & ((1u << a.row_bits) - 1u)) * DIM + lane * D;
  float* W = (t1 ? a.syn1 : a.syn0) + off;
  float w[D];
  RowIO<float, D>::load(W, w);
#pragma unroll
  for (int q = 0; q < D; ++q) w[q] = w[q] - a.lr * g[q];
  store_row<D>(W, w);
  uint16_t* sh = t1 ? a.bf1 : a.bf0;
  if (sh) store_bf16<D>(sh + off, w);
}

// Wave-uniform state of one block's walk over its sorted entries.
struct SegWalk {
  uint64_t b, base;
  uint32_t nv, pkey0, after;
  uint64_t ends;          // bit e: entry e ends a run of equal keys
  uint32_t piece_start;
  bool fin_here;
};

// Sum entries [lo, hi) of the block, all reading rows of element type T at
// `rows + src_e`; kRowsInFlight row loads are issued before they are added,
// in entry order.  A run end stores the run's sum: as a finished row
// gradient (at the run's first entry), as the block's carry (the run
// continues in the next block) or as the block's head part (the run started
// in an earlier block).
template <typename T, int DIM>
__device__ __forceinline__ void seg_run(const SegArgs& a, SegWalk& w, const T* rows, uint32_t lo,
                                        uint32_t hi, uint32_t lane, uint32_t key, uint32_t src,
                                        float scale, float (&acc)[DIM / 64]) {
  constexpr int D = DIM / 64;
  for (uint32_t e0 = lo; e0 < hi; e0 += kRowsInFlight) {
    float v[kRowsInFlight][D];
#pragma unroll
    for (int u = 0; u < kRowsInFlight; ++u) {
      if (e0 + u < hi) {
        const uint64_t se = (uint64_t)lane_u32(src, e0 + u) * DIM;
        RowIO<T, D>::load(rows + se + lane * D, v[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kRowsInFlight; ++u) {
      const uint32_t e = e0 + u;
      if (e < hi) {
        const float ce = lane_f32(scale, e);
#pragma unroll
        for (int q = 0; q < D; ++q) acc[q] += ce * v[u][q];
        if ((w.ends >> e) & 1ull) {
          const uint32_t ke = lane_u32(key, e);
          const bool continued = w.piece_start == 0 && w.base > 0 && w.pkey0 == ke;
          const bool open = e + 1 == w.nv && w.after == ke;
          if (!continued && !open) {
            store_row<D>(a.grad + (w.base + w.piece_start) * DIM + lane * D, acc);
          } else if (open) {
            store_row<D>(a.carry + w.b * DIM + lane * D, acc);
          } else {
            store_row<D>(a.headp + w.b * DIM + lane * D, acc);
            w.fin_here = true;
          }
#pragma unroll
          for (int q = 0; q < D; ++q) acc[q] = 0.0f;
          w.piece_start = e + 1;
        }
      }
    }
  }
}

// Entry e of the sorted list contributes
//   syn0 entry (tuple t):        stage[t]               (sum_j dz_tj y_tj)
//   syn1 entry (t, j):           dz_tj * x[context_t]
// The table bit is the key's top bit, so a block's syn0 entries come first.
template <typename E, int DIM>
__global__ __launch_bounds__(256) void tuple_seg_kernel(SegArgs a) {
  constexpr int D = DIM / 64;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t n_blocks = (a.N + kSegBlock - 1) / kSegBlock;
  const E* xt = static_cast<const E*>(a.xt);
  const uint32_t K1 = a.K + 1, W = a.K + 2;
  for (uint64_t b = wave; b < n_blocks; b += nwaves) {
    SegWalk w;
    w.b = b;
    w.base = b * kSegBlock;
    w.nv = (uint32_t)(a.N - w.base < (uint64_t)kSegBlock ? a.N - w.base : kSegBlock);
    const uint64_t i = w.base + lane;
    const bool valid = lane < w.nv;
    const uint32_t key = valid ? a.keys[i] : 0xffffffffu;
    const uint32_t val = valid ? a.vals[i] : 0u;
    const uint32_t nkey = (i + 1 < a.N) ? a.keys[i + 1] : 0xffffffffu;
    w.pkey0 = w.base > 0 ? a.keys[w.base - 1] : 0xffffffffu;
    w.after = w.base + w.nv < a.N ? a.keys[w.base + w.nv] : 0xffffffffu;
    const bool is0 = (key >> a.row_bits) == 0u;
    uint32_t src = 0;  // row index into the stage (syn0 entries) or into syn0 (syn1 entries)
    float scale = 0.0f;
    if (valid) {
      if (is0) {
        src = val;
        scale = 1.0f;
      } else {
        const uint32_t t = val / K1;
        src = a.tuples[(uint64_t)t * W + 1];
        scale = a.dz[val];
      }
    }
    w.ends = __ballot(valid && (key != nkey || lane + 1 == w.nv));
    const uint32_t n0 = (uint32_t)__popcll(__ballot(valid && is0));
    w.piece_start = 0;
    w.fin_here = false;
    float acc[D];
#pragma unroll
    for (int q = 0; q < D; ++q) acc[q] = 0.0f;
    seg_run<float, DIM>(a, w, a.stage, 0, n0, lane, key, src, scale, acc);
    seg_run<E, DIM>(a, w, xt, n0, w.nv, lane, key, src, scale, acc);
    if (lane == 0) a.fin[b] = w.fin_here ? 1u : 0u;
  }
}

// Rows whose entries cross block edges, and (with apply) the updates.
// Workgroup g covers blocks 8 g .. 8 g + 7: wavefront w < 8 takes the first
// task below for block 8 g + w, wavefront 8 + w the second, so that the two
// chains of dependent loads run side by side.
//  * When block b's first run started in an earlier block and ends in b
//    (fin): the run's first entry is found from the first keys of the
//    blocks before b, its sum is the carries of blocks b0 .. b-1 (b0 = the
//    block of the first entry) plus b's head part, added in block order, and
//    stored at the first entry.  A chain of at most kShortChain carries the
//    wavefront sums alone, all loads in flight at once; a longer one (the
//    hot rows) is summed afterwards by the workgroup's sixteen wavefronts,
//    each a fixed contiguous part of the carries, the parts then added in
//    order.  With apply the row is then updated.
//  * With apply, the rows whose entries all lie in block b are updated from
//    their stored gradients (every read of the old rows ended with the
//    segmented-sum launch).
constexpr int kFixWaves = 16;
constexpr int kShortChain = 16;

// Block b's first key k continues a run from earlier blocks: the run's first
// entry (its sorted position) and the block holding it.  The blocks whose
// first key is k are contiguous and end at b, so the nearest block c < b
// whose first key differs is found by probing 64 blocks at a time at strides
// 1, 64, 4096, ... until a probe differs, then refining at stride / 64.
__device__ void run_start(const SegArgs& a, uint64_t b, uint32_t k, uint64_t& first, uint64_t& b0) {
  const uint32_t lane = threadIdx.x & 63u;
  int64_t hi = (int64_t)b;  // a block whose first key is k; c < hi
  int64_t stride = 1;
  int64_t c;
  for (;;) {
    const int64_t cb = hi - stride * (int64_t)(lane + 1);
    const bool differs = cb < 0 || a.keys[(uint64_t)cb * kSegBlock] != k;
    const uint64_t m = __ballot(differs);
    if (m) {
      const int64_t j = (int64_t)__builtin_ctzll(m);
      if (stride == 1) {
        c = hi - (j + 1);
        break;
      }
      hi -= stride * j;  // probe j - 1 (or hi itself) starts with k; c >= hi - stride
      stride /= 64;
    } else {
      hi -= stride * 64;
      stride *= 64;
    }
  }
  if (c < 0) {
    first = 0;
    b0 = 0;
    return;
  }
  // the run's entries in block c are a suffix of it (possibly empty)
  const uint64_t m = __ballot(a.keys[(uint64_t)c * kSegBlock + lane] == k);
  if (m) {
    first = (uint64_t)c * kSegBlock + (uint32_t)__builtin_ctzll(m);
    b0 = (uint64_t)c;
  } else {
    first = (uint64_t)(c + 1) * kSegBlock;
    b0 = (uint64_t)(c + 1);
  }
}

template <int DIM>
__global__ __launch_bounds__(64 * kFixWaves) void tuple_fix_kernel(SegArgs a) {
  constexpr int D = DIM / 64;
  __shared__ float part[kFixWaves][DIM];
  __shared__ uint32_t s_long;
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint64_t n_blocks = (a.N + kSegBlock - 1) / kSegBlock;
  if (threadIdx.x == 0) s_long = 0u;
  __syncthreads();
  constexpr uint32_t kHalf = kFixWaves / 2;
  const uint64_t b = (uint64_t)blockIdx.x * kHalf + (wv % kHalf);
  const bool completes = wv < kHalf;
  const bool upd = a.syn0 != nullptr && *a.sort_err == 0u;  // the update, unless the sort failed
  if (completes && b < n_blocks && a.fin[b]) {
    const uint32_t k = a.keys[b * kSegBlock];
    uint64_t first, b0;
    run_start(a, b, k, first, b0);
    const uint32_t cnt = (uint32_t)(b - b0);
    if (cnt > (uint32_t)kShortChain) {
      if (lane == 0) atomicOr(&s_long, 1u << wv);
    } else {
      float v[kShortChain][D];
#pragma unroll
      for (int u = 0; u < kShortChain; ++u)
        if ((uint32_t)u < cnt) RowIO<float, D>::load(a.carry + (b0 + u) * DIM + lane * D, v[u]);
      float h[D];
      RowIO<float, D>::load(a.headp + b * DIM + lane * D, h);
      float s[D];
#pragma unroll
      for (int q = 0; q < D; ++q) s[q] = 0.0f;
#pragma unroll
      for (int u = 0; u < kShortChain; ++u)
        if ((uint32_t)u < cnt) {
#pragma unroll
          for (int q = 0; q < D; ++q) s[q] += v[u][q];
        }
#pragma unroll
      for (int q = 0; q < D; ++q) s[q] += h[q];
      store_row<D>(a.grad + first * DIM + lane * D, s);
      if (upd) apply_one<DIM>(a, k, s);
    }
  }
  // rows whose entries all lie in block b
  if (!completes && upd && b < n_blocks) {
    const uint64_t base = b * kSegBlock;
    const uint32_t nv = (uint32_t)(a.N - base < (uint64_t)kSegBlock ? a.N - base : kSegBlock);
    const uint64_t i = base + lane;
    const bool valid = lane < nv;
    const uint32_t key = valid ? a.keys[i] : 0xffffffffu;
    const uint32_t nkey = i + 1 < a.N ? a.keys[i + 1] : 0xffffffffu;
    const uint32_t pkey = i > 0 ? a.keys[i - 1] : 0xffffffffu;
    const bool head = valid && (i == 0 || pkey != key);
    const uint64_t heads = __ballot(head);
    // a run that ends here (not at the block's last entry when it continues)
    const bool end_here = valid && key != nkey;
    // started in this block: a head at or before the end, after the previous end
    const uint64_t all_ends = __ballot(valid && (key != nkey || lane + 1 == nv));
    const uint64_t prev_ends = all_ends & ((1ull << lane) - 1ull);
    const uint32_t ps = prev_ends ? 64u - (uint32_t)__builtin_clzll(prev_ends) : 0u;  // run's first entry in block
    const bool started = ps > 0 || ((heads & 1ull) != 0ull);
    const uint64_t mine = __ballot(end_here && started);
    if (mine) apply_rows<DIM, 16>(a, mine, key, (uint32_t)(base + ps));
  }
  __syncthreads();
  for (uint32_t lm = s_long; lm != 0u; lm &= lm - 1u) {  // uniform over the workgroup
    const uint64_t bl = (uint64_t)blockIdx.x * kHalf + (uint32_t)__builtin_ctz(lm);
    const uint32_t k = a.keys[bl * kSegBlock];
    uint64_t first, b0;
    run_start(a, bl, k, first, b0);
    const uint64_t cnt = bl - b0;
    const uint64_t lo = b0 + cnt * wv / kFixWaves, hi = b0 + cnt * (wv + 1) / kFixWaves;
    float acc[D];
#pragma unroll
    for (int q = 0; q < D; ++q) acc[q] = 0.0f;
    for (uint64_t c0 = lo; c0 < hi; c0 += kShortChain) {
      float v[kShortChain][D];
#pragma unroll
      for (int u = 0; u < kShortChain; ++u)
        if (c0 + u < hi) RowIO<float, D>::load(a.carry + (c0 + u) * DIM + lane * D, v[u]);
#pragma unroll
      for (int u = 0; u < kShortChain; ++u)
        if (c0 + u < hi) {
#pragma unroll
          for (int q = 0; q < D; ++q) acc[q] += v[u][q];
        }
    }
#pragma unroll
    for (int q = 0; q < D; ++q) part[wv][lane * D + q] = acc[q];
    __syncthreads();
    if (wv == 0) {
      float s[D];
#pragma unroll
      for (int q = 0; q < D; ++q) s[q] = 0.0f;
      for (int kk = 0; kk < kFixWaves; ++kk) {
#pragma unroll
        for (int q = 0; q < D; ++q) s[q] += part[kk][lane * D + q];
      }
      float h[D];
      RowIO<float, D>::load(a.headp + bl * DIM + lane * D, h);
#pragma unroll
      for (int q = 0; q < D; ++q) s[q] += h[q];
      store_row<D>(a.grad + first * DIM + lane * D, s);
      if (upd) apply_one<DIM>(a, k, s);
    }
    __syncthreads();
  }
}

// Gradient rows by (table, row): binary search over the sorted keys for the
// row's first entry, where its gradient is stored.
__global__ void grad_rows_kernel(const uint32_t* keys, const float* grad, uint64_t N, uint32_t row_bits,
                                 uint32_t table, const uint32_t* rows, uint64_t n_rows, uint32_t dim,
                                 float* out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t r = wave; r < n_rows; r += nwaves) {
    const uint32_t want = (table << row_bits) | rows[r];
    uint64_t lo = 0, hi = N;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (keys[mid] < want)
        lo = mid + 1;
      else
        hi = mid;
    }
    const bool hit = lo < N && keys[lo] == want;
    for (uint32_t d = lane; d < dim; d += 64)
      out[r * dim + d] = hit ? grad[lo * dim + d] : 0.0f;
  }
}

__global__ void to_bf16_kernel(const float* src, uint64_t n, uint16_t* dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = to_bf16(src[i]);
}

int waves_grid(dg_ctx* ctx, uint64_t waves) {
  uint64_t g = (waves + 3) / 4;  // 4 wavefronts per 256-thread block
  const uint64_t cap = (uint64_t)ctx->num_cus * 32;
  if (g > cap) g = cap;
  return g == 0 ? 1 : (int)g;
}

int row_bits_for(uint32_t rows) {
  int b = 1;
  while ((1ull << b) < rows) ++b;
  return b;
}

}  // namespace

struct dg_tuple_state {
  uint32_t* d_tuples = nullptr;
  size_t tuples_cap = 0;
  float* d_logits = nullptr;
  size_t logits_cap = 0;
  double* d_wloss = nullptr;  // loss partial per forward wavefront
  size_t wloss_cap = 0;
  double* d_loss = nullptr;
  float* d_dl = nullptr;  // upstream dL/dz
  size_t dl_cap = 0;
  float* d_dz = nullptr;
  size_t dz_cap = 0;
  float* d_stage = nullptr;
  size_t stage_cap = 0;
  uint32_t* d_k[2] = {nullptr, nullptr};
  uint32_t* d_v[2] = {nullptr, nullptr};
  size_t k_cap[2] = {0, 0};
  size_t v_cap[2] = {0, 0};
  uint32_t* d_lb = nullptr;     // tickets, then the look-back words of the sort and segment launches
  size_t lb_cap = 0;
  uint32_t* h_err = nullptr;    // host-visible: set by a kernel whose look-back wait or store bound failed
  uint32_t* d_err = nullptr;    // the device's address of h_err
  float* d_grad = nullptr;
  size_t grad_cap = 0;
  float* d_carry = nullptr;
  size_t carry_cap = 0;
  float* d_headp = nullptr;
  size_t headp_cap = 0;
  uint32_t* d_fin = nullptr;
  size_t fin_cap = 0;
  // BF16 copies of the tables and the table version they were made from
  uint16_t* d_bf0 = nullptr;
  uint16_t* d_bf1 = nullptr;
  size_t bf_cap = 0;
  uint64_t bf_version = 0;
  // gradients of the last backward
  bool have_grads = false;
  uint32_t grad_rb = 0;
  uint32_t grad_dim = 0;
  uint32_t grad_rows = 0;
  uint64_t grad_N = 0;
  int sorted = 0;  // which of d_k / d_v holds the sorted pairs
};

void dg_tuples_free(dg_ctx* ctx) {
  dg_tuple_state* t = ctx->tuples;
  if (!t) return;
  void* bufs[] = {t->d_tuples, t->d_logits, t->d_wloss,    t->d_loss,       t->d_dl,    t->d_dz,
                  t->d_stage,  t->d_k[0],   t->d_k[1],     t->d_v[0],       t->d_v[1],
                  t->d_lb,
                  t->d_grad,   t->d_carry,  t->d_headp,    t->d_fin,        t->d_bf0,   t->d_bf1};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (t->h_err) hipHostFree(t->h_err);
  delete t;
  ctx->tuples = nullptr;
}

void dg_tuples_invalidate(dg_ctx* ctx) {
  if (ctx->tuples) ctx->tuples->have_grads = false;
}

namespace {

dg_tuple_state* tuple_state(dg_ctx* ctx) {
  if (!ctx->tuples) ctx->tuples = new dg_tuple_state();
  return ctx->tuples;
}

// Validate a tuple batch on the host (every id is dereferenced on the device)
// and copy it to the device.
int upload_tuples(dg_ctx* ctx, dg_tuple_state* st, const dg_table_view& tv, const uint32_t* tuples,
                  uint64_t n, uint32_t K) {
  if (K > (uint32_t)kMaxTupleNeg)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "K = %u negatives per tuple: at most %d", K, kMaxTupleNeg);
  if (tv.dim != 64 && tv.dim != 128 && tv.dim != 256)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "tuple interface: dim %u (64, 128 or 256)", tv.dim);
  const uint64_t W = K + 2;
  if (n == 0 || n * W >= (1ull << 31))
    return dg_fail(ctx, DG_ERR_ARG, "tuple count %llu: must be in [1, 2^31 / (K + 2))",
                   (unsigned long long)n);
  for (uint64_t i = 0; i < n * W; ++i)
    if (tuples[i] >= tv.rows)
      return dg_fail(ctx, DG_ERR_ARG, "tuple %llu: row %u out of range (%u rows)",
                     (unsigned long long)(i / W), tuples[i], tv.rows);
  DG_CHECK(dg_reserve(ctx, &st->d_tuples, &st->tuples_cap, (size_t)(n * W)));
  DG_HIP(ctx, hipMemcpyAsync(st->d_tuples, tuples, sizeof(uint32_t) * n * W, hipMemcpyHostToDevice,
                             ctx->stream));
  return DG_OK;
}

// BF16 copies of both tables, remade when the FP32 tables changed.
int ensure_bf16(dg_ctx* ctx, dg_tuple_state* st, const dg_table_view& tv) {
  const size_t n = (size_t)tv.rows * tv.dim;
  if (st->d_bf0 && st->bf_cap >= n && st->bf_version == tv.version) return DG_OK;
  if (st->bf_cap < n) {
    if (st->d_bf0) hipFree(st->d_bf0);
    if (st->d_bf1) hipFree(st->d_bf1);
    st->d_bf0 = st->d_bf1 = nullptr;
    st->bf_cap = 0;
    DG_HIP(ctx, hipMalloc((void**)&st->d_bf0, sizeof(uint16_t) * n));
    DG_HIP(ctx, hipMalloc((void**)&st->d_bf1, sizeof(uint16_t) * n));
    st->bf_cap = n;
  }
  const int g = waves_grid(ctx, n / 64 + 1);
  to_bf16_kernel<<<g, 256, 0, ctx->stream>>>(tv.syn0, n, st->d_bf0);
  to_bf16_kernel<<<g, 256, 0, ctx->stream>>>(tv.syn1, n, st->d_bf1);
  DG_HIP(ctx, hipGetLastError());
  st->bf_version = tv.version;
  return DG_OK;
}

// Lanes per tuple: 8 for BF16 rows with K + 2 <= 8 ids per tuple (a BF16
// row is half an FP32 row, so twice the tuples fit the same registers), 16
// for K + 2 <= 16, otherwise 32.  Returns the wavefronts launched (at most
// 128 per CU: the forward writes one loss partial per wavefront).
template <typename E, int DIM, bool BWD, bool OUT>
uint32_t launch_fwd(dg_ctx* ctx, const TupleArgs& a) {
  if (sizeof(E) == 2 && DIM >= 128 && a.K + 2 <= 8) {
    const int g = waves_grid(ctx, (a.n + 7) / 8);  // a tuple per 8 lanes
    tuple_fwd_kernel<E, DIM, 8, BWD, 8, OUT><<<g, 256, 0, ctx->stream>>>(a);
    return 4u * (uint32_t)g;
  }
  if (a.K + 2 <= 16) {
    const int g = waves_grid(ctx, (a.n + 3) / 4);  // a tuple per 16 lanes
    if (a.K + 1 <= 8)
      tuple_fwd_kernel<E, DIM, 8, BWD, 16, OUT><<<g, 256, 0, ctx->stream>>>(a);
    else
      tuple_fwd_kernel<E, DIM, 16, BWD, 16, OUT><<<g, 256, 0, ctx->stream>>>(a);
    return 4u * (uint32_t)g;
  }
  const int g = waves_grid(ctx, (a.n + 1) / 2);  // a tuple per half-wavefront
  if (a.K + 1 <= 8)
    tuple_fwd_kernel<E, DIM, 8, BWD, 32, OUT><<<g, 256, 0, ctx->stream>>>(a);
  else
    tuple_fwd_kernel<E, DIM, 16, BWD, 32, OUT><<<g, 256, 0, ctx->stream>>>(a);
  return 4u * (uint32_t)g;
}

template <bool BWD, bool OUT = !BWD>
uint32_t launch_fwd_form(dg_ctx* ctx, const TupleArgs& a, int bf16, uint32_t dim) {
  if (bf16) {
    if (dim == 64) return launch_fwd<uint16_t, 64, BWD, OUT>(ctx, a);
    if (dim == 128) return launch_fwd<uint16_t, 128, BWD, OUT>(ctx, a);
    return launch_fwd<uint16_t, 256, BWD, OUT>(ctx, a);
  }
  if (dim == 64) return launch_fwd<float, 64, BWD, OUT>(ctx, a);
  if (dim == 128) return launch_fwd<float, 128, BWD, OUT>(ctx, a);
  return launch_fwd<float, 256, BWD, OUT>(ctx, a);
}

template <typename E>
void launch_seg(dg_ctx* ctx, const SegArgs& a, uint32_t dim, int g, int gf) {
  if (dim == 64) {
    tuple_seg_kernel<E, 64><<<g, 256, 0, ctx->stream>>>(a);
    tuple_fix_kernel<64><<<gf, 64 * kFixWaves, 0, ctx->stream>>>(a);
  } else if (dim == 128) {
    tuple_seg_kernel<E, 128><<<g, 256, 0, ctx->stream>>>(a);
    tuple_fix_kernel<128><<<gf, 64 * kFixWaves, 0, ctx->stream>>>(a);
  } else {
    tuple_seg_kernel<E, 256><<<g, 256, 0, ctx->stream>>>(a);
    tuple_fix_kernel<256><<<gf, 64 * kFixWaves, 0, ctx->stream>>>(a);
  }
}

}  // namespace

extern "C" int dg_sgns_tuples_forward(dg_ctx* ctx, const uint32_t* tuples, uint64_t n, uint32_t K,
                                      int bf16, float* logits, double* loss) {
  if (!ctx || !tuples) return DG_ERR_ARG;
  dg_table_view tv;
  DG_CHECK(dg_sgns_view(ctx, &tv));
  DG_HIP(ctx, hipSetDevice(ctx->device));
  dg_tuple_state* st = tuple_state(ctx);
  DG_CHECK(upload_tuples(ctx, st, tv, tuples, n, K));
  if (bf16) DG_CHECK(ensure_bf16(ctx, st, tv));
  const uint32_t K1 = K + 1;
  DG_CHECK(dg_reserve(ctx, &st->d_logits, &st->logits_cap, (size_t)(n * K1)));
  DG_CHECK(dg_reserve(ctx, &st->d_wloss, &st->wloss_cap, (size_t)ctx->num_cus * 128));
  if (!st->d_loss) DG_HIP(ctx, hipMalloc((void**)&st->d_loss, sizeof(double)));
  TupleArgs a{};
  a.tuples = st->d_tuples;
  a.n = n;
  a.K = K;
  a.xt = bf16 ? (const void*)st->d_bf0 : (const void*)tv.syn0;
  a.yt = bf16 ? (const void*)st->d_bf1 : (const void*)tv.syn1;
  a.logits = st->d_logits;
  a.wloss = st->d_wloss;
  size_t tk;
  DG_CHECK(dg_time_begin(ctx, "tuple_fwd", &tk));
  const uint32_t waves = launch_fwd_form<false>(ctx, a, bf16, tv.dim);
  loss_sum_kernel<<<1, kLossThreads, 0, ctx->stream>>>(st->d_wloss, waves, st->d_loss);
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tk));
  if (logits)
    DG_HIP(ctx, hipMemcpyAsync(logits, st->d_logits, sizeof(float) * n * K1, hipMemcpyDeviceToHost,
                               ctx->stream));
  double h_loss = 0.0;
  DG_HIP(ctx, hipMemcpyAsync(&h_loss, st->d_loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_CHECK(dg_time_resolve(ctx));
  if (loss) *loss = h_loss;
  return DG_OK;
}

namespace {

// ---- the allocation-free forward ---------------------------------------------
// Scratch of dg_sgns_tuples_forward_noalloc: the loss partial of each
// forward wavefront (at most kNoAllocMaxWaves doubles), then a range-check
// flag.  Each device holds its own copy; host code reaches it only through
// the address hipGetSymbolAddress gives on that device
// (dg_tuples_scratch_setup).
constexpr uint64_t kNoAllocMaxTuples = 1ull << 20;
constexpr uint32_t kNoAllocMaxWaves = 65536;              // 128 per CU on up to 512 CUs
constexpr int kNoAllocFlag = 2 * kNoAllocMaxWaves;        // float index of the flag
static __device__ __align__(256) float scratch[kNoAllocFlag + 64];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxScratchDevices] = {};

__global__ void noalloc_clear_kernel(uint32_t* flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0u;
}

__global__ void tuple_check_kernel(const uint32_t* tuples, uint64_t count, uint32_t rows, uint32_t* flag) {
  bool bad = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    bad |= tuples[i] >= rows;
  if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

}  // namespace

int dg_tuples_scratch_setup(dg_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxScratchDevices)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "device %d: the allocation-free scratch covers devices "
                   "0..%d", ctx->device, kMaxScratchDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return dg_fail(ctx, DG_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[ctx->device] = static_cast<float*>(p);
  }
  ctx->noalloc_scratch = g_scratch_dev[ctx->device];
  return DG_OK;
}

extern "C" int dg_sgns_tuples_forward_noalloc(dg_ctx* ctx, const uint32_t* d_tuples, uint64_t n,
                                              uint32_t K, float* d_logits, double* d_loss) {
  if (!ctx || !d_tuples || !d_logits || !d_loss) return DG_ERR_ARG;
  if (!ctx->noalloc_scratch) return dg_fail(ctx, DG_ERR_STATE, "no allocation-free scratch on this context");
  dg_table_view tv;
  DG_CHECK(dg_sgns_view(ctx, &tv));
  if (K > (uint32_t)kMaxTupleNeg)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "K = %u negatives per tuple: at most %d", K, kMaxTupleNeg);
  if (tv.dim != 64 && tv.dim != 128 && tv.dim != 256)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "tuple interface: dim %u (64, 128 or 256)", tv.dim);
  if (n == 0 || n > kNoAllocMaxTuples)
    return dg_fail(ctx, DG_ERR_ARG, "%llu tuples: the allocation-free forward takes 1 to %llu",
                   (unsigned long long)n, (unsigned long long)kNoAllocMaxTuples);
  if ((uint64_t)ctx->num_cus * 128 > kNoAllocMaxWaves)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "%d CUs: the allocation-free forward covers at most %u",
                   ctx->num_cus, kNoAllocMaxWaves / 128);
  DG_HIP(ctx, hipSetDevice(ctx->device));
  float* base = ctx->noalloc_scratch;
  double* wloss = reinterpret_cast<double*>(base);
  uint32_t* flag = reinterpret_cast<uint32_t*>(base + kNoAllocFlag);
  // every row id is dereferenced by the forward: check them on the device
  const uint64_t count = n * (K + 2);
  noalloc_clear_kernel<<<1, 64, 0, ctx->stream>>>(flag);
  tuple_check_kernel<<<waves_grid(ctx, count / 64 + 1), 256, 0, ctx->stream>>>(d_tuples, count, tv.rows,
                                                                               flag);
  DG_HIP(ctx, hipGetLastError());
  uint32_t bad = 0;
  DG_HIP(ctx, hipMemcpyAsync(&bad, flag, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (bad) return dg_fail(ctx, DG_ERR_ARG, "a tuple row is out of range (%u rows)", tv.rows);
  TupleArgs a{};
  a.tuples = d_tuples;
  a.n = n;
  a.K = K;
  a.xt = tv.syn0;
  a.yt = tv.syn1;
  a.logits = d_logits;
  a.wloss = wloss;
  const uint32_t waves = launch_fwd_form<false>(ctx, a, 0, tv.dim);
  loss_sum_kernel<<<1, kLossThreads, 0, ctx->stream>>>(wloss, waves, d_loss);
  DG_HIP(ctx, hipGetLastError());
  return DG_OK;
}

namespace {

// The backward (and, fused, the forward's logits and loss from the same pass
// over the rows, with the loss's own gradient).
int tuples_backward_run(dg_ctx* ctx, const uint32_t* tuples, uint64_t n, uint32_t K, int bf16,
                        const float* dlogits, float lr, int apply, bool fused, float* logits,
                        double* loss) {
  if (!ctx || !tuples) return DG_ERR_ARG;
  dg_table_view tv;
  DG_CHECK(dg_sgns_view(ctx, &tv));
  DG_HIP(ctx, hipSetDevice(ctx->device));
  dg_tuple_state* st = tuple_state(ctx);
  st->have_grads = false;
  DG_CHECK(upload_tuples(ctx, st, tv, tuples, n, K));
  if (bf16) DG_CHECK(ensure_bf16(ctx, st, tv));
  const uint32_t K1 = K + 1, dim = tv.dim;
  const uint64_t N = n * (K + 2);
  const int rb = row_bits_for(tv.rows);
  if (N > (uint64_t)kCountMask)  // the sort's look-back words count in 30 bits
    return dg_fail(ctx, DG_ERR_ARG, "%llu gradient entries: at most 2^30 - 1 per backward",
                   (unsigned long long)N);
  if ((rb + 1 + 7) / 8 > kMaxPasses)
    return dg_fail(ctx, DG_ERR_UNSUPPORTED, "%u rows: keys wider than 32 bits", tv.rows);
  DG_CHECK(dg_reserve(ctx, &st->d_dz, &st->dz_cap, (size_t)(n * K1)));
  DG_CHECK(dg_reserve(ctx, &st->d_stage, &st->stage_cap, (size_t)(n * dim)));
  for (int b = 0; b < 2; ++b) {
    DG_CHECK(dg_reserve(ctx, &st->d_k[b], &st->k_cap[b], (size_t)N));
    DG_CHECK(dg_reserve(ctx, &st->d_v[b], &st->v_cap[b], (size_t)N));
  }
  if (dlogits) {
    DG_CHECK(dg_reserve(ctx, &st->d_dl, &st->dl_cap, (size_t)(n * K1)));
    DG_HIP(ctx, hipMemcpyAsync(st->d_dl, dlogits, sizeof(float) * n * K1, hipMemcpyHostToDevice,
                               ctx->stream));
  }
  if (fused) {
    DG_CHECK(dg_reserve(ctx, &st->d_logits, &st->logits_cap, (size_t)(n * K1)));
    DG_CHECK(dg_reserve(ctx, &st->d_wloss, &st->wloss_cap, (size_t)ctx->num_cus * 128));
    if (!st->d_loss) DG_HIP(ctx, hipMalloc((void**)&st->d_loss, sizeof(double)));
  }
  // the sort's sizes and the words it expects zero (cleared by the backward kernel)
  if (!st->h_err) {
    DG_HIP(ctx, hipHostMalloc((void**)&st->h_err, sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent));
    DG_HIP(ctx, hipHostGetDevicePointer((void**)&st->d_err, st->h_err, 0));
  }
  *st->h_err = 0u;  // the previous call's kernels have finished
  const int passes = (rb + 1 + 7) / 8;
  const uint64_t n_tiles = (N + kSortTile - 1) / kSortTile;
  const uint64_t n_groups = (n_tiles + kGroup - 1) / kGroup;
  const uint32_t n_hist = (uint32_t)((N + kHistChunk - 1) / kHistChunk);  // histogram workgroups
  const uint64_t n_blocks = (N + kSegBlock - 1) / kSegBlock;
  // words cleared by the backward kernel, in 32-bit units (64-bit group
  // words at even offsets): the digit counts [copy][pass][digit], kMaxPasses
  // tickets, the device error word, then per pass the [tile][digit] and
  // [group][digit] look-back words.  Tiles are numbered by workgroup id when every tile is resident
  // at once (one 1,024-thread workgroup per CU), by ticket otherwise.
  auto even = [](uint64_t x) { return (x + 1) & ~1ull; };
  const uint64_t o_tick = (uint64_t)kHistCopies * kMaxPasses * 256;
  const uint64_t o_err = o_tick + kMaxPasses;  // the device error word
  const uint64_t o_sort = even(o_err + 1);
  const uint64_t pass_words = even(n_tiles * 256) + 2 * n_groups * 256;
  const uint64_t n_lb = o_sort + (uint64_t)passes * pass_words;
  const bool tickets = n_tiles > (uint64_t)ctx->num_cus;
  DG_CHECK(dg_reserve(ctx, &st->d_lb, &st->lb_cap, (size_t)n_lb));
  TupleArgs a{};
  a.tuples = st->d_tuples;
  a.n = n;
  a.K = K;
  a.xt = bf16 ? (const void*)st->d_bf0 : (const void*)tv.syn0;
  a.yt = bf16 ? (const void*)st->d_bf1 : (const void*)tv.syn1;
  a.dlogits = dlogits ? st->d_dl : nullptr;
  a.dz = st->d_dz;
  a.stage = st->d_stage;
  a.keys = st->d_k[0];
  a.vals = st->d_v[0];
  a.row_bits = (uint32_t)rb;
  a.clear = st->d_lb;
  a.n_clear = n_lb;
  if (fused) {
    a.logits = st->d_logits;
    a.wloss = st->d_wloss;
  }
  size_t tk;
  DG_CHECK(dg_time_begin(ctx, "tuple_prep", &tk));
  const uint32_t waves = fused ? launch_fwd_form<true, true>(ctx, a, bf16, dim)
                               : launch_fwd_form<true, false>(ctx, a, bf16, dim);
  if (fused) loss_sum_kernel<<<1, kLossThreads, 0, ctx->stream>>>(st->d_wloss, waves, st->d_loss);
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tk));
  // sort (key, entry) pairs on rb + 1 key bits
  DG_CHECK(dg_time_begin(ctx, "tuple_sort", &tk));
  radix_hist_kernel<<<n_hist, 256, 0, ctx->stream>>>(st->d_k[0], N, passes, st->d_lb);
  int cur = 0;
  for (int p = 0; p < passes; ++p) {
    uint32_t* pw = st->d_lb + o_sort + (uint64_t)p * pass_words;
    SortPass sp{};
    sp.keys = st->d_k[cur];
    sp.vals = st->d_v[cur];
    sp.n = N;
    sp.pass = p;
    sp.ghist = st->d_lb;
    sp.tile_words = pw;
    sp.group_words = reinterpret_cast<uint64_t*>(pw + even(n_tiles * 256));
    sp.ticket = tickets ? st->d_lb + o_tick + p : nullptr;
    sp.err = SortErr{st->d_err, st->d_lb + o_err};
    sp.okeys = st->d_k[1 - cur];
    sp.ovals = st->d_v[1 - cur];
    radix_pass_kernel<<<(unsigned)n_tiles, kSortThreads, 0, ctx->stream>>>(sp);
    cur = 1 - cur;
  }
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tk));
  // segmented sums, then the rows that cross blocks (and, with apply, the updates)
  DG_CHECK(dg_reserve(ctx, &st->d_grad, &st->grad_cap, (size_t)(N * dim)));
  DG_CHECK(dg_reserve(ctx, &st->d_carry, &st->carry_cap, (size_t)(n_blocks * dim)));
  DG_CHECK(dg_reserve(ctx, &st->d_headp, &st->headp_cap, (size_t)(n_blocks * dim)));
  DG_CHECK(dg_reserve(ctx, &st->d_fin, &st->fin_cap, (size_t)n_blocks));
  SegArgs sa{};
  sa.keys = st->d_k[cur];
  sa.vals = st->d_v[cur];
  sa.N = N;
  sa.n = n;
  sa.K = K;
  sa.row_bits = (uint32_t)rb;
  sa.tuples = st->d_tuples;
  sa.dz = st->d_dz;
  sa.stage = st->d_stage;
  sa.xt = bf16 ? (const void*)st->d_bf0 : (const void*)tv.syn0;
  sa.grad = st->d_grad;
  sa.carry = st->d_carry;
  sa.headp = st->d_headp;
  sa.fin = st->d_fin;
  sa.sort_err = st->d_lb + o_err;
  const bool shadow = st->d_bf0 && st->bf_version == tv.version;
  if (apply) {
    sa.lr = lr;
    sa.syn0 = tv.syn0;
    sa.syn1 = tv.syn1;
    sa.bf0 = shadow ? st->d_bf0 : nullptr;
    sa.bf1 = shadow ? st->d_bf1 : nullptr;
  }
  const int g = waves_grid(ctx, n_blocks);
  const int gf = (int)((n_blocks + kFixWaves / 2 - 1) / (kFixWaves / 2));
  DG_CHECK(dg_time_begin(ctx, "tuple_seg", &tk));
  if (bf16)
    launch_seg<uint16_t>(ctx, sa, dim, g, gf);
  else
    launch_seg<float>(ctx, sa, dim, g, gf);
  DG_HIP(ctx, hipGetLastError());
  DG_CHECK(dg_time_end(ctx, tk));
  st->sorted = cur;
  st->grad_rb = (uint32_t)rb;
  st->grad_dim = dim;
  st->grad_rows = tv.rows;
  st->grad_N = N;
  if (apply) {
    const uint64_t v = dg_sgns_written(ctx);
    if (shadow) st->bf_version = v;  // the shadow rows were updated with the FP32 rows
  }
  double h_loss = 0.0;
  if (fused) {
    if (logits)
      DG_HIP(ctx, hipMemcpyAsync(logits, st->d_logits, sizeof(float) * n * K1, hipMemcpyDeviceToHost,
                                 ctx->stream));
    DG_HIP(ctx, hipMemcpyAsync(&h_loss, st->d_loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_CHECK(dg_time_resolve(ctx));
  if (*st->h_err)
    return dg_fail(ctx, DG_ERR_HIP,
                   "tuple sort: a look-back wait or an output bound failed (%u); tables left unchanged",
                   *st->h_err);
  st->have_grads = true;
  if (fused && loss) *loss = h_loss;
  return DG_OK;
}

}  // namespace

extern "C" int dg_sgns_tuples_backward(dg_ctx* ctx, const uint32_t* tuples, uint64_t n, uint32_t K,
                                       int bf16, const float* dlogits, float lr, int apply) {
  return tuples_backward_run(ctx, tuples, n, K, bf16, dlogits, lr, apply, false, nullptr, nullptr);
}

extern "C" int dg_sgns_tuples_step(dg_ctx* ctx, const uint32_t* tuples, uint64_t n, uint32_t K, int bf16,
                                   float lr, int apply, float* logits, double* loss) {
  return tuples_backward_run(ctx, tuples, n, K, bf16, nullptr, lr, apply, true, logits, loss);
}

extern "C" int dg_sgns_grad_rows(dg_ctx* ctx, int table, const uint32_t* rows, uint64_t n_rows,
                                 float* out) {
  if (!ctx || (n_rows && (!rows || !out)) || (table != 0 && table != 1)) return DG_ERR_ARG;
  dg_tuple_state* st = ctx->tuples;
  if (!st || !st->have_grads) return dg_fail(ctx, DG_ERR_STATE, "no gradients: run dg_sgns_tuples_backward");
  if (n_rows == 0) return DG_OK;
  for (uint64_t i = 0; i < n_rows; ++i)
    if (rows[i] >= st->grad_rows)
      return dg_fail(ctx, DG_ERR_ARG, "row %u out of range (%u rows)", rows[i], st->grad_rows);
  DG_HIP(ctx, hipSetDevice(ctx->device));
  uint32_t* d_rows = nullptr;
  float* d_out = nullptr;
  DG_HIP(ctx, hipMalloc((void**)&d_rows, sizeof(uint32_t) * n_rows));
  DG_HIP(ctx, hipMalloc((void**)&d_out, sizeof(float) * n_rows * st->grad_dim));
  int rc = DG_OK;
  hipError_t e = hipMemcpy(d_rows, rows, sizeof(uint32_t) * n_rows, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    grad_rows_kernel<<<waves_grid(ctx, n_rows), 256, 0, ctx->stream>>>(
        st->d_k[st->sorted], st->d_grad, st->grad_N, st->grad_rb, (uint32_t)table, d_rows, n_rows,
        st->grad_dim, d_out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpy(out, d_out, sizeof(float) * n_rows * st->grad_dim, hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = dg_fail(ctx, DG_ERR_HIP, "grad_rows: %s", hipGetErrorString(e));
  hipFree(d_rows);
  hipFree(d_out);
  return rc;
}
