// walker.hip - persistent information-oriented walker for gfx950, and the
// per-round corpus packing and relative-entropy kernels.
//
// Reference: WalkEngine::internal_random_walk / internal_walk_epoch
// (include/walk.hpp:608-782, :789-1043) with Correlation (:260-387), for one
// partition.  PathCollector/PathSet (include/path.hpp) become a device token
// buffer in walk-id order.
//
// Design (MI355X-first, not a translation of the OpenMP/MPI loop):
//   * One lane per walk, 64 walks per wavefront, in one persistent launch.
//     Walks have uneven lengths; a lane whose walk ends takes the next walk id
//     from its wave's pool, and a wave refills its pool with one 64-bit atomic
//     on the device work queue, so every lane stays busy until the queue
//     drains and every wave then leaves the loop.
//   * Each lane keeps its walk in LDS (kPathLds vertices, lane-major with an odd
//     stride so lane-strided and cooperative accesses are bank-conflict free)
//     and spills longer walks to a private global area.  The visit count f_i
//     of walk.hpp:816 (an unordered_map per walk) is a scan of that buffer.
//   * A finished walk is written out cooperatively by its wave into its own
//     slot (walk index * kPathLds tokens) with one 64-lane store; only walks
//     longer than a slot take space from a shared overflow cursor.
//   * Visit counts are not taken in the walk loop: the packing pass counts
//     every token, the highest-degree vertices in per-workgroup LDS counters
//     and the rest with uncontended global atomics.
//   * Every draw comes from the walk's own Philox stream (dg_philox.h), and the
//     FP64 entropy/acceptance math uses dg_math.h, so the corpus is the same
//     bits as the CPU oracle's whatever the schedule.
#include <algorithm>

#include "dg_internal.h"
#include "dg_math.h"
#include "dg_philox.h"

namespace {

constexpr int kWalkBlock = 256;
constexpr int kPathLds = 23;    // vertices per lane kept in LDS (odd); reference walks hold 22
constexpr int kTermTable = 32;  // entropy terms tabulated for walks shorter than 32
constexpr int kTermStride = kTermTable + 1;
constexpr uint32_t kPool = 64;
constexpr uint32_t kDefaultLengthLimit = DG_WALK_LENGTH_LIMIT_DEFAULT;

struct WalkArgs {
  const uint64_t* row_ptr;
  const uint2* adj;  // {neighbour, bounce threshold}
  const double* term_table;
  uint32_t v_num;
  uint32_t round;
  uint64_t seed;
  uint32_t walker_begin;
  uint32_t walker_num;
  uint32_t min_length;
  uint32_t max_length;
  uint32_t length_limit;
  uint32_t spill_per_slot;
  double mu;
  uint32_t* tokens;   // raw buffer: walker_num slots of kPathLds, then the overflow region
  uint64_t tok_cap;
  uint64_t slots_end;  // walker_num * kPathLds
  uint64_t* walk_off;
  uint32_t* walk_len;
  uint32_t* spill;
  dg_walk_queue* q;
};

// Running statistics of Correlation (include/walk.hpp:260-387).  Sn and N
// both equal the number of vertices on the walk, which the kernel keeps as an
// integer, so only the seven FP64 moments live in registers.
struct WalkState {
  double H, EH, ES, EHS, EHH, ESS;
};

// log2 term of Correlation::iterateHn with step_k = 1 (include/walk.hpp:
// 284-304) for a walk of Sn vertices appending a vertex seen fi times before:
// log2(pow(Sn/(Sn+1), Sn) * pow(1/(Sn+1), 1) [* pow((fi+1)/fi, fi) * pow(fi+1, 1)]).
__device__ __forceinline__ double entropy_term(double Sn, double fi) {
  const double one = 1.0;
  const double Sn1 = Sn + one;
  const double prod = dg_pow(Sn / Sn1, Sn) * dg_pow(one / Sn1, one);
  if (fi == 0.0) return dg_log2(prod);
  const double fj = fi + one;
  const double fac = dg_pow(fj / fi, fi) * dg_pow(fj, one);
  return dg_log2(prod * fac);
}

// Rest of Correlation::iterate (include/walk.hpp:317-354), same operation
// order.  The correlation R itself is only read by the termination test
// (walk.hpp:979-985), which runs after min_length steps, so it is formed from
// the moments there (state_corr) instead of on every step.
__device__ __forceinline__ void state_iterate(WalkState& w, double Sn, double t) {
  const double one = 1.0;
  const double H1 = (Sn * w.H - t) / (Sn + one);
  const double S1 = Sn + one;
  const double N1 = Sn + 1.0;
  w.EH = w.EH + (H1 - w.EH) / N1;
  w.ES = w.ES + (S1 - w.ES) / N1;
  w.EHS = w.EHS + (H1 * S1 - w.EHS) / N1;
  w.EHH = w.EHH + (H1 * H1 - w.EHH) / N1;
  w.ESS = w.ESS + (S1 * S1 - w.ESS) / N1;
  w.H = H1;
}

// Pearson R of the (H, S) history from the running moments (walk.hpp:342-354).
__device__ __forceinline__ double state_corr(const WalkState& w) {
  const double DH = w.EHH - w.EH * w.EH;
  const double DS = w.ESS - w.ES * w.ES;
  const double cov = w.EHS - w.EH * w.ES;
  return cov == 0.0 ? 0.0 : cov / dg_sqrt(DH * DS);
}

// term_table[S * kTermStride + f] = entropy_term(S, f) for 1 <= S < kTermTable,
// 0 <= f <= S; computed by the same device code the walker would run inline.
__global__ void term_table_kernel(double* tbl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kTermTable * kTermStride) return;
  const int S = i / kTermStride, f = i % kTermStride;
  tbl[i] = (S >= 1 && f <= S) ? entropy_term((double)S, (double)f) : 0.0;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__global__ __launch_bounds__(kWalkBlock) void walk_kernel(WalkArgs a) {
  __shared__ uint32_t s_path[kWalkBlock * kPathLds];
  __shared__ double s_term[kTermTable * kTermStride];
  for (int i = threadIdx.x; i < kTermTable * kTermStride; i += kWalkBlock) s_term[i] = a.term_table[i];
  __syncthreads();

  const uint32_t lane = threadIdx.x & 63u;
  uint32_t* path = s_path + threadIdx.x * kPathLds;
  const uint64_t slot = (uint64_t)blockIdx.x * kWalkBlock + threadIdx.x;
  uint32_t* spill = a.spill + slot * a.spill_per_slot;

  bool active = false;
  bool drained = false;                   // wave-uniform
  uint64_t pool_next = 0, pool_end = 0;   // wave-uniform

  uint32_t wid = 0, cur = 0, cur_deg = 0, len = 0;
  uint64_t cur_rb = 0;
  uint32_t ndraw = 0, nmove = 0;
  // this lane's totals over its walks, added to the queue once per wave at the end
  uint64_t tot_draw = 0, tot_move = 0, tot_tok = 0;
  uint32_t max_l = 0;
  bool pending_back = false;
  WalkState ws = {0, 0, 0, 0, 0, 0};
  dg_stream rs;
  dg_stream_init(&rs, a.seed, 0, 0, DG_TAG_WALK);

  while (true) {
    // ---- hand new walk ids to idle lanes --------------------------------
    if (!drained) {
      const uint64_t want = __ballot(!active);
      if (want != 0) {
        const uint32_t nwant = (uint32_t)__popcll(want);
        const uint32_t rank = lanes_below(want);
        const uint64_t avail = pool_end - pool_next;
        uint64_t id;
        if (avail >= nwant) {
          id = pool_next + rank;
          pool_next += nwant;
        } else {
          unsigned long long base = 0;
          if (lane == 0) base = atomicAdd(&a.q->next, (unsigned long long)kPool);
          base = __shfl(base, 0, 64);
          id = rank < avail ? pool_next + rank : base + (rank - avail);
          pool_next = base + (nwant - avail);
          pool_end = base + kPool;
        }
        if (!active && id < a.walker_num) {
          // walk.hpp:529-589 (start = id mod V), :661-669 (footprint, step 0),
          // :849-856 (first-visit state)
          wid = (uint32_t)id;
          const uint32_t w = a.walker_begin + wid;
          cur = w % a.v_num;
          cur_rb = a.row_ptr[cur];
          cur_deg = (uint32_t)(a.row_ptr[cur + 1] - cur_rb);
          path[0] = cur;
          len = 1;
          ndraw = 0;
          nmove = 1;
          pending_back = false;
          ws.H = 0.0;
          ws.EH = 0.0;
          ws.ES = 1.0;
          ws.EHS = 0.0;
          ws.EHH = 0.0;
          ws.ESS = 1.0;
          dg_stream_init(&rs, a.seed, w, a.round, DG_TAG_WALK);
          active = true;
        }
        if (__ballot(!active) != 0) drained = true;
      }
    }
    if (__ballot(active) == 0) {
      if (drained) break;
      continue;
    }

    // ---- one appended vertex per active lane ------------------------------
    bool finish = false;
    if (active) {
      if (a.max_length != 0 && len >= a.max_length) {
        finish = true;
      } else {
        bool append = false;
        bool trace = false;
        uint32_t v_app = 0;
        uint64_t next_rb = 0, next_re = 0;  // row of the vertex moved to, loaded early
        if (pending_back) {
          v_app = cur;  // walk.hpp:986-1017: back to the current vertex
          append = true;
        } else if (cur_deg == 0) {
          finish = true;  // walk.hpp:887-892
        } else {
          const uint32_t i = dg_stream_uniform(&rs, cur_deg);  // walk.hpp:899
          const uint2 e = a.adj[cur_rb + i];
          ++ndraw;
          if (cur_deg == 2 && a.adj[cur_rb].x == cur) {
            finish = true;  // walk.hpp:900-902
          } else if (!(len == 1 && e.x == cur)) {  // walk.hpp:904-906 redraw at step 0
            const uint32_t u = dg_stream_u31(&rs);  // walk.hpp:928
            trace = (u >= e.y);                    // walk.hpp:933-938 via the threshold
            v_app = e.x;
            append = true;
            if (!trace) {  // the walk moves to e.x unless it ends: its row, in flight during the math
              next_rb = a.row_ptr[e.x];
              next_re = a.row_ptr[e.x + 1];
            }
          }
        }
        if (append) {
          // walk.hpp:961-985: footprint, f_i = visits so far, entropy update
          uint32_t fi = 0;
          const uint32_t n1 = len < (uint32_t)kPathLds ? len : (uint32_t)kPathLds;
          for (uint32_t j = 0; j < n1; ++j) fi += (path[j] == v_app);
          for (uint32_t j = kPathLds; j < len; ++j) fi += (spill[j - kPathLds] == v_app);
          const double t = len < (uint32_t)kTermTable ? s_term[len * kTermStride + fi]
                                                       : entropy_term((double)len, (double)fi);
          state_iterate(ws, (double)len, t);
          if (len < (uint32_t)kPathLds)
            path[len] = v_app;
          else
            spill[len - kPathLds] = v_app;
          ++len;
          const uint32_t step = len - 1;
          bool ends = false;
          if (step > a.min_length) {
            const double R = state_corr(ws);
            ends = R * R < a.mu || R < 0.0;
          }
          if (ends) {
            finish = true;
          } else if (a.max_length != 0 && len >= a.max_length) {
            finish = true;
          } else if (len >= a.length_limit) {
            finish = true;
            atomicOr(&a.q->flags, DG_QFLAG_LENGTH_LIMIT);
          } else if (pending_back) {
            pending_back = false;
          } else if (trace) {
            pending_back = true;
          } else {
            cur = v_app;  // walk.hpp:1018-1023
            cur_rb = next_rb;
            cur_deg = (uint32_t)(next_re - next_rb);
            ++nmove;
          }
        }
      }
    }

    // ---- write finished walks out, cooperatively per wave ------------------
    // A walk of up to kPathLds vertices goes to its own slot of the raw buffer
    // (walk index * kPathLds); longer walks take space after the slots from an
    // overflow cursor.  No wave waits on a shared counter for a short walk.
    const uint64_t fin = __ballot(finish);
    if (fin != 0) {
      uint64_t my_off = 0;
      bool fits = true;
      if (finish) {
        tot_draw += ndraw;
        tot_move += nmove;
        tot_tok += len;
        max_l = len > max_l ? len : max_l;
        if (len <= (uint32_t)kPathLds) {
          my_off = (uint64_t)wid * kPathLds;
        } else {
          my_off = a.slots_end + atomicAdd(&a.q->cursor, (unsigned long long)len);
          fits = my_off + len <= a.tok_cap;
          if (!fits) atomicOr(&a.q->flags, DG_QFLAG_TOKEN_OVERFLOW);
        }
        a.walk_off[wid] = my_off;
        a.walk_len[wid] = len;
      }
      __builtin_amdgcn_wave_barrier();
      uint64_t mm = fin & __ballot(fits);
      const uint32_t* wave_path = s_path + (threadIdx.x - lane) * kPathLds;
      while (mm != 0) {
        const int L = __builtin_ctzll(mm);
        mm &= mm - 1;
        const uint32_t n = __shfl(len, L, 64);
        const unsigned long long off = __shfl((unsigned long long)my_off, L, 64);
        const uint32_t nl = n < (uint32_t)kPathLds ? n : (uint32_t)kPathLds;
        const uint32_t* src = wave_path + L * kPathLds;
        // lane j < nl holds token j (nl <= kPathLds < 64)
        const uint32_t v = (uint32_t)lane < nl ? src[lane] : 0xffffffffu;
        if ((uint32_t)lane < nl) a.tokens[off + lane] = v;
      }
      if (finish && fits)  // the spilled tail, by its own lane
        for (uint32_t j = kPathLds; j < len; ++j) a.tokens[my_off + j] = spill[j - kPathLds];
      if (finish) active = false;
    }
  }
  // ---- the lane totals, one atomic each per wave
  const uint64_t wd = wave_sum_u64(tot_draw), wm = wave_sum_u64(tot_move), wt = wave_sum_u64(tot_tok);
  uint32_t wmax = max_l;
  for (int d = 32; d >= 1; d >>= 1) {
    const uint32_t o = __shfl_xor(wmax, d, 64);
    wmax = o > wmax ? o : wmax;
  }
  if (lane == 0) {
    if (wd) atomicAdd(&a.q->draws, (unsigned long long)wd);
    if (wm) atomicAdd(&a.q->moves, (unsigned long long)wm);
    if (wt) atomicAdd(&a.q->tokens, (unsigned long long)wt);
    if (wmax) atomicMax(&a.q->max_len, wmax);
  }
}

// Empties a round that overflowed the token buffer or hit the length limit
// (every walk length 0), so the kernels after it leave the corpus, the visit
// counts and the tables as they were.
__global__ void round_guard_kernel(const dg_walk_queue* q, uint32_t* walk_len, uint32_t n) {
  if (!(q->flags & (DG_QFLAG_TOKEN_OVERFLOW | DG_QFLAG_LENGTH_LIMIT))) return;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    walk_len[i] = 0;
}

__global__ void zero_first_kernel(uint64_t* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = 0;
}

__global__ void queue_reset_kernel(dg_walk_queue* q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    q->next = 0;
    q->cursor = 0;
    q->flags = 0;
    q->max_len = 0;
    q->zero_terms = 0;
    q->re = 0.0;
    q->draws = 0;
    q->moves = 0;
    q->tokens = 0;
  }
}

__global__ void add_u32_kernel(uint32_t* dst, const uint32_t* src, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] += src[i];
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- exclusive scan of walk lengths (three phases, 1024 items per block) ----
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* s_wave, uint64_t* total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint64_t incl = v;
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t t = __shfl_up(incl, d, 64);
    if ((int)lane >= d) incl += t;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint64_t wave_prefix = 0, sum = 0;
  for (uint32_t w = 0; w < kScanBlock / 64; ++w) {
    if (w < wave) wave_prefix += s_wave[w];
    sum += s_wave[w];
  }
  __syncthreads();
  *total = sum;
  return wave_prefix + incl - v;
}

__global__ __launch_bounds__(kScanBlock) void scan_tile_sums(const uint32_t* len, uint64_t n,
                                                             uint64_t* tile_sums) {
  __shared__ uint64_t s_wave[kScanBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint64_t s = 0;
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += len[base + k];
  uint64_t total;
  block_exclusive_scan(s, s_wave, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Exclusive scan of the tile sums in one block; tile_sums[n_tiles] gets the total.
__global__ __launch_bounds__(kScanBlock) void scan_tile_prefix(uint64_t* tile_sums,
                                                               uint64_t n_tiles) {
  __shared__ uint64_t s_wave[kScanBlock / 64];
  const uint64_t per = (n_tiles + kScanBlock - 1) / kScanBlock;
  const uint64_t lo = threadIdx.x * per;
  const uint64_t hi = lo + per < n_tiles ? lo + per : n_tiles;
  uint64_t s = 0;
  for (uint64_t i = lo; i < hi; ++i) s += tile_sums[i];
  uint64_t total;
  uint64_t run = block_exclusive_scan(s, s_wave, &total);
  for (uint64_t i = lo; i < hi; ++i) {
    const uint64_t t = tile_sums[i];
    tile_sums[i] = run;
    run += t;
  }
  if (threadIdx.x == 0) tile_sums[n_tiles] = total;
}

__global__ __launch_bounds__(kScanBlock) void scan_tile_apply(const uint32_t* len, uint64_t n,
                                                              const uint64_t* tile_prefix,
                                                              uint64_t n_tiles, uint64_t* off) {
  __shared__ uint64_t s_wave[kScanBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
  uint64_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? len[base + k] : 0;
    s += v[k];
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan(s, s_wave, &total) + tile_prefix[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) off[base + k] = run;
    run += v[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) off[n] = tile_prefix[n_tiles];
}

// ---- corpus packing in walk-id order, and the round's visit counts ------
// (path.hpp:91,95: context_map_freq / vertex_cn).  A half-wavefront (32
// lanes) copies one walk at a time, two walks per pass, in chunks of 32
// tokens.  It sorts the chunk across its lanes (bitonic network), so equal
// vertices become runs, and counts each distinct vertex once with its
// multiplicity (walks revisit their hub at every other step): a hot vertex
// (hot_slot != kNotHot) in this workgroup's LDS counters, written out as one
// partial row at the end, any other vertex with a global atomic.  The
// partial rows are gridDim.x x n_hot.
constexpr int kPackBlock = 256;
constexpr int kPackWalks = 2;  // walks per half-wavefront per pass

__device__ __forceinline__ uint32_t bitonic32(uint32_t x, uint32_t hl) {
#pragma unroll
  for (int k = 2; k <= 32; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint32_t y = __shfl_xor(x, j, 32);
      const bool up = (hl & (uint32_t)k) == 0;
      const bool lower = (hl & (uint32_t)j) == 0;
      x = (lower == up) ? (x < y ? x : y) : (x < y ? y : x);
    }
  }
  return x;
}

__global__ __launch_bounds__(kPackBlock) void pack_count_kernel(
    const uint32_t* raw, const uint64_t* walk_off, const uint32_t* walk_len, const uint64_t* out_off,
    uint64_t n_walks, uint32_t* tokens, const uint16_t* hot_slot, uint32_t n_hot, uint32_t* freq,
    uint32_t* hot_part) {
  __shared__ uint32_t s_hot[kHotMax];
  for (uint32_t k = threadIdx.x; k < n_hot; k += kPackBlock) s_hot[k] = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, hl = lane & 31u, half = lane >> 5;
  const uint64_t wave = ((uint64_t)blockIdx.x * kPackBlock + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * kPackBlock) >> 6;
  const uint64_t per_pass = 2 * kPackWalks;  // walks per wavefront per pass
  for (uint64_t w0 = wave * per_pass; w0 < n_walks; w0 += nwaves * per_pass) {
    uint64_t w[kPackWalks], src[kPackWalks], dst[kPackWalks];
    uint32_t len[kPackWalks];
#pragma unroll
    for (int u = 0; u < kPackWalks; ++u) {
      w[u] = w0 + (uint64_t)(2 * u + half);
      const bool ok = w[u] < n_walks;
      len[u] = ok ? walk_len[w[u]] : 0u;
      src[u] = ok ? walk_off[w[u]] : 0ull;
      dst[u] = ok ? out_off[w[u]] : 0ull;
    }
    // chunks of 32 tokens; the count is made wave-uniform so that both halves
    // take part in every shuffle
    uint32_t nch = 0;
#pragma unroll
    for (int u = 0; u < kPackWalks; ++u) nch = max(nch, (len[u] + 31u) >> 5);
    nch = max(nch, (uint32_t)__shfl_xor((int)nch, 32, 64));
    for (uint32_t c = 0; c < nch; ++c) {
      uint32_t x[kPackWalks], nvalid[kPackWalks];
#pragma unroll
      for (int u = 0; u < kPackWalks; ++u) {
        const uint32_t j = c * 32u + hl;
        nvalid[u] = len[u] > c * 32u ? min(len[u] - c * 32u, 32u) : 0u;
        x[u] = 0xffffffffu;
        if (j < len[u]) {
          x[u] = raw[src[u] + j];
          tokens[dst[u] + j] = x[u];
        }
      }
#pragma unroll
      for (int u = 0; u < kPackWalks; ++u) {
        const uint32_t y = bitonic32(x[u], hl);
        const uint32_t prev = __shfl_up(y, 1, 32);
        const bool head = hl < nvalid[u] && (hl == 0 || prev != y);
        const uint32_t hmask = (uint32_t)(__ballot(head) >> (32 * half));
        if (head) {
          const uint32_t above = hmask & ~((2u << hl) - 1u);
          const uint32_t next = above ? (uint32_t)__builtin_ctz(above) : nvalid[u];
          const uint32_t cnt = next - hl;
          const uint16_t hs = hot_slot[y];
          if (hs != kNotHot)
            atomicAdd(&s_hot[hs], cnt);
          else
            atomicAdd(&freq[y], cnt);
        }
      }
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < n_hot; k += kPackBlock)
    hot_part[(uint64_t)blockIdx.x * n_hot + k] = s_hot[k];
}

// Hot-vertex counts: the partial rows summed per slot.  Grid (slots / 64,
// rows / 64): each thread adds 64 partial rows of one slot (coalesced across
// the slots of a wavefront) and adds that to the vertex's count.
constexpr uint32_t kHotRows = 64;

__global__ __launch_bounds__(64) void hot_reduce_kernel(const uint32_t* hot_part, uint32_t n_parts,
                                                        uint32_t n_hot, const uint32_t* hot_id,
                                                        uint32_t* freq) {
  const uint32_t k = blockIdx.x * 64 + threadIdx.x;
  if (k >= n_hot) return;
  const uint32_t b0 = blockIdx.y * kHotRows;
  const uint32_t b1 = b0 + kHotRows < n_parts ? b0 + kHotRows : n_parts;
  uint32_t c = 0;
  for (uint32_t b = b0; b < b1; ++b) c += hot_part[(uint64_t)b * n_hot + k];
  if (c) atomicAdd(&freq[hot_id[k]], c);
}

// ---- relative entropy (walk.hpp:693-725) in a fixed summation order -------
// terms[i] = p_i (ln p_i - ln q_i); chunk sums of 1024 consecutive terms added
// left to right, then the chunk sums added left to right.
constexpr uint32_t kReChunk = 1024;

// wsum = the visits counted before this round plus the round's tokens (from
// the queue, so the host need not know the round's token count).
__global__ void re_terms_kernel(const uint32_t* deg, const uint32_t* freq, uint32_t v_num,
                                uint64_t dsum, uint64_t wsum_before, double* terms,
                                dg_walk_queue* q) {
  const uint64_t wsum = wsum_before + q->tokens;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < v_num;
       i += gridDim.x * blockDim.x) {
    const uint32_t d = deg[i], f = freq[i];
    double t = 0.0;
    if (d == 0 || f == 0) {
      atomicAdd(&q->zero_terms, 1u);
    } else {
      const double pi = (double)d / (double)dsum;
      const double qi = (double)f / (double)wsum;
      t = pi * (dg_log(pi) - dg_log(qi));
    }
    terms[i] = t;
  }
}

// One wavefront per chunk: the 64 lanes load 64 consecutive terms at a time
// (coalesced) and the sum is taken in index order from the lane registers,
// left to right as the order above requires.
__global__ __launch_bounds__(256) void re_chunks_kernel(const double* terms, const uint32_t* deg,
                                                        const uint32_t* freq, uint32_t v_num,
                                                        double* chunks, uint32_t n_chunks) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (c >= n_chunks) return;  // wave-uniform
  const uint32_t lo = c * kReChunk;
  const uint32_t hi = lo + kReChunk < v_num ? lo + kReChunk : v_num;
  double part = 0.0;
  for (uint32_t i0 = lo; i0 < hi; i0 += 64) {
    const uint32_t i = i0 + lane;
    const bool in = i < hi && deg[i] != 0 && freq[i] != 0;
    const double t = in ? terms[i] : 0.0;
    const uint64_t live = __ballot(in);
    const uint32_t n = hi - i0 < 64u ? hi - i0 : 64u;
    for (uint32_t k = 0; k < n; ++k)
      if ((live >> k) & 1ull) part = part + __shfl(t, (int)k, 64);
  }
  if (lane == 0) chunks[c] = part;
}

// The chunk sums left to right, by one wavefront in the same way.
__global__ __launch_bounds__(64) void re_final_kernel(const double* chunks, uint32_t n_chunks,
                                                      dg_walk_queue* q) {
  const uint32_t lane = threadIdx.x & 63u;
  double total = 0.0;
  for (uint32_t c0 = 0; c0 < n_chunks; c0 += 64) {
    const double t = c0 + lane < n_chunks ? chunks[c0 + lane] : 0.0;
    const uint32_t n = n_chunks - c0 < 64u ? n_chunks - c0 : 64u;
    for (uint32_t k = 0; k < n; ++k) total = total + __shfl(t, (int)k, 64);
  }
  if (lane == 0) q->re = total;
}

__global__ void offsets_shift_kernel(const uint64_t* src, uint64_t n, uint64_t add,
                                     uint64_t* dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i] + add;
}

int grid_for(dg_ctx* ctx, uint64_t n, int block) {
  uint64_t g = (n + block - 1) / block;
  uint64_t cap = (uint64_t)ctx->num_cus * 16;
  if (g > cap) g = cap;
  return g == 0 ? 1 : (int)g;
}

}  // namespace

// A planned round (see plan_round below).
struct dg_walk_plan {
  WalkArgs a;
  uint64_t blocks = 0;        // walk_kernel grid
  uint32_t parts = 0;         // pack_count_kernel grid
  uint32_t round = 0;
  uint32_t walker_num = 0;
  uint32_t length_limit = 0;
  uint64_t freq_total_before = 0;
};


void dg_walk_free(dg_ctx* ctx) {
  void* bufs[] = {ctx->d_freq,   ctx->d_tok_raw,  ctx->d_walk_off, ctx->d_walk_len,
                  ctx->d_tokens, ctx->d_offsets,  ctx->d_spill,    ctx->d_scan_tmp,
                  ctx->d_re_terms, ctx->d_re_chunks, ctx->d_train_tokens, ctx->d_train_offsets,
                  ctx->d_term_table, ctx->d_freq_round};
  for (void* p : bufs)
    if (p) hipFree(p);
  ctx->d_freq = nullptr;
  ctx->d_tok_raw = nullptr;
  ctx->d_walk_off = nullptr;
  ctx->d_walk_len = nullptr;
  ctx->d_tokens = nullptr;
  ctx->d_offsets = nullptr;
  ctx->d_spill = nullptr;
  ctx->d_scan_tmp = nullptr;
  ctx->d_re_terms = nullptr;
  ctx->d_re_chunks = nullptr;
  ctx->d_train_tokens = nullptr;
  ctx->d_train_offsets = nullptr;
  ctx->d_term_table = nullptr;
  ctx->d_freq_round = nullptr;
  ctx->freq_round_cap = 0;
  delete ctx->wplan;
  ctx->wplan = nullptr;
  ctx->tok_raw_cap = ctx->walk_cap = ctx->walk_len_cap = ctx->tokens_cap = ctx->offsets_cap = 0;
  ctx->spill_cap = ctx->scan_tmp_cap = ctx->re_cap = 0;
  ctx->train_tokens_cap = ctx->train_offsets_cap = 0;
  ctx->train_walks = ctx->train_tokens = 0;
  ctx->cur_walks = ctx->cur_tokens = 0;
  ctx->freq_total = 0;
  ctx->prev_re = 0.0;
  ctx->rounds_done = 0;
}

int dg_scan_lengths(dg_ctx* ctx, const uint32_t* d_len, uint64_t n, uint64_t* d_off) {
  const uint64_t n_tiles = (n + kScanTile - 1) / kScanTile;
  DG_CHECK(dg_reserve(ctx, &ctx->d_scan_tmp, &ctx->scan_tmp_cap, n_tiles + 1));
  if (n_tiles > 0) {
    scan_tile_sums<<<(unsigned)n_tiles, kScanBlock, 0, ctx->stream>>>(d_len, n, ctx->d_scan_tmp);
    scan_tile_prefix<<<1, kScanBlock, 0, ctx->stream>>>(ctx->d_scan_tmp, n_tiles);
    scan_tile_apply<<<(unsigned)n_tiles, kScanBlock, 0, ctx->stream>>>(d_len, n, ctx->d_scan_tmp,
                                                                       n_tiles, d_off);
  } else {
    zero_first_kernel<<<1, 64, 0, ctx->stream>>>(d_off);
  }
  DG_HIP(ctx, hipGetLastError());
  return DG_OK;
}

extern "C" void dg_walk_params_default(dg_walk_params* p, uint32_t v_num) {
  if (!p) return;
  p->seed = 1;
  p->walkers_per_round = v_num;
  p->init_round = 5;
  p->min_length = 20;
  p->max_length = 0;
  p->mu = 0.999;
  p->delta = 0.001;
  p->max_rounds = 30;
  p->rounds_mode = DG_ROUNDS_REFERENCE;
  p->length_limit = 0;
  p->reserved = 0;
}

extern "C" int dg_walk_reset(dg_ctx* ctx) {
  if (!ctx) return DG_ERR_ARG;
  if (!ctx->d_row_ptr) return dg_fail(ctx, DG_ERR_STATE, "no graph loaded");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  if (!ctx->d_freq) {
    DG_HIP(ctx, hipMalloc((void**)&ctx->d_freq, sizeof(uint32_t) * ctx->v_num));
  }
  if (!ctx->d_term_table) {
    DG_HIP(ctx, hipMalloc((void**)&ctx->d_term_table, sizeof(double) * kTermTable * kTermStride));
    term_table_kernel<<<(kTermTable * kTermStride + 255) / 256, 256, 0, ctx->stream>>>(
        ctx->d_term_table);
    DG_HIP(ctx, hipGetLastError());
  }
  fill_u32_kernel<<<grid_for(ctx, ctx->v_num, 256), 256, 0, ctx->stream>>>(ctx->d_freq, ctx->v_num,
                                                                           0u);
  DG_HIP(ctx, hipGetLastError());
  ctx->freq_total = 0;
  ctx->prev_re = 0.0;
  ctx->rounds_done = 0;
  return DG_OK;
}

// ---- one round, in three parts ---------------------------------------------
// plan: validation, launch shapes and buffers (may allocate); enqueue: kernels
// only, no host waits, so a stream capture can record them; collect: the
// round's statistics once the stream has finished.  dg_walk_round runs them
// with a host check of the token buffer between the walk and the packing;
// the captured pipeline step (pipeline.hip) puts round_guard_kernel there.

static int plan_round(dg_ctx* ctx, const dg_walk_params* p, uint32_t round,
                      uint32_t walker_begin, uint32_t walker_num, dg_walk_plan* pl) {
  if (!ctx->d_freq) DG_CHECK(dg_walk_reset(ctx));
  const uint32_t length_limit = p->length_limit ? p->length_limit : kDefaultLengthLimit;
  if (p->max_length != 0 && p->max_length > length_limit)
    return dg_fail(ctx, DG_ERR_ARG, "max_length %u exceeds length_limit %u", p->max_length,
                   length_limit);
  if (p->min_length + 2 > length_limit)
    return dg_fail(ctx, DG_ERR_ARG, "min_length %u needs length_limit > %u", p->min_length,
                   p->min_length + 1);
  // the buffers below may be reallocated (their contents dropped): the
  // previous round's corpus is given up here, so a round that fails leaves an
  // empty corpus behind, never one whose buffers no longer hold it
  ctx->cur_walks = 0;
  ctx->cur_tokens = 0;

  // persistent grid: every resident block slot on the device, fewer for small rounds
  int per_cu = 0;
  DG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, walk_kernel, kWalkBlock, 0));
  if (per_cu < 1) per_cu = 1;
  uint64_t blocks = (uint64_t)ctx->num_cus * per_cu;
  const uint64_t need_blocks = ((uint64_t)walker_num + kWalkBlock - 1) / kWalkBlock;
  if (blocks > need_blocks) blocks = need_blocks ? need_blocks : 1;
  const uint32_t spill_per_slot =
      length_limit > (uint32_t)kPathLds ? length_limit - (uint32_t)kPathLds : 0;

  DG_CHECK(dg_reserve(ctx, &ctx->d_walk_off, &ctx->walk_cap, (size_t)walker_num));
  DG_CHECK(dg_reserve(ctx, &ctx->d_walk_len, &ctx->walk_len_cap, (size_t)walker_num + 1));
  DG_CHECK(dg_reserve(ctx, &ctx->d_offsets, &ctx->offsets_cap, (size_t)walker_num + 1));
  DG_CHECK(dg_reserve(ctx, &ctx->d_spill, &ctx->spill_cap,
                      (size_t)blocks * kWalkBlock * spill_per_slot));
  // raw buffer: one slot of kPathLds tokens per walk, then room for longer walks
  const uint64_t slots_end = (uint64_t)walker_num * kPathLds;
  size_t tok_need = (size_t)slots_end + (size_t)walker_num / 8 + 4096;
  if (ctx->tok_raw_cap < tok_need) DG_CHECK(dg_reserve(ctx, &ctx->d_tok_raw, &ctx->tok_raw_cap, tok_need));
  // this round's visits go to a separate count array, added to the cumulative
  // counts once the round is packed
  DG_CHECK(dg_reserve(ctx, &ctx->d_freq_round, &ctx->freq_round_cap, (size_t)ctx->v_num));
  {
    // packing: every resident workgroup slot (LDS: n_hot counters each), fewer for small rounds
    int pcu = 0;
    DG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&pcu, pack_count_kernel, kPackBlock, 0));
    if (pcu < 1) pcu = 1;
    const uint64_t per_wg = (uint64_t)(kPackBlock / 64) * 2 * kPackWalks;  // walks per workgroup pass
    const uint64_t need = ((uint64_t)walker_num + per_wg - 1) / per_wg;
    pl->parts = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * pcu, need ? need : 1);
    DG_CHECK(dg_reserve(ctx, &ctx->d_hot_part, &ctx->hot_part_cap,
                        (size_t)pl->parts * (ctx->n_hot ? ctx->n_hot : 1)));
  }
  const uint32_t n_chunks = (ctx->v_num + kReChunk - 1) / kReChunk;
  DG_CHECK(dg_reserve(ctx, &ctx->d_re_terms, &ctx->re_cap, (size_t)ctx->v_num + n_chunks));
  ctx->d_re_chunks = nullptr;  // chunks live after the terms
  DG_CHECK(dg_reserve(ctx, &ctx->d_scan_tmp, &ctx->scan_tmp_cap,
                      (size_t)((walker_num + kScanTile - 1) / kScanTile) + 1));

  WalkArgs& a = pl->a;
  a.row_ptr = ctx->d_row_ptr;
  a.adj = ctx->d_adj;
  a.term_table = ctx->d_term_table;
  a.v_num = ctx->v_num;
  a.round = round;
  a.seed = p->seed;
  a.walker_begin = walker_begin;
  a.walker_num = walker_num;
  a.min_length = p->min_length;
  a.max_length = p->max_length;
  a.length_limit = length_limit;
  a.spill_per_slot = spill_per_slot;
  a.mu = p->mu;
  a.walk_off = ctx->d_walk_off;
  a.walk_len = ctx->d_walk_len;
  a.spill = ctx->d_spill;
  a.q = ctx->d_queue;
  a.slots_end = slots_end;
  a.tokens = ctx->d_tok_raw;
  a.tok_cap = ctx->tok_raw_cap;
  pl->blocks = blocks;
  pl->round = round;
  pl->walker_num = walker_num;
  pl->length_limit = length_limit;
  pl->freq_total_before = ctx->freq_total;
  return DG_OK;
}

// The walk itself: queue reset and the persistent launch.
static int enqueue_walk(dg_ctx* ctx, const dg_walk_plan& pl, bool timed) {
  queue_reset_kernel<<<1, 64, 0, ctx->stream>>>(ctx->d_queue);
  size_t tk = 0;
  if (timed) DG_CHECK(dg_time_begin(ctx, "walk", &tk));
  walk_kernel<<<(unsigned)pl.blocks, kWalkBlock, 0, ctx->stream>>>(pl.a);
  DG_HIP(ctx, hipGetLastError());
  if (timed) DG_CHECK(dg_time_end(ctx, tk));
  return DG_OK;
}

// Everything after the walk: packing in walk-id order with the round's visit
// counts, the cumulative counts, and the relative entropy.  Reads only device
// state, so it needs no host knowledge of the round's token count.
static int enqueue_tail(dg_ctx* ctx, const dg_walk_plan& pl, bool timed) {
  const uint32_t walker_num = pl.walker_num;
  size_t tp = 0;
  if (timed) DG_CHECK(dg_time_begin(ctx, "pack", &tp));
  DG_CHECK(dg_scan_lengths(ctx, ctx->d_walk_len, walker_num, ctx->d_offsets));
  fill_u32_kernel<<<grid_for(ctx, ctx->v_num, 256), 256, 0, ctx->stream>>>(ctx->d_freq_round,
                                                                           ctx->v_num, 0u);
  pack_count_kernel<<<pl.parts, kPackBlock, 0, ctx->stream>>>(
      ctx->d_tok_raw, ctx->d_walk_off, ctx->d_walk_len, ctx->d_offsets, walker_num, ctx->d_tokens,
      ctx->d_hot_slot, ctx->n_hot, ctx->d_freq_round, ctx->d_hot_part);
  if (ctx->n_hot) {
    const dim3 grid((ctx->n_hot + 63) / 64, (pl.parts + kHotRows - 1) / kHotRows);
    hot_reduce_kernel<<<grid, 64, 0, ctx->stream>>>(ctx->d_hot_part, pl.parts, ctx->n_hot,
                                                    ctx->d_hot_id, ctx->d_freq_round);
  }
  DG_HIP(ctx, hipGetLastError());
  if (timed) DG_CHECK(dg_time_end(ctx, tp));
  add_u32_kernel<<<grid_for(ctx, ctx->v_num, 256), 256, 0, ctx->stream>>>(
      ctx->d_freq, ctx->d_freq_round, ctx->v_num);
  DG_HIP(ctx, hipGetLastError());

  // relative entropy against the cumulative visit counts
  const uint32_t n_chunks = (ctx->v_num + kReChunk - 1) / kReChunk;
  double* chunks = ctx->d_re_terms + ctx->v_num;
  size_t tr = 0;
  if (timed) DG_CHECK(dg_time_begin(ctx, "re", &tr));
  re_terms_kernel<<<grid_for(ctx, ctx->v_num, 256), 256, 0, ctx->stream>>>(
      ctx->d_deg, ctx->d_freq, ctx->v_num, ctx->m, pl.freq_total_before, ctx->d_re_terms,
      ctx->d_queue);
  re_chunks_kernel<<<(n_chunks + 3) / 4, 256, 0, ctx->stream>>>(ctx->d_re_terms, ctx->d_deg,
                                                                ctx->d_freq, ctx->v_num, chunks,
                                                                n_chunks);
  re_final_kernel<<<1, 64, 0, ctx->stream>>>(chunks, n_chunks, ctx->d_queue);
  DG_HIP(ctx, hipGetLastError());
  if (timed) DG_CHECK(dg_time_end(ctx, tr));
  return DG_OK;
}

// The round's statistics from the queue mirror (the stream has finished and
// h_queue holds the final queue).  Advances the host-side round state.
static int collect_round(dg_ctx* ctx, const dg_walk_plan& pl, double walk_ms, double total_ms,
                         dg_round_stats* out) {
  const dg_walk_queue& q = *ctx->h_queue;
  const uint64_t n_tok = q.tokens;
  ctx->freq_total = pl.freq_total_before + n_tok;
  ctx->cur_walks = pl.walker_num;
  ctx->cur_tokens = n_tok;
  const double re = q.re;
  const double delta = ctx->rounds_done == 0 ? re : re - ctx->prev_re;
  ctx->prev_re = re;
  ctx->rounds_done += 1;
  if (out) {
    out->round = pl.round;
    out->walkers = pl.walker_num;
    out->tokens = n_tok;
    out->steps = n_tok - pl.walker_num;
    out->re = re;
    out->delta_re = delta;
    out->walk_ms = walk_ms;
    out->total_ms = total_ms;
    out->max_walk_len = q.max_len;
    out->draws = q.draws;
    out->moves = q.moves;
    out->stop = 0;
  }
  if (q.zero_terms != 0)
    return dg_fail(ctx, DG_ERR_RE_UNDEFINED,
                   "relative entropy undefined: %u vertices with degree 0 or no visits",
                   q.zero_terms);
  return DG_OK;
}

static int read_queue(dg_ctx* ctx) {
  DG_HIP(ctx, hipMemcpyAsync(ctx->h_queue, ctx->d_queue, sizeof(dg_walk_queue),
                             hipMemcpyDeviceToHost, ctx->stream));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DG_OK;
}

static int walk_round_impl(dg_ctx* ctx, const dg_walk_params* p, uint32_t round,
                           uint32_t walker_begin, uint32_t walker_num, dg_round_stats* out) {
  dg_walk_plan pl;
  DG_CHECK(plan_round(ctx, p, round, walker_begin, walker_num, &pl));
  for (int attempt = 0; attempt < 3; ++attempt) {
    pl.a.tokens = ctx->d_tok_raw;
    pl.a.tok_cap = ctx->tok_raw_cap;
    DG_CHECK(enqueue_walk(ctx, pl, true));
    DG_CHECK(read_queue(ctx));
    if (ctx->h_queue->flags & DG_QFLAG_LENGTH_LIMIT)
      return dg_fail(ctx, DG_ERR_WALK_TOO_LONG,
                     "round %u: a walk reached the length limit %u (raise length_limit)", round,
                     pl.length_limit);
    if (!(ctx->h_queue->flags & DG_QFLAG_TOKEN_OVERFLOW)) break;
    const size_t need =
        (size_t)(pl.a.slots_end + ctx->h_queue->cursor + ctx->h_queue->cursor / 8 + 1024);
    DG_CHECK(dg_reserve(ctx, &ctx->d_tok_raw, &ctx->tok_raw_cap, need));
    if (attempt == 2) return dg_fail(ctx, DG_ERR_STATE, "token buffer sizing did not converge");
  }
  DG_CHECK(dg_reserve(ctx, &ctx->d_tokens, &ctx->tokens_cap, (size_t)ctx->h_queue->tokens));
  DG_CHECK(enqueue_tail(ctx, pl, true));
  DG_CHECK(read_queue(ctx));

  double walk_ms = 0.0, total_ms = 0.0;
  for (auto& pe : ctx->pending) {
    float ms = 0.f;
    DG_HIP(ctx, hipEventElapsedTime(&ms, pe.start, pe.stop));
    if (pe.name == "walk") walk_ms = ms;  // the successful (last) attempt
    total_ms += ms;
  }
  DG_CHECK(dg_time_resolve(ctx));
  return collect_round(ctx, pl, walk_ms, total_ms, out);
}

extern "C" int dg_walk_round(dg_ctx* ctx, const dg_walk_params* p, uint32_t round,
                             uint32_t walker_begin, uint32_t walker_num, dg_round_stats* out) {
  if (!ctx || !p) return DG_ERR_ARG;
  if (!ctx->d_row_ptr) return dg_fail(ctx, DG_ERR_STATE, "no graph loaded");
  if (!(p->mu > 0.0)) return dg_fail(ctx, DG_ERR_ARG, "mu must be positive");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  return walk_round_impl(ctx, p, round, walker_begin, walker_num, out);
}

// ---- the round as the captured pipeline step runs it -----------------------

int dg_walk_graph_plan(dg_ctx* ctx, const dg_walk_params* p, uint32_t round,
                       uint32_t walker_begin, uint32_t walker_num, uint64_t* tok_bound,
                       uint64_t* slot_tokens, uint64_t* overflow_tokens) {
  if (!ctx->d_row_ptr) return dg_fail(ctx, DG_ERR_STATE, "no graph loaded");
  if (!(p->mu > 0.0)) return dg_fail(ctx, DG_ERR_ARG, "mu must be positive");
  if (!ctx->wplan) ctx->wplan = new dg_walk_plan();
  DG_CHECK(plan_round(ctx, p, round, walker_begin, walker_num, ctx->wplan));
  // the packed corpus can hold every token the raw buffer can
  DG_CHECK(dg_reserve(ctx, &ctx->d_tokens, &ctx->tokens_cap, ctx->tok_raw_cap));
  *tok_bound = ctx->tok_raw_cap;
  *slot_tokens = kPathLds;
  *overflow_tokens = ctx->tok_raw_cap - ctx->wplan->a.slots_end;
  return DG_OK;
}

int dg_walk_graph_enqueue(dg_ctx* ctx, bool timed) {
  const dg_walk_plan& pl = *ctx->wplan;
  DG_CHECK(enqueue_walk(ctx, pl, timed));
  // a round that overflowed the token buffer or hit the length limit becomes
  // empty here, so the rest of the step changes nothing and can run again
  round_guard_kernel<<<grid_for(ctx, pl.walker_num, 256), 256, 0, ctx->stream>>>(
      ctx->d_queue, ctx->d_walk_len, pl.walker_num);
  DG_HIP(ctx, hipGetLastError());
  return enqueue_tail(ctx, pl, timed);
}

int dg_walk_graph_collect(dg_ctx* ctx, double walk_ms, double total_ms, dg_round_stats* out,
                          bool* retry) {
  const dg_walk_plan& pl = *ctx->wplan;
  *retry = false;
  DG_CHECK(read_queue(ctx));
  const dg_walk_queue& q = *ctx->h_queue;
  if (q.flags & DG_QFLAG_LENGTH_LIMIT)
    return dg_fail(ctx, DG_ERR_WALK_TOO_LONG,
                   "round %u: a walk reached the length limit %u (raise length_limit)", pl.round,
                   pl.length_limit);
  if (q.flags & DG_QFLAG_TOKEN_OVERFLOW) {
    const size_t need = (size_t)(pl.a.slots_end + q.cursor + q.cursor / 8 + 1024);
    DG_CHECK(dg_reserve(ctx, &ctx->d_tok_raw, &ctx->tok_raw_cap, need));
    *retry = true;
    return DG_OK;
  }
  return collect_round(ctx, pl, walk_ms, total_ms, out);
}

extern "C" int dg_corpus_append(dg_ctx* ctx) {
  if (!ctx) return DG_ERR_ARG;
  DG_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t nw = ctx->cur_walks, nt = ctx->cur_tokens;
  DG_CHECK(dg_grow_keep(ctx, &ctx->d_train_tokens, &ctx->train_tokens_cap,
                        (size_t)(ctx->train_tokens + nt), (size_t)ctx->train_tokens));
  DG_CHECK(dg_grow_keep(ctx, &ctx->d_train_offsets, &ctx->train_offsets_cap,
                        (size_t)(ctx->train_walks + nw + 1), (size_t)(ctx->train_walks + 1)));
  if (ctx->train_walks == 0)
    DG_HIP(ctx, hipMemsetAsync(ctx->d_train_offsets, 0, sizeof(uint64_t), ctx->stream));
  if (nt)
    DG_HIP(ctx, hipMemcpyAsync(ctx->d_train_tokens + ctx->train_tokens, ctx->d_tokens,
                               sizeof(uint32_t) * nt, hipMemcpyDeviceToDevice, ctx->stream));
  if (nw)
    offsets_shift_kernel<<<grid_for(ctx, nw, 256), 256, 0, ctx->stream>>>(
        ctx->d_offsets + 1, nw, ctx->train_tokens, ctx->d_train_offsets + ctx->train_walks + 1);
  DG_HIP(ctx, hipGetLastError());
  ctx->train_walks += nw;
  ctx->train_tokens += nt;
  return DG_OK;
}

extern "C" int dg_train_corpus_info(dg_ctx* ctx, uint64_t* n_walks, uint64_t* n_tokens) {
  if (!ctx) return DG_ERR_ARG;
  if (n_walks) *n_walks = ctx->train_walks;
  if (n_tokens) *n_tokens = ctx->train_tokens;
  return DG_OK;
}

extern "C" int dg_train_corpus_clear(dg_ctx* ctx) {
  if (!ctx) return DG_ERR_ARG;
  ctx->train_walks = 0;
  ctx->train_tokens = 0;
  return DG_OK;
}

extern "C" int dg_walk_run(dg_ctx* ctx, const dg_walk_params* p, dg_round_stats* per_round,
                           uint32_t max_stats, uint32_t* n_rounds, int keep_corpus) {
  if (!ctx || !p) return DG_ERR_ARG;
  if (!ctx->d_row_ptr) return dg_fail(ctx, DG_ERR_STATE, "no graph loaded");
  if (p->walkers_per_round == 0 || p->init_round == 0)
    return dg_fail(ctx, DG_ERR_ARG, "walkers_per_round and init_round must be positive");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  DG_CHECK(dg_walk_reset(ctx));
  uint32_t iter = 1;
  uint32_t done = 0;
  const uint32_t max_rounds = p->max_rounds ? p->max_rounds : 30;
  while (true) {
    // walk.hpp:614 (round 1: walker_num * init_round), :764-775 (later rounds)
    const uint64_t nw64 = (iter == 1) ? (uint64_t)p->walkers_per_round * p->init_round
                                      : (uint64_t)p->walkers_per_round;
    if (nw64 > 0xffffffffull) return dg_fail(ctx, DG_ERR_ARG, "too many walkers per round");
    dg_round_stats st;
    DG_CHECK(walk_round_impl(ctx, p, iter, 0, (uint32_t)nw64, &st));
    if (keep_corpus) DG_CHECK(dg_corpus_append(ctx));
    // walk.hpp:735-749: stop after round max_rounds; optionally on |delta RE| <= delta
    bool stop = (iter + 1 > max_rounds);
    if (p->rounds_mode == DG_ROUNDS_CONVERGE && (st.delta_re < 0 ? -st.delta_re : st.delta_re) <= p->delta)
      stop = true;
    st.stop = stop ? 1u : 0u;
    if (per_round && done < max_stats) per_round[done] = st;
    ++done;
    ++iter;
    if (stop) break;
  }
  if (n_rounds) *n_rounds = done;
  return DG_OK;
}

extern "C" int dg_corpus_info(dg_ctx* ctx, uint64_t* n_walks, uint64_t* n_tokens) {
  if (!ctx) return DG_ERR_ARG;
  if (n_walks) *n_walks = ctx->cur_walks;
  if (n_tokens) *n_tokens = ctx->cur_tokens;
  return DG_OK;
}

extern "C" int dg_corpus_copy(dg_ctx* ctx, uint32_t* tokens, uint64_t* offsets) {
  if (!ctx) return DG_ERR_ARG;
  DG_HIP(ctx, hipSetDevice(ctx->device));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (tokens && ctx->cur_tokens)
    DG_HIP(ctx, hipMemcpy(tokens, ctx->d_tokens, sizeof(uint32_t) * ctx->cur_tokens,
                          hipMemcpyDeviceToHost));
  if (offsets && ctx->d_offsets)
    DG_HIP(ctx, hipMemcpy(offsets, ctx->d_offsets, sizeof(uint64_t) * (ctx->cur_walks + 1),
                          hipMemcpyDeviceToHost));
  return DG_OK;
}

extern "C" int dg_vertex_freq(dg_ctx* ctx, uint32_t* freq) {
  if (!ctx || !freq) return DG_ERR_ARG;
  if (!ctx->d_freq) return dg_fail(ctx, DG_ERR_STATE, "no walk has run");
  DG_HIP(ctx, hipSetDevice(ctx->device));
  DG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DG_HIP(ctx, hipMemcpy(freq, ctx->d_freq, sizeof(uint32_t) * ctx->v_num, hipMemcpyDeviceToHost));
  return DG_OK;
}

// Walk text file of include/path.hpp:67-119.  Formatting is done into a large
// buffer with a hand-rolled decimal conversion; the bytes equal fprintf's.
extern "C" int dg_corpus_write_text(dg_ctx* ctx, const char* path, uint32_t threads) {
  if (!ctx || !path) return DG_ERR_ARG;
  if (threads == 0) threads = 1;
  std::vector<uint32_t> tok(ctx->cur_tokens);
  std::vector<uint64_t> off(ctx->cur_walks + 1);
  DG_CHECK(dg_corpus_copy(ctx, tok.data(), off.data()));
  FILE* f = fopen(path, "w");
  if (!f) return dg_fail(ctx, DG_ERR_IO, "cannot open %s for writing", path);
  std::vector<char> buf;
  buf.reserve(1 << 22);
  auto flush = [&]() -> bool {
    bool ok = buf.empty() || fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    buf.clear();
    return ok;
  };
  bool ok = true;
  for (uint32_t t = 0; t < threads && ok; ++t) {
    for (uint64_t w = t; w < ctx->cur_walks && ok; w += threads) {
      for (uint64_t i = off[w]; i < off[w + 1]; ++i) {
        char digits[12];
        int nd = 0;
        uint32_t v = tok[i];
        do {
          digits[nd++] = (char)('0' + v % 10);
          v /= 10;
        } while (v);
        buf.push_back(' ');
        while (nd) buf.push_back(digits[--nd]);
      }
      buf.push_back('\n');
      if (buf.size() > (1u << 22) - 4096) ok = flush();
    }
  }
  ok = flush() && ok;
  if (fclose(f) != 0) ok = false;
  return ok ? DG_OK : dg_fail(ctx, DG_ERR_IO, "write to %s failed", path);
}
