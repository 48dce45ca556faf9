/*
 * dg_philox.h - counter-based random streams shared by the gfx950 kernels and
 * the CPU oracle.
 *
 * The reference walker draws from per-thread std::mt19937 seeded from
 * std::random_device (include/util.hpp:22-48) and from libc rand() shared by
 * all OpenMP threads (include/walk.hpp:928), so no two runs agree.  Here every
 * draw comes from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11), keyed by the user seed, with a counter
 * that names the stream (walk id, round, purpose) and the block index.  A walk
 * therefore consumes the same words whichever lane, thread or device runs it.
 *
 * Stream contract (used by both sides):
 *   key     = (seed & 0xffffffff, seed >> 32)
 *   counter = (block, c1, c2, tag);  block = 0, 1, 2, ... ;  4 words per block,
 *             consumed in order out[0], out[1], out[2], out[3].
 *   uniform integer in [0, n): Lemire's multiply-shift with rejection
 *             ("Fast random integer generation in an interval", 2019).
 *   31-bit draw (stand-in for rand(), RAND_MAX = 2^31 - 1): next() >> 1.
 */
#ifndef DG_PHILOX_H
#define DG_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DG_RNG_HD static __host__ __device__ __forceinline__
#else
#define DG_RNG_HD static inline
#endif

#define DG_PHILOX_M0 0xD2511F53u
#define DG_PHILOX_M1 0xCD9E8D57u
#define DG_PHILOX_W0 0x9E3779B9u
#define DG_PHILOX_W1 0xBB67AE85u

/* Stream tags (counter word 3). */
#define DG_TAG_WALK 0x57414C4Bu   /* walker transitions and acceptance */
#define DG_TAG_SAMPLE 0x53554253u /* SGNS frequent-word subsampling */
#define DG_TAG_SGNS 0x53474E53u   /* SGNS window shrink and negatives */
#define DG_TAG_INIT 0x494E4954u   /* random embedding initialisation */
#define DG_TAG_GEN 0x47454E52u    /* synthetic graph generators */
#define DG_TAG_SPLIT 0x53504C54u  /* link-prediction split */
#define DG_TAG_TUPLE 0x54555045u  /* synthetic tuples and embedding tables */
#define DG_TAG_DSGL_WIN 0x44535749u /* DSGL trainer: window shrink per (worker, step) */
#define DG_TAG_DSGL_NEG 0x4453474Eu /* DSGL trainer: negative batch per (worker, batch) */
#define DG_TAG_SGNS_REUSE 0x52555345u /* negative-reuse trainer: negatives per sentence */

typedef struct {
  uint32_t x0, x1, x2, x3;
} dg_u32x4;

DG_RNG_HD dg_u32x4 dg_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                    uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)DG_PHILOX_M0 * (uint64_t)c0;
    uint64_t p1 = (uint64_t)DG_PHILOX_M1 * (uint64_t)c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += DG_PHILOX_W0;
    k1 += DG_PHILOX_W1;
  }
  dg_u32x4 o;
  o.x0 = c0;
  o.x1 = c1;
  o.x2 = c2;
  o.x3 = c3;
  return o;
}

/* A sequential stream of 32-bit words.  The buffered words shift down as they
 * are consumed, so no dynamically indexed array lives in registers. */
typedef struct {
  uint32_t k0, k1, c1, c2, tag;
  uint32_t block;
  uint32_t b0, b1, b2, b3;
  uint32_t avail;
} dg_stream;

DG_RNG_HD void dg_stream_init(dg_stream* s, uint64_t seed, uint32_t c1, uint32_t c2,
                              uint32_t tag) {
  s->k0 = (uint32_t)seed;
  s->k1 = (uint32_t)(seed >> 32);
  s->c1 = c1;
  s->c2 = c2;
  s->tag = tag;
  s->block = 0;
  s->b0 = s->b1 = s->b2 = s->b3 = 0;
  s->avail = 0;
}

DG_RNG_HD uint32_t dg_stream_next(dg_stream* s) {
  if (s->avail == 0) {
    dg_u32x4 o = dg_philox4x32_10(s->block, s->c1, s->c2, s->tag, s->k0, s->k1);
    s->block += 1;
    s->b0 = o.x0;
    s->b1 = o.x1;
    s->b2 = o.x2;
    s->b3 = o.x3;
    s->avail = 4;
  }
  uint32_t r = s->b0;
  s->b0 = s->b1;
  s->b1 = s->b2;
  s->b2 = s->b3;
  s->avail -= 1;
  return r;
}

/* Key of a stream that belongs to one training epoch: the epoch number goes
 * into the high key word, so the counters of different epochs never need to
 * share bits (each epoch is its own family of streams under the user seed). */
DG_RNG_HD uint64_t dg_epoch_key(uint64_t seed, uint32_t epoch) {
  return seed ^ ((uint64_t)epoch << 32);
}

/* Uniform integer in [0, n), n >= 1 (Lemire 2019, with rejection). */
DG_RNG_HD uint32_t dg_stream_uniform(dg_stream* s, uint32_t n) {
  uint32_t x = dg_stream_next(s);
  uint64_t m = (uint64_t)x * (uint64_t)n;
  uint32_t l = (uint32_t)m;
  if (l < n) {
    uint32_t t = (uint32_t)(0u - n) % n;
    while (l < t) {
      x = dg_stream_next(s);
      m = (uint64_t)x * (uint64_t)n;
      l = (uint32_t)m;
    }
  }
  return (uint32_t)(m >> 32);
}

/* 31-bit draw standing in for libc rand() with RAND_MAX = 2^31 - 1. */
DG_RNG_HD uint32_t dg_stream_u31(dg_stream* s) { return dg_stream_next(s) >> 1; }

#endif /* DG_PHILOX_H */
