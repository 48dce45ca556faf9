/*
 * dg_math.h - deterministic FP64 elementary functions shared by the gfx950
 * kernels and the CPU oracle.
 *
 * The information-oriented walker decides walk termination from FP64 entropy
 * statistics (include/walk.hpp:284-354, `pow`, `log2`, `sqrt`) and walk
 * acceptance from `tanh` (include/walk.hpp:933-938); the round driver's
 * relative entropy uses `log` (include/walk.hpp:724).  Those calls go to glibc
 * in the reference and to ROCm's device library on the GPU, whose last bits
 * differ.  Every function here is built only from IEEE-754 correctly rounded
 * operations (+ - * / and fused multiply-add) plus integer bit manipulation, so
 * a host build and a device build of the same call return identical bits.
 * Every translation unit that includes this header must be compiled with
 * -ffp-contract=off (gcc and hipcc), otherwise the compiler may fuse a*b+c and
 * change the rounding.
 *
 * Internals use double-double arithmetic (Dekker / Knuth error-free
 * transformations), giving results within about one unit in the last place
 * and almost always correctly rounded; the tests measure agreement with glibc.
 */
#ifndef DG_MATH_H
#define DG_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DG_HD static __host__ __device__ __forceinline__
#define DG_FMA(a, b, c) __builtin_fma((a), (b), (c))
#else
#include <math.h>
#define DG_HD static inline
#define DG_FMA(a, b, c) fma((a), (b), (c))
#endif

#if defined(__clang__)
#pragma STDC FP_CONTRACT OFF
#endif

typedef struct {
  double hi;
  double lo;
} dg_dd;

DG_HD uint64_t dg_bits(double x) {
  union {
    double d;
    uint64_t u;
  } c;
  c.d = x;
  return c.u;
}

DG_HD double dg_from_bits(uint64_t u) {
  union {
    double d;
    uint64_t u;
  } c;
  c.u = u;
  return c.d;
}

DG_HD dg_dd dd_make(double hi, double lo) {
  dg_dd r;
  r.hi = hi;
  r.lo = lo;
  return r;
}

/* Knuth two-sum: s + e == a + b exactly. */
DG_HD dg_dd dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return dd_make(s, e);
}

/* Dekker fast two-sum, requires |a| >= |b|. */
DG_HD dg_dd dd_fast_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return dd_make(s, e);
}

/* p + e == a * b exactly (fma gives the rounding error of the product). */
DG_HD dg_dd dd_two_prod(double a, double b) {
  double p = a * b;
  double e = DG_FMA(a, b, -p);
  return dd_make(p, e);
}

DG_HD dg_dd dd_add(dg_dd a, dg_dd b) {
  dg_dd s = dd_two_sum(a.hi, b.hi);
  dg_dd t = dd_two_sum(a.lo, b.lo);
  double lo = s.lo + t.hi;
  s = dd_fast_two_sum(s.hi, lo);
  lo = s.lo + t.lo;
  return dd_fast_two_sum(s.hi, lo);
}

DG_HD dg_dd dd_add_d(dg_dd a, double b) {
  dg_dd s = dd_two_sum(a.hi, b);
  double lo = s.lo + a.lo;
  return dd_fast_two_sum(s.hi, lo);
}

DG_HD dg_dd dd_mul(dg_dd a, dg_dd b) {
  dg_dd p = dd_two_prod(a.hi, b.hi);
  double cross = a.hi * b.lo + a.lo * b.hi;
  double lo = p.lo + cross;
  return dd_fast_two_sum(p.hi, lo);
}

DG_HD dg_dd dd_mul_d(dg_dd a, double b) {
  dg_dd p = dd_two_prod(a.hi, b);
  double lo = p.lo + a.lo * b;
  return dd_fast_two_sum(p.hi, lo);
}

DG_HD dg_dd dd_div(dg_dd a, dg_dd b) {
  double q1 = a.hi / b.hi;
  dg_dd r = dd_add(a, dd_mul_d(b, -q1));
  double q2 = r.hi / b.hi;
  r = dd_add(r, dd_mul_d(b, -q2));
  double q3 = r.hi / b.hi;
  dg_dd q = dd_fast_two_sum(q1, q2);
  return dd_add_d(q, q3);
}

/* ln 2 and log2(e) as double-double (hi, lo). */
#define DG_LN2_HI 0x1.62e42fefa39efp-1
#define DG_LN2_LO 0x1.abc9e3b39803fp-56
#define DG_LOG2E_HI 0x1.71547652b82fep+0
#define DG_LOG2E_LO 0x1.777d0ffda0d24p-56
#define DG_SQRT2 0x1.6a09e667f3bcdp+0

/* ln(m) for m in [sqrt(1/2), sqrt(2)], via ln(m) = 2 atanh(s), s = (m-1)/(m+1),
 * |s| <= 0.1716.  The first two series terms are carried in double-double,
 * the tail s^5/5 + ... + s^27/27 (relative size < 2e-4) in double. */
DG_HD dg_dd dd_log_reduced(double m) {
  dg_dd num = dd_make(m - 1.0, 0.0); /* exact by Sterbenz' lemma */
  dg_dd den = dd_two_sum(m, 1.0);
  dg_dd s = dd_div(num, den);
  dg_dd s2 = dd_mul(s, s);
  dg_dd s3 = dd_mul(s2, s);
  dg_dd t3 = dd_div(s3, dd_make(3.0, 0.0));
  double z = s2.hi;
  double poly = 1.0 / 27.0;
  poly = 1.0 / 25.0 + z * poly;
  poly = 1.0 / 23.0 + z * poly;
  poly = 1.0 / 21.0 + z * poly;
  poly = 1.0 / 19.0 + z * poly;
  poly = 1.0 / 17.0 + z * poly;
  poly = 1.0 / 15.0 + z * poly;
  poly = 1.0 / 13.0 + z * poly;
  poly = 1.0 / 11.0 + z * poly;
  poly = 1.0 / 9.0 + z * poly;
  poly = 1.0 / 7.0 + z * poly;
  poly = 1.0 / 5.0 + z * poly;
  double tail = (s3.hi * z) * poly;
  dg_dd r = dd_add(s, t3);
  r = dd_add_d(r, tail);
  r.hi *= 2.0;
  r.lo *= 2.0;
  return r;
}

/* Split a positive finite x into x = m * 2^e with m in [sqrt(1/2), sqrt(2)). */
DG_HD double dd_reduce_log_arg(double x, int* e_out) {
  uint64_t b = dg_bits(x);
  int eb = (int)((b >> 52) & 0x7ff);
  int adj = 0;
  if (eb == 0) { /* subnormal: scale into the normal range first */
    x = x * 0x1p54;
    b = dg_bits(x);
    eb = (int)((b >> 52) & 0x7ff);
    adj = -54;
  }
  int e = eb - 1023 + adj;
  double m = dg_from_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > DG_SQRT2) {
    m *= 0.5;
    e += 1;
  }
  *e_out = e;
  return m;
}

/* Natural log of positive finite x as double-double. */
DG_HD dg_dd dd_log(double x) {
  int e;
  double m = dd_reduce_log_arg(x, &e);
  dg_dd lm = dd_log_reduced(m);
  dg_dd le = dd_mul_d(dd_make(DG_LN2_HI, DG_LN2_LO), (double)e);
  return dd_add(le, lm);
}

/* log2 of positive finite x as double-double. */
DG_HD dg_dd dd_log2(double x) {
  int e;
  double m = dd_reduce_log_arg(x, &e);
  dg_dd lm = dd_log_reduced(m);
  dg_dd l2 = dd_mul(lm, dd_make(DG_LOG2E_HI, DG_LOG2E_LO));
  return dd_add_d(l2, (double)e);
}

/* expm1(r) for |r| <= 0.36: evaluate at r/8 by Taylor series, then undo the
 * scaling with three exact-form doublings u <- u * (2 + u). */
DG_HD dg_dd dd_expm1_core(dg_dd r) {
  dg_dd q = dd_make(r.hi * 0.125, r.lo * 0.125);
  double z = q.hi;
  dg_dd q2 = dd_mul(q, q);
  double poly = 1.0 / 39916800.0;   /* 1/11! */
  poly = 1.0 / 3628800.0 + z * poly; /* 1/10! */
  poly = 1.0 / 362880.0 + z * poly;
  poly = 1.0 / 40320.0 + z * poly;
  poly = 1.0 / 5040.0 + z * poly;
  poly = 1.0 / 720.0 + z * poly;
  poly = 1.0 / 120.0 + z * poly;
  poly = 1.0 / 24.0 + z * poly;
  poly = 1.0 / 6.0 + z * poly;
  double tail = (q2.hi * z) * poly; /* q^3/6 + q^4/24 + ... */
  dg_dd u = dd_add(q, dd_make(q2.hi * 0.5, q2.lo * 0.5));
  u = dd_add_d(u, tail);
  for (int i = 0; i < 3; ++i) {
    u = dd_mul(u, dd_add_d(u, 2.0));
  }
  return u;
}

/* 2^n for integer n in the normal exponent range. */
DG_HD double dg_pow2i(int n) {
  return dg_from_bits((uint64_t)(n + 1023) << 52);
}

/* Nearest integer of |y| < 2^31 (ties away from zero); exact conversions only. */
DG_HD double dg_round_int(double y) {
  int64_t n = (int64_t)(y >= 0.0 ? y + 0.5 : y - 0.5);
  return (double)n;
}

/* 2^y for double-double y with |y| < 1000. */
DG_HD dg_dd dd_exp2(dg_dd y) {
  double n = dg_round_int(y.hi);
  dg_dd f = dd_add_d(y, -n);
  dg_dd r = dd_mul(f, dd_make(DG_LN2_HI, DG_LN2_LO));
  dg_dd em1 = dd_expm1_core(r);
  dg_dd v = dd_add_d(em1, 1.0);
  int ni = (int)n;
  /* split the scaling so that each factor stays a normal number */
  int n1 = ni / 2;
  int n2 = ni - n1;
  v = dd_mul_d(v, dg_pow2i(n1));
  v = dd_mul_d(v, dg_pow2i(n2));
  return v;
}

/* ---------------------------------------------------------------------------
 * Public functions.
 * ------------------------------------------------------------------------ */

/* Natural logarithm, x > 0 finite. */
DG_HD double dg_log(double x) { return dd_log(x).hi; }

/* Base-2 logarithm, x > 0 finite. */
DG_HD double dg_log2(double x) { return dd_log2(x).hi; }

/* x^y for x > 0 finite; exact for y == 0 and y == 1 (as C99 pow is). */
DG_HD double dg_pow(double x, double y) {
  if (y == 0.0 || x == 1.0) return 1.0;
  if (y == 1.0) return x;
  dg_dd l = dd_log2(x);
  dg_dd p = dd_mul_d(l, y);
  return dd_exp2(p).hi;
}

/* tanh for x > 0. */
DG_HD double dg_tanh_pos(double x) {
  if (x >= 20.0) return 1.0;
  double y = 2.0 * x;
  dg_dd em1;
  if (y <= 0.34) {
    em1 = dd_expm1_core(dd_make(y, 0.0));
  } else {
    dg_dd yy = dd_mul_d(dd_make(DG_LOG2E_HI, DG_LOG2E_LO), y);
    em1 = dd_add_d(dd_exp2(yy), -1.0);
  }
  dg_dd t = dd_div(em1, dd_add_d(em1, 2.0));
  return t.hi;
}

/* Hyperbolic tangent: tanh(x) = expm1(2x) / (expm1(2x) + 2) for x > 0, odd
 * symmetry below zero.  Rounds to 1 above x = 19.06; +inf gives 1. */
DG_HD double dg_tanh(double x) {
  if (x == 0.0 || x != x) return x;
  if (x < 0.0) return -dg_tanh_pos(-x);
  return dg_tanh_pos(x);
}

/* Square root: IEEE-754 requires it correctly rounded, and both compilers
 * emit a correctly rounded sequence (sqrtsd on the host; on gfx950 the
 * v_rsq_f64 estimate refined by fma Newton steps with a final residual
 * correction), so host and device agree bit for bit; the device side is
 * checked against the host on sampled inputs by the tests. */
DG_HD double dg_sqrt(double x) { return __builtin_sqrt(x); }

#endif /* DG_MATH_H */
