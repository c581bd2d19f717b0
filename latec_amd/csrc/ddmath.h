// ddmath.h -- double-double exp / log2 used to build the Laplace-mixture
// tables on the GPU.
//
// The tables (latec/rans.py:247-267, 676-722) are sensitive to the last bit of
// every exp(): mirror-image bins get nearly equal fractional parts and the
// reference breaks those ties by a stable sort.  These routines evaluate exp
// and log2 in double-double arithmetic (relative error far below 2^-90) and
// round once, i.e. they return the correctly rounded double.  Written from the
// mathematics (Cody-Waite reduction by ln2, a halving/squaring step and a
// Taylor series; log2 by one Newton step on exp), no library constants.
//
// Compiled for both the host (tests) and gfx950 (table kernels).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define DD_FN __host__ __device__ inline
#else
#define DD_FN static inline
#endif

#include <math.h>

struct dd {
  double hi, lo;
};

DD_FN dd dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return dd{s, e};
}
DD_FN dd dd_fast_two_sum(double a, double b) {
  double s = a + b;
  return dd{s, b - (s - a)};
}
DD_FN dd dd_two_prod(double a, double b) {
  double p = a * b;
  return dd{p, fma(a, b, -p)};
}
DD_FN dd dd_add(dd a, dd b) {
  dd s = dd_two_sum(a.hi, b.hi);
  dd t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}
DD_FN dd dd_mul(dd a, dd b) {
  dd p = dd_two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return dd_fast_two_sum(p.hi, p.lo);
}
DD_FN dd dd_mul_d(dd a, double b) {
  dd p = dd_two_prod(a.hi, b);
  p.lo += a.lo * b;
  return dd_fast_two_sum(p.hi, p.lo);
}
DD_FN dd dd_div_d(dd a, double b) {
  double q1 = a.hi / b;
  dd p = dd_two_prod(q1, b);
  dd r = dd_add(a, dd{-p.hi, -p.lo});
  double q2 = r.hi / b;
  return dd_fast_two_sum(q1, q2);
}

// ln2 as a triple of doubles: LN2_0 + LN2_1 + LN2_2 (about 160 bits), the
// binary expansion of ln 2 split into 53-bit pieces.
#define DD_LN2_0 0x1.62e42fefa39efp-1
#define DD_LN2_1 0x1.abc9e3b39803fp-56
#define DD_LN2_2 0x1.7b57a079a1934p-111

// exp(x) as a double-double, |x| <= 709.
DD_FN dd dd_exp(double x) {
  double kd = rint(x * 0x1.71547652b82fep0);  // x / ln2, any rounding is fine
  // r = x - k ln2 in double-double (k*LN2_0 exact via two_prod)
  dd p0 = dd_two_prod(kd, DD_LN2_0);
  dd p1 = dd_two_prod(kd, DD_LN2_1);
  dd r = dd_add(dd{x, 0.0}, dd{-p0.hi, -p0.lo});
  r = dd_add(r, dd{-p1.hi, -p1.lo});
  r = dd_add(r, dd{-kd * DD_LN2_2, 0.0});
  // exp(r) = exp(r / 2^10)^(2^10); Taylor to degree 10 at |t| < 3.4e-4
  dd t{ldexp(r.hi, -10), ldexp(r.lo, -10)};
  dd sum{1.0, 0.0}, term{1.0, 0.0};
  for (int i = 1; i <= 10; i++) {
    term = dd_div_d(dd_mul(term, t), (double)i);
    sum = dd_add(sum, term);
  }
  for (int i = 0; i < 10; i++) sum = dd_mul(sum, sum);
  int k = (int)kd;
  return dd{ldexp(sum.hi, k), ldexp(sum.lo, k)};
}

// Correctly rounded exp for the table's argument range.
DD_FN double cr_exp(double x) {
  dd e = dd_exp(x);
  return e.hi + e.lo;
}

// log(a) for a > 0 (normal): y0 = log(a) in double, then one Newton step
// y = y0 + (a - exp(y0)) / exp(y0) in double-double.
DD_FN dd dd_log(double a) {
  int ex;
  double m = frexp(a, &ex);  // a = m 2^ex, m in [0.5, 1)
  double y0 = log(m);
  dd e = dd_exp(y0);
  dd diff = dd_add(dd{m, 0.0}, dd{-e.hi, -e.lo});
  double corr = (diff.hi + diff.lo) / (e.hi + e.lo);
  dd y = dd_two_sum(y0, corr);
  // + ex * ln2
  dd l0 = dd_two_prod((double)ex, DD_LN2_0);
  dd l1 = dd_two_prod((double)ex, DD_LN2_1);
  y = dd_add(y, l0);
  y = dd_add(y, l1);
  return y;
}

// Correctly rounded log2(a) (up to one ulp on the rare exact-midpoint path).
DD_FN double cr_log2(double a) {
  dd l = dd_log(a);
  // divide by ln2 (double-double)
  dd ln2{DD_LN2_0, DD_LN2_1};
  double q1 = l.hi / ln2.hi;
  dd p = dd_mul_d(ln2, q1);
  dd r = dd_add(l, dd{-p.hi, -p.lo});
  double q2 = (r.hi + r.lo) / ln2.hi;
  return q1 + q2;
}

// ---------------------------------------------------------------------------
// latec_exp: the table-driven exp of ARM's optimized-routines (MIT licence;
// glibc's double exp since 2.28), evaluated with the fused multiply-adds of
// its FMA build, which is the variant x86-64 glibc selects on every CPU with
// FMA.  Table: exp_table.h (generated from the mathematics).  Valid for
// finite x with exp(x) normal (|x| < 708); the Laplace tables use [-700, 9].
// ---------------------------------------------------------------------------
#include "exp_table.h"

DD_FN uint64_t dd_asu64(double x) {
  union { double d; uint64_t u; } v;
  v.d = x;
  return v.u;
}
DD_FN double dd_asf64(uint64_t x) {
  union { double d; uint64_t u; } v;
  v.u = x;
  return v.d;
}

DD_FN double latec_exp(double x) {
  const double InvLn2N = 0x1.71547652b82fep0 * 128.0;
  const double Shift = 0x1.8p52;
  const double NegLn2hiN = -0x1.62e42fefa0000p-8;
  const double NegLn2loN = -0x1.cf79abc9e3b3ap-47;
  const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3;
  const double C4 = 0x1.55555cf172b91p-5, C5 = 0x1.1111167a4d017p-7;
  double kd = fma(InvLn2N, x, Shift);
  uint64_t ki = dd_asu64(kd);
  kd -= Shift;
  double r = fma(kd, NegLn2loN, fma(kd, NegLn2hiN, x));
  uint64_t idx = 2 * (ki % 128);
  uint64_t top = ki << 45;
  double tail = dd_asf64(latec_exp_tab[idx]);
  uint64_t sbits = latec_exp_tab[idx + 1] + top;
  double r2 = r * r;
  double tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), tail + r));
  uint32_t abstop = (uint32_t)(dd_asu64(x) >> 52) & 0x7ff;
  if (abstop >= 0x408) {  // |x| >= 512: rescale to keep the exponent in range
    if ((ki & 0x80000000u) == 0) {
      sbits -= 1009ull << 52;
      double sc = dd_asf64(sbits);
      return 0x1p1009 * fma(sc, tmp, sc);
    }
    sbits += 1022ull << 52;
    double sc = dd_asf64(sbits);
    double y = sc * tmp;  // this branch is not fused in the FMA build
    return 0x1p-1022 * (sc + y);
  }
  double scale = dd_asf64(sbits);
  return fma(scale, tmp, scale);
}
