/* legacy_expf.h -- expf as the reference's C renderer gets it from the C
 * library on x86-64 (c_api/src/lumen_c.c:514, glibc with FMA), restated so
 * that the GPU reproduces its bits.
 *
 * The published table-driven method (Szabolcs Nagy's expf for the Arm
 * optimized-routines, adopted by glibc 2.27+): with N = 32,
 *   z = x N / ln2 in double, k = round-to-nearest-even(z), r = z - k (one
 *   fused multiply-add from x, as the library's FMA build computes it),
 *   exp(x) = 2^(k / N) 2^(r / N) ~ s (c0 r^3 + c1 r^2 + c2 r + 1),
 * s = 2^((k mod N) / N) (a table of 32 doubles) scaled by 2^(k div N) in
 * its exponent field, the cubic a minimax fit of 2^(r/N) on |r| <= 1/2,
 * evaluated with fused multiply-adds, and one final rounding to float.  Every step is an IEEE double operation
 * (or an exact integer step), so host and device builds return the same
 * bits; it equals glibc's expf at every float x <= 0 (checked exhaustively
 * by tests/test_legacy_oracle.py).  Inputs above 0 are not used by the
 * legacy renderer (its exponent is clamped to <= 0); they are handled for
 * completeness up to the overflow bound.
 *
 * Plain C/C++: LEGACY_HD marks functions for both host and device builds. */
#ifndef LUMEN_LEGACY_EXPF_H
#define LUMEN_LEGACY_EXPF_H

#include <stdint.h>
#include <string.h>

#ifndef LEGACY_HD
#define LEGACY_HD static inline
#endif

/* bits of 2^(i/32) rounded to the nearest double, minus i << 47 (so that
 * adding k << 47 for k = 32 q + i gives 2^(q + i/32)) */
#define LUMEN_LEGACY_EXP2_TABLE                                                                \
  {0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull, \
   0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull, \
   0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull, \
   0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull, \
   0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull, \
   0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull, \
   0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull, \
   0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull}

LEGACY_HD float legacy_expf(float x) {
  const uint64_t table[32] = LUMEN_LEGACY_EXP2_TABLE;
  uint32_t ux;
  memcpy(&ux, &x, 4);
  const uint32_t abstop = (ux >> 20) & 0x7ffu;
  if (abstop >= 0x42bu) { /* |x| >= 88, inf or nan */
    if (ux == 0xff800000u) return 0.0f;       /* -inf */
    if (abstop >= 0x7f8u) return x + x;       /* nan, +inf */
    if (x > 0x1.62e42ep6f) return x * 0x1p127f * 0x1p127f;  /* overflow: +inf */
    if (x < -0x1.9fe368p6f) return 0.0f;      /* underflow */
  }
  const double xd = (double)x;
  const double z = 0x1.71547652b82fep+0 * 32.0 * xd; /* x N / ln2 */
  double kd = z + 0x1.8p52;                          /* round to nearest even */
  uint64_t ki;
  memcpy(&ki, &kd, 8);
  kd -= 0x1.8p52;
  const double r = __builtin_fma(0x1.71547652b82fep+0 * 32.0, xd, -kd); /* z - k, unrounded product */
  uint64_t t = table[ki % 32u] + (ki << 47);
  double s;
  memcpy(&s, &t, 8);
  const double c0 = 0x1.c6af84b912394p-5 / (32.0 * 32.0 * 32.0);
  const double c1 = 0x1.ebfce50fac4f3p-3 / (32.0 * 32.0);
  const double c2 = 0x1.62e42ff0c52d6p-1 / 32.0;
  const double p = __builtin_fma(c0, r, c1);
  const double r2 = r * r;
  double y = __builtin_fma(c2, r, 1.0);
  y = __builtin_fma(p, r2, y);
  y = y * s;
  return (float)y;
}

#endif
