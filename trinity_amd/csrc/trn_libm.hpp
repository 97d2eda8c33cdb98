// trn_libm.hpp -- FP64 exp, log and pow that return what the host C library
// the reference links (glibc 2.35, x86-64) returns.
//
// Why: the reference's metric (src/metric.cpp:129, pow) and its smoothing
// (src/numeric.cpp:139-140, 198-199 log/exp; :496 log; :304 pow(., 3)) feed
// these values into branches and eigen solves that amplify a one-ulp
// difference far beyond 1e-10 (a clamped eigenvalue of a nearly isotropic
// tensor turns it into a rotation of the eigenvector).  The device library's
// functions are faithful but differ from glibc in the last bit often enough
// to show at 16.8 M vertices.
//
// exp and pow restate the published algorithm glibc has shipped since 2.28
// (sysdeps/ieee754/dbl-64/e_exp.c, e_pow.c, contributed from Arm's
// optimized-routines, Szabolcs Nagy 2018): exp(x) = 2^(k/128) exp(r) with a
// 128-entry table and a degree-5 polynomial; pow(x, y) = exp(y log x) with
// log x in double-double from a 128-entry table of short reciprocals and a
// degree-7 polynomial, and the exp step fed the low part.  The x86-64 build
// selects the FMA variant of each, so the products that variant fuses are
// fused here too (__builtin_fma); this file is compiled without any other
// contraction.  The tables come from gen_libm_tables.py (their defining
// formulas, evaluated at 300 bits); the polynomial coefficients are the
// published ones.  Equality with the host libm is tested bit for bit on the
// CPU (tests/test_libm.py, millions of inputs over the ranges the path
// uses, plus the special values) and on the GPU.
//
// log is not glibc's algorithm (its table comes from a search that is not
// reproducible from its description): it is an accurate logarithm, relative
// error below about 2^-74 before the final rounding, built on the pow
// table: x = 2^k z, r = z/c - 1 exact, log x = k ln2 + log c + log1p(r)
// with the r, r^2/2 and r^3/3 terms carried in double-double and the Taylor
// series of log1p to r^12.  It is therefore correctly rounded except for
// about one input in 10^6, and equal to glibc's log wherever glibc's is
// correctly rounded (all but about 2 in 10^5 inputs, measured against
// 300-bit references).
//
// Usable from host and device code (the CPU tests compile this header with
// g++).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define TRN_LIBM_FN __host__ __device__ __forceinline__
#define TRN_LIBM_COLD static __host__ __device__ __noinline__
#else
#define TRN_LIBM_FN static inline
#define TRN_LIBM_COLD static __attribute__((noinline))
#endif

namespace trn {
namespace libm {

#include "trn_libm_tables.inc"

TRN_LIBM_FN uint64_t asu(const double d) { return __builtin_bit_cast(uint64_t, d); }
TRN_LIBM_FN double asd(const uint64_t u) { return __builtin_bit_cast(double, u); }
TRN_LIBM_FN uint32_t top12(const double x) { return (uint32_t)(asu(x) >> 52); }
TRN_LIBM_FN double fma_(const double a, const double b, const double c) {
  return __builtin_fma(a, b, c);
}

constexpr int kTableBits = 7;
constexpr int kN = 1 << kTableBits;
constexpr uint64_t kSignBias = 0x800ULL << kTableBits;

// exp reduction and polynomial (|r| <= ln2/256).
constexpr double kInvLn2N = 0x1.71547652b82fep0 * kN;
constexpr double kShift = 0x1.8p52;
constexpr double kNegLn2hiN = -0x1.62e42fefa0000p-8;
constexpr double kNegLn2loN = -0x1.cf79abc9e3b3ap-47;
constexpr double kC2 = 0x1.ffffffffffdbdp-2;
constexpr double kC3 = 0x1.555555555543cp-3;
constexpr double kC4 = 0x1.55555cf172b91p-5;
constexpr double kC5 = 0x1.1111167a4d017p-7;

// log reduction (shared by pow and log).
constexpr uint64_t kLogOff = 0x3fe6955500000000ULL;
constexpr double kLn2hi = 0x1.62e42fefa3800p-1;   // 42 bits: k*kLn2hi is exact
constexpr double kLn2lo = 0x1.ef35793c76730p-45;
// pow's log1p(r) - r + r^2/2 polynomial, coefficients scaled for its
// evaluation in powers of ar2 = -r^2/2.
constexpr double kA0 = -0x1p-1;
constexpr double kA1 = 0x1.555555555556p-2 * -2;
constexpr double kA2 = -0x1.0000000000006p-2 * -2;
constexpr double kA3 = 0x1.999999959554ep-3 * 4;
constexpr double kA4 = -0x1.555555529a47ap-3 * 4;
constexpr double kA5 = 0x1.2495b9b4845e9p-3 * -8;
constexpr double kA6 = -0x1.0002b8b263fc3p-3 * -8;

// Results outside the finite range, with the same values the C library
// returns (+-inf, +-0 after the underflowing product).
TRN_LIBM_FN double oflow(const uint32_t sign) {
  return (sign ? -0x1p769 : 0x1p769) * 0x1p769;
}
TRN_LIBM_FN double uflow(const uint32_t sign) {
  return (sign ? -0x1p-767 : 0x1p-767) * 0x1p-767;
}

// Scale by 2^(k/N) when k/N is outside the normal exponent range of the
// table word: |x| in [512, 1024).  Handles the subnormal results with one
// extra rounding step so that they round once.
TRN_LIBM_FN double exp_special(const double tmp, uint64_t sbits, const uint64_t ki) {
  if ((ki & 0x80000000ULL) == 0) {
    sbits -= 1009ULL << 52;
    const double scale = asd(sbits);
    return 0x1p1009 * fma_(scale, tmp, scale);
  }
  // k < 0: the product scale * tmp has two uses below, so the FMA build
  // rounds it separately (a product is fused only into its single sum).
  sbits += 1022ULL << 52;
  const double scale = asd(sbits);
  const double st = scale * tmp;
  double y = scale + st;
  double ay = y < 0 ? -y : y;
  if (ay < 1.0) {
    const double one = y < 0.0 ? -1.0 : 1.0;
    double lo = scale - y + st;
    const double hi = one + y;
    lo = one - hi + y + lo;
    y = (hi + lo) - one;
    if (y == 0) y = asd(sbits & 0x8000000000000000ULL);
  }
  return 0x1p-1022 * y;
}

// exp(x + xtail) * (sign_bias ? -1 : 1); xtail is the low part of pow's
// product (0 for exp itself).
TRN_LIBM_FN double exp_core(const double x, const double xtail, const uint32_t sign_bias,
                            const bool is_pow) {
  uint32_t abstop = top12(x) & 0x7ff;
  if (abstop - top12(0x1p-54) >= top12(512.0) - top12(0x1p-54)) {
    if (abstop - top12(0x1p-54) >= 0x80000000u) {
      const double one = 1.0 + x;
      return sign_bias ? -one : one;
    }
    if (abstop >= top12(1024.0)) {
      if (!is_pow) {
        if (asu(x) == 0xfff0000000000000ULL) return 0.0;   // -inf
        if (abstop >= top12(__builtin_inf())) return 1.0 + x;
      }
      return (asu(x) >> 63) ? uflow(sign_bias) : oflow(sign_bias);
    }
    abstop = 0;   // |x| in [512, 1024): scaled below
  }
  double kd = fma_(kInvLn2N, x, kShift);
  const uint64_t ki = asu(kd);
  kd -= kShift;
  double r = fma_(kd, kNegLn2loN, fma_(kd, kNegLn2hiN, x));
  r += xtail;
  const uint64_t idx = 2 * (ki % kN);
  const uint64_t top = (ki + (sign_bias ? kSignBias : 0)) << (52 - kTableBits);
  const double tail = asd(kExpTab[idx]);
  const uint64_t sbits = kExpTab[idx + 1] + top;
  const double r2 = r * r;
  const double tmp = fma_(r2 * r2, fma_(r, kC5, kC4), fma_(r2, fma_(r, kC3, kC2), tail + r));
  if (abstop == 0) return exp_special(tmp, sbits, ki);
  const double scale = asd(sbits);
  return fma_(scale, tmp, scale);
}

// exp for arguments outside the main range, out of line (cold).
TRN_LIBM_COLD double exp_rare(const double x) { return exp_core(x, 0.0, 0, false); }

// glibc exp: the main range 2^-54 <= |x| < 512 inline and without branches
// (the tables and arithmetic of exp_core), everything else out of line.
TRN_LIBM_FN double exp(const double x) {
  const uint32_t abstop = top12(x) & 0x7ff;
  double kd = fma_(kInvLn2N, x, kShift);
  const uint64_t ki = asu(kd);
  kd -= kShift;
  const double r = fma_(kd, kNegLn2loN, fma_(kd, kNegLn2hiN, x)) + 0.0;
  const uint32_t idx = 2 * ((uint32_t)ki % kN);
  const uint64_t top = ki << (52 - kTableBits);
  const double tail = asd(kExpTab[idx]);
  const uint64_t sbits = kExpTab[idx + 1] + top;
  const double r2 = r * r;
  const double tmp = fma_(r2 * r2, fma_(r, kC5, kC4), fma_(r2, fma_(r, kC3, kC2), tail + r));
  const double scale = asd(sbits);
  const double y = fma_(scale, tmp, scale);
  if (abstop - top12(0x1p-54) >= top12(512.0) - top12(0x1p-54)) return exp_rare(x);
  return y;
}

// log(x) = hi + *tail for a normal positive x given by its bits, relative
// error about 2^-68 (pow's inner logarithm).
TRN_LIBM_FN double log_pow_inline(const uint64_t ix, double* tail) {
  const uint64_t tmp = ix - kLogOff;
  const int i = (int)((tmp >> (52 - kTableBits)) % kN);
  const int k = (int)((int64_t)tmp >> 52);
  const uint64_t iz = ix - (tmp & (0xfffULL << 52));
  const double z = asd(iz);
  const double kd = (double)k;
  const double invc = kLogTab[i][0], logc = kLogTab[i][1], logctail = kLogTab[i][2];
  const double r = fma_(z, invc, -1.0);
  const double t1 = fma_(kd, kLn2hi, logc);
  const double t2 = t1 + r;
  const double lo1 = fma_(kd, kLn2lo, logctail);
  const double lo2 = t1 - t2 + r;
  const double ar = kA0 * r;
  const double ar2 = r * ar;
  const double ar3 = r * ar2;
  const double hi = t2 + ar2;
  const double lo3 = fma_(ar, r, -ar2);
  const double lo4 = t2 - hi + ar2;
  const double p = ar3 * fma_(ar2, fma_(ar2, fma_(r, kA6, kA5), fma_(r, kA4, kA3)),
                              fma_(r, kA2, kA1));
  const double lo = lo1 + lo2 + lo3 + lo4 + p;
  const double y = hi + lo;
  *tail = hi - y + lo;
  return y;
}

// 0: not an integer, 1: odd integer, 2: even integer.
TRN_LIBM_FN int checkint(const uint64_t iy) {
  const int e = (int)(iy >> 52 & 0x7ff);
  if (e < 0x3ff) return 0;
  if (e > 0x3ff + 52) return 2;
  if (iy & ((1ULL << (0x3ff + 52 - e)) - 1)) return 0;
  if (iy & (1ULL << (0x3ff + 52 - e))) return 1;
  return 2;
}

TRN_LIBM_FN bool zeroinfnan(const uint64_t i) {
  return 2 * i - 1 >= 2 * 0x7ff0000000000000ULL - 1;
}

// glibc pow.
TRN_LIBM_FN double pow(const double x, const double y) {
  uint32_t sign_bias = 0;
  uint64_t ix = asu(x);
  const uint64_t iy = asu(y);
  uint32_t topx = top12(x);
  const uint32_t topy = top12(y);
  if (topx - 0x001 >= 0x7ff - 0x001 || (topy & 0x7ff) - 0x3be >= 0x43e - 0x3be) {
    if (zeroinfnan(iy)) {
      if (2 * iy == 0) return 1.0;
      if (ix == asu(1.0)) return 1.0;
      if (2 * ix > 2 * 0x7ff0000000000000ULL || 2 * iy > 2 * 0x7ff0000000000000ULL)
        return x + y;
      if (2 * ix == 2 * asu(1.0)) return 1.0;
      if ((2 * ix < 2 * asu(1.0)) == !(iy >> 63)) return 0.0;
      return y * y;
    }
    if (zeroinfnan(ix)) {
      double x2 = x * x;
      if ((ix >> 63) && checkint(iy) == 1) x2 = -x2;
      return (iy >> 63) ? 1 / x2 : x2;
    }
    if (ix >> 63) {
      const int yint = checkint(iy);
      if (yint == 0) return (x - x) / (x - x);
      if (yint == 1) sign_bias = 1;
      ix &= 0x7fffffffffffffffULL;
      topx &= 0x7ff;
    }
    if ((topy & 0x7ff) - 0x3be >= 0x43e - 0x3be) {
      if (ix == asu(1.0)) return 1.0;
      if ((topy & 0x7ff) < 0x3be) return ix > asu(1.0) ? 1.0 + y : 1.0 - y;
      return (ix > asu(1.0)) == (topy < 0x800) ? oflow(0) : uflow(0);
    }
    if (topx == 0) {
      ix = asu(x * 0x1p52);
      ix &= 0x7fffffffffffffffULL;
      ix -= 52ULL << 52;
    }
  }
  double lo;
  const double hi = log_pow_inline(ix, &lo);
  const double ehi = y * hi;
  const double elo = fma_(y, lo, fma_(y, hi, -ehi));
  return exp_core(ehi, elo, sign_bias, true);
}

// Accurate log (see the file comment): x = 2^k z, r = z/c - 1 (exact),
// log x = (k ln2_hi + log_hi c) + r - r^2/2 + r^3/3 + [k ln2_lo + log_lo c
// + r^4 (-1/4 + r/5 - ... - r^8/12)], with the leading sums error-free.
TRN_LIBM_FN double log_main(uint64_t ix);

// log of zero, negative, subnormal, infinite and NaN arguments, out of line.
TRN_LIBM_COLD double log_rare(const double x) {
  const uint64_t ix = asu(x);
  const uint32_t top = (uint32_t)(ix >> 48);
  if (2 * ix == 0) return -__builtin_inf();
  if (ix == 0x7ff0000000000000ULL) return x;
  if ((top & 0x8000) || (top & 0x7ff0) == 0x7ff0) return (x - x) / (x - x);
  return log_main(asu(x * 0x1p52) - (52ULL << 52));   // subnormal: normalise
}

TRN_LIBM_FN double log(const double x) {
  const uint64_t ix = asu(x);
  const uint32_t top = (uint32_t)(ix >> 48);
  const double y = log_main(ix);
  if (top - 0x0010 >= 0x7ff0 - 0x0010) return log_rare(x);
  return y;
}

// The logarithm of a positive normal number given by its bits.
TRN_LIBM_FN double log_main(uint64_t ix) {
  const uint64_t tmp = ix - kLogOff;
  const int i = (int)((tmp >> (52 - kTableBits)) % kN);
  const int k = (int)((int64_t)tmp >> 52);
  const double z = asd(ix - (tmp & (0xfffULL << 52)));
  const double kd = (double)k;
  const double invc = kLogTab[i][0], logc = kLogTab[i][1], logctail = kLogTab[i][2];
  const double r = fma_(z, invc, -1.0);           // exact, |r| < 0x1.7p-8
  const double t1 = fma_(kd, kLn2hi, logc);       // exact
  const double lo1 = fma_(kd, kLn2lo, logctail);
  const double t2 = t1 + r;                       // |t1| >= |r| or t1 == 0
  const double lo2 = (t1 - t2) + r;
  const double ar = -0.5 * r;
  const double ar2 = r * ar;                      // -r^2/2 = ar2 + lo3
  const double lo3 = fma_(ar, r, -ar2);
  const double hi2 = t2 + ar2;                    // |t2| >= |ar2|
  const double lo4 = (t2 - hi2) + ar2;
  const double s = r * r;                         // r^2 = s + s_lo
  const double s_lo = fma_(r, r, -s);
  const double u = s * r;                         // r^3 = u + u_lo
  const double u_lo = fma_(s, r, -u) + s_lo * r;
  constexpr double kThirdHi = 0x1.5555555555555p-2;
  constexpr double kThirdLo = 0x1.5555555555555p-56;
  const double c = u * kThirdHi;                  // r^3/3 = c + c_lo
  const double c_lo = fma_(u, kThirdHi, -c) + (u * kThirdLo + u_lo * kThirdHi);
  const double hi3 = hi2 + c;                     // |hi2| >> |c|
  const double lo5 = (hi2 - hi3) + c;
  double q = -1.0 / 12;
  q = fma_(q, r, 1.0 / 11);
  q = fma_(q, r, -1.0 / 10);
  q = fma_(q, r, 1.0 / 9);
  q = fma_(q, r, -1.0 / 8);
  q = fma_(q, r, 1.0 / 7);
  q = fma_(q, r, -1.0 / 6);
  q = fma_(q, r, 1.0 / 5);
  q = fma_(q, r, -1.0 / 4);
  q = (s * s) * q;
  const double lo = ((((q + c_lo) + lo5) + lo4) + lo3) + (lo2 + lo1);
  return hi3 + lo;
}

// pow(x, 3.) as pow above (and the C library) returns it, mostly without
// evaluating pow.  x^3 = q + lo in double-double (x^2 = p + e and p x =
// q + f exactly with FMA, lo = f + e x, error below 2^-100 of x^3), so
// r = q + lo is the correctly rounded cube and d = x^3 - r its residual.
// pow is within 0.52 ulp of x^3: when |d| stays below 0.48 ulp of x^3's
// binade, the other neighbour of x^3 is more than 0.52 ulp away and pow
// returns r.  The rest (a cube within 0.02 ulp of a midpoint, about 4 in
// 100 arguments, and zero, subnormal or non-finite cubes) takes pow.
TRN_LIBM_FN double cube(const double x) {
  const double p = x * x;
  const double e = fma_(x, x, -p);
  const double q = p * x;
  const double f = fma_(p, x, -q);
  const double lo = f + e * x;
  const double r = q + lo;
  const double d = (q - r) + lo;
  const uint64_t er = asu(r) & 0x7ff0000000000000ULL;
  if (er <= (64ULL << 52) || er == 0x7ff0000000000000ULL) return pow(x, 3.);
  // ulp of x^3: r's, or half of it when r is a power of two and x^3 < |r|
  double u = asd(er) * 0x1p-52;
  if ((asu(r) & 0x000fffffffffffffULL) == 0 && (r > 0 ? d < 0 : d > 0)) u *= 0.5;
  const double ad = d < 0 ? -d : d;
  if (ad < 0.48 * u) return r;
  return pow(x, 3.);
}

// ---- functions of the adjoint's own derivatives ----------------------------
// Not on the reference's path: the divided differences of log and exp that
// the adjoint's backward needs (the gradient moves by about 3e-10 of its
// largest entry per ulp of them).  The CPU reference adjoint calls these same
// functions, compiled from this header, so both sides round identically.

// log(1 + r) for |r| <= 2^-10: r - r^2/2 in double-double plus the Taylor
// tail to r^8, relative error below 2^-70.
TRN_LIBM_FN double log1p_small(const double r) {
  const double ar = -0.5 * r;
  const double ar2 = r * ar;                 // -r^2/2 = ar2 + lo3 exactly
  const double lo3 = fma_(ar, r, -ar2);
  const double hi = r + ar2;                 // |r| >= |ar2|
  const double lo4 = (r - hi) + ar2;
  double q = -1.0 / 8;
  q = fma_(q, r, 1.0 / 7);
  q = fma_(q, r, -1.0 / 6);
  q = fma_(q, r, 1.0 / 5);
  q = fma_(q, r, -1.0 / 4);
  q = fma_(q, r, 1.0 / 3);
  q = (r * r) * r * q;
  return hi + ((q + lo3) + lo4);
}

// exp(d) - 1 within about half an ulp more than exact rounding: the Taylor
// series (d + d^2/2 in double-double) for |d| < 1/16; for |d| < 0.7,
// exp(d) = scale (1 + tmp) with scale near 1, so scale - 1 is exact and one
// fused step gives the result; beyond, exp(d) - 1.
TRN_LIBM_FN double expm1(const double d) {
  const double ad = d < 0 ? -d : d;
  if (!(ad < 0.7)) return exp(d) - 1.0;
  if (ad < 0x1p-4) {
    const double h = 0.5 * d;
    const double d2 = d * h;                 // d^2/2 = d2 + d2_lo exactly
    const double d2_lo = fma_(d, h, -d2);
    const double s = d + d2;                 // |d| >= d^2/2
    const double s_lo = (d - s) + d2;
    double q = 1.0 / 479001600;              // 1/12!
    q = fma_(q, d, 1.0 / 39916800);
    q = fma_(q, d, 1.0 / 3628800);
    q = fma_(q, d, 1.0 / 362880);
    q = fma_(q, d, 1.0 / 40320);
    q = fma_(q, d, 1.0 / 5040);
    q = fma_(q, d, 1.0 / 720);
    q = fma_(q, d, 1.0 / 120);
    q = fma_(q, d, 1.0 / 24);
    q = fma_(q, d, 1.0 / 6);
    q = (d * d) * d * q;
    return s + ((q + d2_lo) + s_lo);
  }
  double kd = fma_(kInvLn2N, d, kShift);
  const uint64_t ki = asu(kd);
  kd -= kShift;
  const double r = fma_(kd, kNegLn2loN, fma_(kd, kNegLn2hiN, d));
  const uint64_t idx = 2 * (ki % kN);
  const uint64_t top = ki << (52 - kTableBits);
  const double tail = asd(kExpTab[idx]);
  const double scale = asd(kExpTab[idx + 1] + top);
  const double r2 = r * r;
  const double tmp = fma_(r2 * r2, fma_(r, kC5, kC4), fma_(r2, fma_(r, kC3, kC2), tail + r));
  return fma_(scale, tmp, scale - 1.0);
}

}  // namespace libm
}  // namespace trn
