// cr_sincos.hpp — sin and cos in binary64, correctly rounded, for the device
// side of expmap's large-angle branch (dynamics.cpp:26-102 calls the host
// libm's sin and cos there; the device libm's results can differ from it in
// the last bit, the host's are correctly rounded for these arguments).
//
// Method: double-double arithmetic (Dekker's and Knuth's error-free sum and
// product, the product through a fused multiply-add), Cody-Waite reduction
// by pi/2 split into 33 + 33 + 53 bits (k pi/2 for |k| < 2^20 subtracted
// with the first two products exact), and the Taylor series of sin and cos
// on |r| <= pi/4 to the term of degree 29 and 28 (truncation below 2^-110
// relative).  The double-double result carries about 2^-100 relative error,
// so rounding its head gives the correctly rounded value except for
// arguments within ~2^-50 of a multiple of pi/2 (where the third part's
// truncation, 1e-37 k, matters) and the astronomically rare cases whose
// exact value lies within 2^-100 of a rounding boundary.  Arguments of 2^20
// and more (or non-finite) return the device libm's value.  The constants are
// pi/2 and the reciprocal factorials, computed exactly in rational
// arithmetic and rounded to binary64 pairs.
#pragma once
#include <cmath>

#include "exact_math.hpp"

namespace dh {

struct DD {
  double hi, lo;
};

DH_HD DD dd_two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

DH_HD DD dd_quick(double a, double b) {  // |a| >= |b|
  const double s = a + b;
  return {s, b - (s - a)};
}

DH_HD DD dd_mul(const DD& x, const DD& y) {
  using std::fma;
  const double p = x.hi * y.hi;
  double e = fma(x.hi, y.hi, -p);
  e += x.hi * y.lo + x.lo * y.hi;
  return dd_quick(p, e);
}

DH_HD DD dd_add(const DD& x, const DD& y) {
  DD s = dd_two_sum(x.hi, y.hi);
  s.lo += x.lo + y.lo;
  return dd_quick(s.hi, s.lo);
}

// 1/(2k+1)! and 1/(2k)!, (head, tail)
constexpr double kInvOddFact[15][2] = {
    {0x1.0000000000000p+0, 0.0},
    {0x1.5555555555555p-3, 0x1.5555555555555p-57},
    {0x1.1111111111111p-7, 0x1.1111111111111p-63},
    {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73},
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
    {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80},
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87},
    {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97},
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103},
    {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112},
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120},
    {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130},
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139},
    {0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149},
    {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157}};
constexpr double kInvEvenFact[15][2] = {
    {0x1.0000000000000p+0, 0.0},
    {0x1.0000000000000p-1, 0.0},
    {0x1.5555555555555p-5, 0x1.5555555555555p-59},
    {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65},
    {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
    {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76},
    {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83},
    {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92},
    {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101},
    {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107},
    {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120},
    {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124},
    {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135},
    {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143},
    {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153}};

// sum_k (-1)^k z^k c[k] by Horner in double-double
DH_HD DD dd_series(const DD& z, const double (*c)[2]) {
  DD s = {c[14][0], c[14][1]};
#pragma unroll 1
  for (int k = 13; k >= 0; --k) {
    s = dd_mul(s, z);
    s = {-s.hi, -s.lo};
    s = dd_add(s, DD{c[k][0], c[k][1]});
  }
  return s;
}

// sin(x) (want_cos false) or cos(x), correctly rounded as described above.
DH_HD double cr_sincos(double x, bool want_cos) {
  using std::fabs;
  if (!(fabs(x) < 0x1p20)) return want_cos ? std::cos(x) : std::sin(x);
  const double kPio2a = 0x1.921fb54400000p+0, kPio2b = 0x1.0b4611a600000p-34,
               kPio2c = 0x1.3198a2e037073p-69, kTwoOverPi = 0x1.45f306dc9c883p-1;
  const double k = std::nearbyint(x * kTwoOverPi);
  // r = x - k pi/2 in double-double: the first two products are exact (33-bit
  // parts, |k| < 2^20), the third is split by a fused multiply-add
  const double t1 = x - k * kPio2a;
  const DD s = dd_two_sum(t1, -(k * kPio2b));
  const double ph = k * kPio2c;
  const double pl = std::fma(k, kPio2c, -ph);
  DD r = dd_add(s, DD{-ph, -pl});
  const DD z = dd_mul(r, r);
  const DD sr = dd_mul(r, dd_series(z, kInvOddFact));
  const DD cr = dd_series(z, kInvEvenFact);
  // quadrant: sin(r + q pi/2), cos(r + q pi/2)
  const long long q = static_cast<long long>(k) & 3;
  const long long qq = want_cos ? ((q + 1) & 3) : q;
  const DD v = (qq == 0) ? sr : (qq == 1) ? cr : (qq == 2) ? DD{-sr.hi, -sr.lo} : DD{-cr.hi, -cr.lo};
  return v.hi;
}

}  // namespace dh
