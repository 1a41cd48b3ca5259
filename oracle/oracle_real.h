/*
 * oracle_real.h -- scalar type of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker of the HIP tracker; it
 * is never linked into libbmx and never runs on the product path.
 *
 * Plain build (C):      real == double, math from <math.h>.
 * Counting build (C++): -DORACLE_COUNT_FLOPS makes real a class that counts
 *   FP64 operations with the convention add/sub/mul/div/sqrt = 1 and every
 *   transcendental call (sin, cos, tan, sinh, cosh, asin, atan2) = 1.  The
 *   oracle source is written in the common subset of C and C++ so the same
 *   text is compiled both ways; the count per particle-element-turn feeds the
 *   algorithmic-FLOP figure of the roofline.
 */
#ifndef ORACLE_REAL_H
#define ORACLE_REAL_H

#ifndef ORACLE_COUNT_FLOPS

#include <math.h>
typedef double real;
#define RV(x) (x)
#define FLOP_COUNT_ADD(n) ((void)0)

#else /* counting build, C++ only */

#include <cmath>
extern "C" {
extern long long oracle_flop_counter;
}
#define FLOP_COUNT_ADD(n) (oracle_flop_counter += (n))

struct real {
  double v;
  real() : v(0) {}
  real(double x) : v(x) {}
  real &operator+=(const real &o) { v += o.v; FLOP_COUNT_ADD(1); return *this; }
  real &operator-=(const real &o) { v -= o.v; FLOP_COUNT_ADD(1); return *this; }
  real &operator*=(const real &o) { v *= o.v; FLOP_COUNT_ADD(1); return *this; }
  real &operator/=(const real &o) { v /= o.v; FLOP_COUNT_ADD(1); return *this; }
};
#define RV(x) ((x).v)
inline real operator+(real a, real b) { FLOP_COUNT_ADD(1); return real(a.v + b.v); }
inline real operator-(real a, real b) { FLOP_COUNT_ADD(1); return real(a.v - b.v); }
inline real operator*(real a, real b) { FLOP_COUNT_ADD(1); return real(a.v * b.v); }
inline real operator/(real a, real b) { FLOP_COUNT_ADD(1); return real(a.v / b.v); }
inline real operator-(real a) { return real(-a.v); }
inline bool operator<(real a, real b) { return a.v < b.v; }
inline bool operator>(real a, real b) { return a.v > b.v; }
inline bool operator<=(real a, real b) { return a.v <= b.v; }
inline bool operator>=(real a, real b) { return a.v >= b.v; }
inline bool operator==(real a, real b) { return a.v == b.v; }
inline bool operator!=(real a, real b) { return a.v != b.v; }
inline real sqrt(real a) { FLOP_COUNT_ADD(1); return real(std::sqrt(a.v)); }
inline real sin(real a) { FLOP_COUNT_ADD(1); return real(std::sin(a.v)); }
inline real cos(real a) { FLOP_COUNT_ADD(1); return real(std::cos(a.v)); }
inline real tan(real a) { FLOP_COUNT_ADD(1); return real(std::tan(a.v)); }
inline real sinh(real a) { FLOP_COUNT_ADD(1); return real(std::sinh(a.v)); }
inline real cosh(real a) { FLOP_COUNT_ADD(1); return real(std::cosh(a.v)); }
inline real asin(real a) { FLOP_COUNT_ADD(1); return real(std::asin(a.v)); }
inline real atan2(real a, real b) { FLOP_COUNT_ADD(1); return real(std::atan2(a.v, b.v)); }
inline real fabs(real a) { return real(std::fabs(a.v)); }
inline real floor(real a) { return real(std::floor(a.v)); }
inline real fmod(real a, real b) { return real(std::fmod(a.v, b.v)); }
inline real pow(real a, int n) { FLOP_COUNT_ADD(n > 1 ? n - 1 : 0); return real(std::pow(a.v, n)); }

#endif

#endif /* ORACLE_REAL_H */
