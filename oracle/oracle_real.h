/*
 * oracle_real.h -- scalar type of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker of the HIP tracker; it
 * is never linked into libbmx and never runs on the product path.
 *
 * Plain build (C):      real == double, math from <math.h>.
 * Counting build (C++): -DORACLE_COUNT_FLOPS makes real a class that counts
 *   FP64 operations with the convention add/sub/mul/div/sqrt = 1 and every
 *   transcendental call (sin, cos, tan, sinh, cosh, asin, atan2) = 1, on
 *   values that depend on the particle's coordinates.  The
 *   oracle source is written in the common subset of C and C++ so the same
 *   text is compiled both ways; the count per particle-element-turn feeds the
 *   algorithmic-FLOP figure of the roofline.
 */
#ifndef ORACLE_REAL_H
#define ORACLE_REAL_H

#if defined(ORACLE_DUAL)

/* Forward-mode dual numbers: a value and its six partial derivatives with
 * respect to the initial (x, px, y, py, z, pz).  Used to form the exact
 * one-turn Jacobian of the oracle's map for the symplecticity check. */
#include <cmath>
#define FLOP_COUNT_ADD(n) ((void)0)
struct real {
  double v;
  double d[6];
  real() : v(0) { for (int i = 0; i < 6; i++) d[i] = 0; }
  real(double x) : v(x) { for (int i = 0; i < 6; i++) d[i] = 0; }
  real &operator+=(const real &o) { v += o.v; for (int i = 0; i < 6; i++) d[i] += o.d[i]; return *this; }
  real &operator-=(const real &o) { v -= o.v; for (int i = 0; i < 6; i++) d[i] -= o.d[i]; return *this; }
  real &operator*=(const real &o) { for (int i = 0; i < 6; i++) d[i] = d[i] * o.v + v * o.d[i]; v *= o.v; return *this; }
  real &operator/=(const real &o) { for (int i = 0; i < 6; i++) d[i] = (d[i] * o.v - v * o.d[i]) / (o.v * o.v); v /= o.v; return *this; }
};
#define RV(x) ((x).v)
#define PDEP(x) real((double)(x))
inline real chain(double v, double dv, const real &a) { real r(v); for (int i = 0; i < 6; i++) r.d[i] = dv * a.d[i]; return r; }
inline real operator+(real a, real b) { a += b; return a; }
inline real operator-(real a, real b) { a -= b; return a; }
inline real operator*(real a, real b) { a *= b; return a; }
inline real operator/(real a, real b) { a /= b; return a; }
inline real operator-(real a) { real r(-a.v); for (int i = 0; i < 6; i++) r.d[i] = -a.d[i]; return r; }
inline bool operator<(real a, real b) { return a.v < b.v; }
inline bool operator>(real a, real b) { return a.v > b.v; }
inline bool operator<=(real a, real b) { return a.v <= b.v; }
inline bool operator>=(real a, real b) { return a.v >= b.v; }
inline bool operator==(real a, real b) { return a.v == b.v; }
inline bool operator!=(real a, real b) { return a.v != b.v; }
inline real sqrt(real a) { double s = std::sqrt(a.v); return chain(s, 0.5 / s, a); }
inline real sin(real a) { return chain(std::sin(a.v), std::cos(a.v), a); }
inline real cos(real a) { return chain(std::cos(a.v), -std::sin(a.v), a); }
inline real tan(real a) { double t = std::tan(a.v); return chain(t, 1 + t * t, a); }
inline real sinh(real a) { return chain(std::sinh(a.v), std::cosh(a.v), a); }
inline real cosh(real a) { return chain(std::cosh(a.v), std::sinh(a.v), a); }
inline real asin(real a) { return chain(std::asin(a.v), 1.0 / std::sqrt(1 - a.v * a.v), a); }
inline real atan2(real a, real b) {
  double den = a.v * a.v + b.v * b.v;
  real r(std::atan2(a.v, b.v));
  for (int i = 0; i < 6; i++) r.d[i] = (b.v * a.d[i] - a.v * b.d[i]) / den;
  return r;
}
inline real fabs(real a) { return a.v < 0 ? -a : a; }
inline real floor(real a) { return real(std::floor(a.v)); }
inline real fmod(real a, real b) { real r = a; r.v = std::fmod(a.v, b.v); return r; }

#elif !defined(ORACLE_COUNT_FLOPS)

#include <math.h>
typedef double real;
#define RV(x) (x)
#define PDEP(x) (x)
#define FLOP_COUNT_ADD(n) ((void)0)

#else /* counting build, C++ only */

#include <cmath>
extern "C" {
extern long long oracle_flop_counter;
}
#define FLOP_COUNT_ADD(n) (oracle_flop_counter += (n))

/* A value remembers whether it depends on particle coordinates; only such
 * operations are counted, so per-element constants that the reference
 * recomputes for every particle (multipole bookkeeping, tilt sines, edge
 * coefficients) are excluded -- the count is of the hoisted algorithm. */
struct real {
  double v;
  bool d;
  real() : v(0), d(false) {}
  real(double x) : v(x), d(false) {}
  real(double x, bool dep) : v(x), d(dep) {}
  real &operator+=(const real &o) { v += o.v; d = d || o.d; if (d) FLOP_COUNT_ADD(1); return *this; }
  real &operator-=(const real &o) { v -= o.v; d = d || o.d; if (d) FLOP_COUNT_ADD(1); return *this; }
  real &operator*=(const real &o) { v *= o.v; d = d || o.d; if (d) FLOP_COUNT_ADD(1); return *this; }
  real &operator/=(const real &o) { v /= o.v; d = d || o.d; if (d) FLOP_COUNT_ADD(1); return *this; }
};
#define RV(x) ((x).v)
#define PDEP(x) real((double)(x), true)
inline real op2(double v, bool d) { if (d) FLOP_COUNT_ADD(1); return real(v, d); }
inline real operator+(real a, real b) { return op2(a.v + b.v, a.d || b.d); }
inline real operator-(real a, real b) { return op2(a.v - b.v, a.d || b.d); }
inline real operator*(real a, real b) { return op2(a.v * b.v, a.d || b.d); }
inline real operator/(real a, real b) { return op2(a.v / b.v, a.d || b.d); }
inline real operator-(real a) { return real(-a.v, a.d); }
inline bool operator<(real a, real b) { return a.v < b.v; }
inline bool operator>(real a, real b) { return a.v > b.v; }
inline bool operator<=(real a, real b) { return a.v <= b.v; }
inline bool operator>=(real a, real b) { return a.v >= b.v; }
inline bool operator==(real a, real b) { return a.v == b.v; }
inline bool operator!=(real a, real b) { return a.v != b.v; }
inline real sqrt(real a) { return op2(std::sqrt(a.v), a.d); }
inline real sin(real a) { return op2(std::sin(a.v), a.d); }
inline real cos(real a) { return op2(std::cos(a.v), a.d); }
inline real tan(real a) { return op2(std::tan(a.v), a.d); }
inline real sinh(real a) { return op2(std::sinh(a.v), a.d); }
inline real cosh(real a) { return op2(std::cosh(a.v), a.d); }
inline real asin(real a) { return op2(std::asin(a.v), a.d); }
inline real atan2(real a, real b) { return op2(std::atan2(a.v, b.v), a.d || b.d); }
inline real fabs(real a) { return real(std::fabs(a.v), a.d); }
inline real floor(real a) { return real(std::floor(a.v), a.d); }
inline real fmod(real a, real b) { return real(std::fmod(a.v, b.v), a.d || b.d); }

#endif

#endif /* ORACLE_REAL_H */
