// bmx_tangent.hip -- exact transfer matrices on the device.
//
// Bmad forms an element's 6x6 transfer matrix alongside the tracking, in the
// make_matrix branches of the track_a_* routines (e.g.
// bmad/low_level/track_a_drift.f90:94-114, collected by make_mat6,
// bmad/code/make_mat6.f90:23): hand-written derivatives of each map.  Here the
// derivative comes from the tracking code itself, run in forward-mode
// automatic differentiation: bmx_ops.h is compiled a second time with its
// particle arithmetic (`double`) redefined as a dual number carrying the six
// partial derivatives with respect to the initial coordinates.  The chain
// over elements and turns is then exact to rounding, and every branch the
// value takes (series or exact forms, the bend's common or general path) is
// the one the tracking kernels take, so the matrix is the Jacobian of the
// map libbmx tracks.  One particle per lane, the op program read straight
// from global memory; this is not a throughput path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/bmx.h"
#include "bmx_internal.h"
#include "bmx_tangent.h"

namespace bmx {
namespace tangent {

// value and partial derivatives d value / d (x, px, y, py, z, pz) at entry
struct Dual {
  double v;
  double d[6];
  Dual() = default;
  __host__ __device__ constexpr Dual(double x) : v(x), d{0.0, 0.0, 0.0, 0.0, 0.0, 0.0} {}
};

__device__ __forceinline__ Dual scaled(const Dual& a, double val, double f) {  // val, f * a'
  Dual r;
  r.v = val;
  for (int k = 0; k < 6; k++) r.d[k] = f * a.d[k];
  return r;
}
__device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v + b.v;
  for (int k = 0; k < 6; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v - b.v;
  for (int k = 0; k < 6; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator-(const Dual& a) { return scaled(a, -a.v, -1.0); }
__device__ __forceinline__ Dual operator+(const Dual& a) { return a; }
__device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v * b.v;
  for (int k = 0; k < 6; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v / b.v;
  for (int k = 0; k < 6; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
__device__ __forceinline__ Dual& operator+=(Dual& a, const Dual& b) { return a = a + b; }
__device__ __forceinline__ Dual& operator-=(Dual& a, const Dual& b) { return a = a - b; }
__device__ __forceinline__ Dual& operator*=(Dual& a, const Dual& b) { return a = a * b; }
__device__ __forceinline__ Dual& operator/=(Dual& a, const Dual& b) { return a = a / b; }
__device__ __forceinline__ bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
__device__ __forceinline__ bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

__device__ __forceinline__ Dual sqrt(const Dual& a) {
  double s = ::sqrt(a.v);
  return scaled(a, s, 0.5 / s);
}
__device__ __forceinline__ Dual fabs(const Dual& a) { return scaled(a, ::fabs(a.v), a.v < 0.0 ? -1.0 : 1.0); }
__device__ __forceinline__ Dual sin(const Dual& a) { return scaled(a, ::sin(a.v), ::cos(a.v)); }
__device__ __forceinline__ Dual cos(const Dual& a) { return scaled(a, ::cos(a.v), -::sin(a.v)); }
__device__ __forceinline__ void sincos(const Dual& a, Dual* s, Dual* c) {
  double sv, cv;
  ::sincos(a.v, &sv, &cv);
  *s = scaled(a, sv, cv);
  *c = scaled(a, cv, -sv);
}
__device__ __forceinline__ Dual sinh(const Dual& a) { return scaled(a, ::sinh(a.v), ::cosh(a.v)); }
__device__ __forceinline__ Dual cosh(const Dual& a) { return scaled(a, ::cosh(a.v), ::sinh(a.v)); }
__device__ __forceinline__ Dual expm1(const Dual& a) { return scaled(a, ::expm1(a.v), ::exp(a.v)); }
__device__ __forceinline__ Dual asin(const Dual& a) { return scaled(a, ::asin(a.v), 1.0 / ::sqrt(1.0 - a.v * a.v)); }
__device__ __forceinline__ Dual atan2(const Dual& y, const Dual& x) {
  Dual r;
  r.v = ::atan2(y.v, x.v);
  double q = 1.0 / (x.v * x.v + y.v * y.v);
  for (int k = 0; k < 6; k++) r.d[k] = (x.v * y.d[k] - y.v * x.d[k]) * q;
  return r;
}
// fmod(x, p) and rint: piecewise x - k p and a constant
__device__ __forceinline__ Dual fmod(const Dual& x, const Dual& p) {
  double m = ::fmod(x.v, p.v);
  Dual r = x - Dual((x.v - m) / p.v) * p;
  r.v = m;
  return r;
}
__device__ __forceinline__ Dual rint(const Dual& a) { return Dual(::rint(a.v)); }

#define BMX_TANGENT 1
#define double Dual
#include "bmx_ops.h"
#undef double

// One particle per lane through [lo1, hi1) then [lo2, hi2) of the op program
// for n_turns turns, as the bunch kernel tracks it (orbit_too_large at entry,
// then every op); on loss the lane stops with the state and the coordinates
// at the op that lost it, and a zero matrix.
__global__ __launch_bounds__(64) void bmx_transfer_kernel(const DevOp* __restrict__ ops, int lo1, int hi1, int lo2,
                                                          int hi2, LatticeConsts lc, int n_turns,
                                                          const bmx_f64* __restrict__ in, bmx_f64* __restrict__ out,
                                                          bmx_f64* __restrict__ mat, int32_t* __restrict__ state,
                                                          int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Lane L;
  Dual* c[6] = {&L.x, &L.px, &L.y, &L.py, &L.z, &L.pz};
  for (int k = 0; k < 6; k++) {
    *c[k] = Dual(in[i * 6 + k]);
    c[k]->d[k] = 1.0;
  }
  // beta from pz as init_coord forms it (the reference beta at pz = 0)
  if (in[i * 6 + 5] == 0.0) {
    L.beta = Dual(lc.beta0);
  } else {
    Dual pc = Dual(lc.p0c) * (Dual(1.0) + L.pz);
    L.beta = pc / sqrt(pc * pc + Dual(lc.mc2sq));
  }
  L.dt = Dual(0.0);
  L.state = BMX_ALIVE;
  refresh_invariants(L, lc);
  TurnTime tt;
  tt.k = 0;
  tt.r = 0.0;
  tt.turn = 0;
  bool alive = !too_large(L, lc);
  for (int t = 0; t < n_turns && alive; t++) {
    tt.turn = t;
    for (int seg = 0; seg < 2 && alive; seg++) {
      int lo = seg ? lo2 : lo1, hi = seg ? hi2 : hi1;
      for (int k = lo; k < hi; k++) {
        const DevOp& op = ops[k];
        run_op(L, op, op.type, op.flags, op.n, lc, tt);
        if (L.state != BMX_ALIVE) {
          alive = false;
          break;
        }
      }
    }
  }
  for (int r = 0; r < 6; r++) {
    out[i * 6 + r] = c[r]->v;
    for (int q = 0; q < 6; q++) mat[i * 36 + r * 6 + q] = alive ? c[r]->d[q] : 0.0;
  }
  state[i] = L.state;
}

}  // namespace tangent

hipError_t launch_transfer(hipStream_t stream, const DevOp* ops, int lo1, int hi1, int lo2, int hi2,
                           const LatticeConsts& lc, int n_turns, const double* d_in, double* d_out, double* d_mat,
                           int32_t* d_state, int64_t n) {
  if (n <= 0) return hipSuccess;
  const int block = 64;
  unsigned grid = (unsigned)((n + block - 1) / block);
  hipLaunchKernelGGL(tangent::bmx_transfer_kernel, dim3(grid), dim3(block), 0, stream, ops, lo1, hi1, lo2, hi2, lc,
                     n_turns, d_in, d_out, d_mat, d_state, n);
  return hipGetLastError();
}

}  // namespace bmx
