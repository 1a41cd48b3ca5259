// bmx_track.hip -- gfx950 tracking kernels and the libbmx C ABI.
//
// One particle per lane.  Each lane keeps its 6-D coordinates, beta and two
// per-particle invariants of the drift map in registers for the whole launch
// (all turns, all elements): the structure-of-arrays store in HBM is read once
// and written once per launch.  The element program (bmx_internal.h) is walked
// with a wave-uniform index, so op constants come through the scalar cache and
// the op dispatch is a scalar branch.  Lost lanes leave the loop early and
// record Bmad's loss state, element and turn.
//
// The arithmetic of every op restates the reference routine cited at it in
// FP64, with particle-independent factors hoisted by the lattice compiler and
// divisions shared where the algebra allows (1/(1+pz) and beta/beta0 - 1 are
// per-particle constants between energy kicks).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <functional>
#include <vector>

#include "../../include/bmx.h"
#include "bmx_internal.h"

namespace bmx {
bool compile_variant(const bmx_ele_rec* eles, int n_ele, const bmx_ref& ref, const bmx_config& cfg,
                     std::vector<DevOp>& prog, std::vector<int32_t>& ele_op_begin, LatticeConsts* lc,
                     std::string* err);
}

using namespace bmx;

#define D_C_LIGHT 2.99792458e8
#define D_TWOPI 6.283185307179586476925286766559
#define D_PI 3.141592653589793238462643383279

// ---------------------------------------------------------------------------
// per-lane particle state
// ---------------------------------------------------------------------------
struct Lane {
  double x, px, y, py, z, pz, beta;
  double irp;  // 1 / (1 + pz)
  double bb;   // beta/beta0 - 1 as track_a_drift forms it (track_a_drift.f90:68-69)
  int state;
};

// Series coefficients.  FP64 VALU instructions take no 64-bit literal, so a
// coefficient has to sit in a register pair.  Left alone, the compiler hoists
// every one of them out of the element loop into VGPRs and spills them.
// Instead each series reads its coefficients at the point of use through the
// scalar cache: series() hands out the table address through an empty asm,
// so the loads cannot be hoisted and become a few s_load_dwordx8 into
// SGPRs, and sg() keeps each loaded value in its SGPR pair, where the FMA
// takes it as an operand.  Per block: the highest-order coefficient last.
typedef __attribute__((address_space(4))) const double cdouble;
enum : int {
  CS_DRIFT = 0,   // 5/16, 35/128, 63/256: binomial series of 1/sqrt(1-u) - 1 from u^3
  CS_QUAD = 3,    // (4m)!, (4m+2)!, (4m+1)!, (4m+3)! reciprocals, m = 0..5
  CS_ASIN = 27,   // asin(s) = s + s^3 P(s^2)
  CS_SIN = 33,    // sin(e) = e + e^3 P(e^2)
  CS_COS = 38,    // cos(e) = 1 - e^2/2 + e^4 P(e^2)
  CS_ATAN = 43,   // atan(t) = t + t^3 P(t^2)
  CS_ISINC = 49,  // x / sin(x) = 1 + x^2 P(x^2)
  CS_END = 56
};
__constant__ double kSeries[CS_END] = {
    // CS_DRIFT
    0.3125, 0.2734375, 0.24609375,
    // CS_QUAD
    1.0, 0.5, 1.0, 1.0 / 6,
    1.0 / 24, 1.0 / 720, 1.0 / 120, 1.0 / 5040,
    1.0 / 40320, 1.0 / 3628800, 1.0 / 362880, 1.0 / 39916800,
    1.0 / 479001600, 1.0 / 87178291200.0, 1.0 / 6227020800.0, 1.0 / 1307674368000.0,
    1.0 / 20922789888000.0, 1.0 / 6402373705728000.0, 1.0 / 355687428096000.0, 1.0 / 121645100408832000.0,
    1.0 / 2432902008176640000.0, 1.0 / 1124000727777607680000.0, 1.0 / 51090942171709440000.0,
    1.0 / 25852016738884976640000.0,
    // CS_ASIN
    1.0 / 6, 3.0 / 40, 5.0 / 112, 35.0 / 1152, 63.0 / 2816, 231.0 / 13312,
    // CS_SIN
    -1.0 / 6, 1.0 / 120, -1.0 / 5040, 1.0 / 362880, -1.0 / 39916800,
    // CS_COS
    1.0 / 24, -1.0 / 720, 1.0 / 40320, -1.0 / 3628800, 1.0 / 479001600,
    // CS_ATAN
    -1.0 / 3, 1.0 / 5, -1.0 / 7, 1.0 / 9, -1.0 / 11, 1.0 / 13,
    // CS_ISINC
    1.0 / 6, 7.0 / 360, 31.0 / 15120, 127.0 / 604800, 73.0 / 3421440, 1414477.0 / 653837184000.0,
    8191.0 / 37362124800.0,
};

__device__ __forceinline__ const cdouble* series(int block) {
  const cdouble* p = (const cdouble*)kSeries + block;
  asm volatile("" : "+s"(p));
  return p;
}

__device__ __forceinline__ double sg(double c) {
  asm("" : "+s"(c));
  return c;
}

// c[0] + x (c[1] + x (... + x c[N-1]))
template <int N>
__device__ __forceinline__ double poly(const cdouble* c, double x) {
  double a = sg(c[N - 1]);
#pragma unroll
  for (int i = N - 2; i >= 0; i--) a = a * x + sg(c[i]);
  return a;
}

// Keeps the compiler from merging the lane-dependent tests of a block into
// the wave-uniform test that guards it (which would evaluate them, and mask
// exec, on every op whether the block runs or not).
__device__ __forceinline__ void uniform_block() { asm volatile(""); }

// true on every lane of the wave when b holds on any active lane
__device__ __forceinline__ bool wave_any(bool b) { return __builtin_amdgcn_ballot_w64(b) != 0; }

// sqrt_one (sim_utils/special_functions/sqrt_one.f90:18-41)
__device__ __forceinline__ double d_sqrt_one(double x) { return x / (sqrt(1.0 + x) + 1.0); }

// beta/beta0 - 1 = sqrt_one(mc2^2 (2 pz + pz^2) / ((p0c (1+pz))^2 + mc2^2))
__device__ __forceinline__ void refresh_invariants(Lane& L, const LatticeConsts& lc) {
  double rel = 1.0 + L.pz;
  L.irp = 1.0 / rel;
  double pc = lc.p0c * rel;
  L.bb = d_sqrt_one((lc.mc2sq * (2.0 * L.pz + L.pz * L.pz)) / (pc * pc + lc.mc2sq));
}

// modulo2 (sim_utils/special_functions/modulo2_mod.f90:85-100)
__device__ __forceinline__ double d_modulo2(double x, double amp) {
  double p = 2.0 * amp;
  double m = fmod(x, p);
  if (m != 0.0 && m < 0.0) m += p;
  if (m >= amp) m -= p;
  return m;
}

// ---------------------------------------------------------------------------
// ops
// ---------------------------------------------------------------------------

// track_a_drift (bmad/low_level/track_a_drift.f90:25-116), forward, with reference motion.
// With u = (px^2 + py^2)/(1+pz)^2 and S = 1/sqrt(1-u) - 1:
//   x += L px_rel (1 + S),  z += L (beta/beta0 - 1 - S)   [sqrt_one(-u)/ps = -S].
// For u < 1e-4 (angles below 10 mrad) S is summed as its binomial series to
// u^5 (truncation < 3e-25), which needs no division or square root; larger
// angles take the exact form.
__device__ __forceinline__ void drift_apply(Lane& L, double len, double pxr, double pyr, double S) {
  double lps = len * (1.0 + S);
  L.x += lps * pxr;
  L.y += lps * pyr;
  L.z += len * (L.bb - S);
}

// The series is evaluated for every lane; only a wave holding a lane at
// u >= 1e-4 takes the exact form (and the lost_pz test, u >= 1, which leaves
// the particle where it is) for those lanes.
__device__ __forceinline__ void drift_c(Lane& L, double len, double d0, double d1, double d2) {
  double pxr = L.px * L.irp, pyr = L.py * L.irp;
  double u = pxr * pxr + pyr * pyr;
  double S = u * (0.5 + u * (0.375 + u * (d0 + u * (d1 + u * d2))));
  bool wide = u >= 1e-4;
  if (wave_any(wide)) {
    bool move = true;
    if (wide) {
      if (u >= 1.0) {
        L.state = BMX_LOST_PZ;
        move = false;
      } else {
        double ps = sqrt(1.0 - u);
        S = u / (ps * (ps + 1.0));
      }
    }
    if (move) drift_apply(L, len, pxr, pyr, S);
    return;
  }
  drift_apply(L, len, pxr, pyr, S);
}

__device__ __forceinline__ void op_drift(Lane& L, double len) {
  const cdouble* d = series(CS_DRIFT);
  drift_c(L, len, sg(d[0]), sg(d[1]), sg(d[2]));
}

// drift with include_ref_motion = false (offset_particle's drift to the edge)
__device__ __forceinline__ void op_drift_noref(Lane& L, double len) {
  double pxr = L.px * L.irp, pyr = L.py * L.irp;
  double pxy2 = pxr * pxr + pyr * pyr;
  if (pxy2 >= 1.0) { L.state = BMX_LOST_PZ; return; }
  double ips = 1.0 / sqrt(1.0 - pxy2);
  L.x += len * pxr * ips;
  L.y += len * pyr * ips;
  L.z += -len * ips;
}

// (kr + i ki) *= (ar + i ai)
__device__ __forceinline__ void cmul(double& kr, double& ki, double ar, double ai) {
  double tr = kr * ar - ki * ai;
  double ti = kr * ai + ki * ar;
  kr = tr;
  ki = ti;
}

// Thin multipole kick, ab form (bmad/modules/multipole_mod.f90:82-174, :314-401):
// K = sum_n c_n (x + i y)^n, dpx = -Re K, dpy = Im K, Horner in complex FP64.
// c holds nh + 1 complex coefficients for orders lo .. lo + nh; the common
// factor (x + i y)^lo is applied as powers of (x + i y)^2, without a loop for
// lo <= 3 (nh and lo are wave-uniform).
__device__ __forceinline__ void kick_horner(const double* c, int nh, int lo, double x, double y, double& kr,
                                            double& ki) {
  kr = c[2 * nh];
  ki = c[2 * nh + 1];
  for (int j = nh - 1; j >= 0; j--) {
    cmul(kr, ki, x, y);
    kr += c[2 * j];
    ki += c[2 * j + 1];
  }
  if (lo & 1) cmul(kr, ki, x, y);
  if (lo >= 2) {
    double wr = x * x - y * y, wi = 2.0 * x * y;  // (x + i y)^2
    cmul(kr, ki, wr, wi);
    for (int m = lo - 2; m >= 2; m -= 2) cmul(kr, ki, wr, wi);
  }
}

__device__ __forceinline__ void kick_eval(const DevOp& op, double x, double y, double& kr, double& ki) {
  kick_horner(op.c, __builtin_amdgcn_readfirstlane(op.flags) & 0xff, __builtin_amdgcn_readfirstlane(op.n), x, y, kr,
              ki);
}

__device__ __forceinline__ void op_kick(Lane& L, const DevOp& op) {
  double kr, ki;
  kick_eval(op, L.x, L.y, kr, ki);
  L.px -= kr;
  L.py += ki;
}

// Body of a thick multipole (track_a_thick_multipole.f90:91-121): kick x 1/2,
// then n_step x (drift, kick), the last kick x 1/2.  As in the reference the
// loop runs on after a drift has set lost_pz; the element end checks state.
// c: 0 step_len, 1.. full-step kick coefficients; n = n_step; flags = nh | lo << 8
// The body for a single multipole order lo = LO (2: sextupole, 3: octupole)
// with the kick of kick_horner written out: the coefficients and the drift
// series are read once for all slices.  cb, ca: full-step b, a; the half
// kicks use 0.5 cb, 0.5 ca, which scales the kick exactly by 1/2.
template <int LO>
__device__ __forceinline__ void kd_single(Lane& L, double len, int n_step, double cb, double ca) {
  const cdouble* d = series(CS_DRIFT);
  double d0 = sg(d[0]), d1 = sg(d[1]), d2 = sg(d[2]);
  auto kick = [&](double b, double a) {
    double x = L.x, y = L.y;
    double kr = b, ki = a;
    if (LO & 1) cmul(kr, ki, x, y);
    cmul(kr, ki, x * x - y * y, 2.0 * x * y);
    L.px -= kr;
    L.py += ki;
  };
  kick(0.5 * cb, 0.5 * ca);
  for (int i = 1; i <= n_step; i++) {
    int st = L.state;
    drift_c(L, len, d0, d1, d2);
    L.state = (st != BMX_ALIVE) ? st : L.state;  // first loss state wins
    if (i == n_step) kick(0.5 * cb, 0.5 * ca);
    else kick(cb, ca);
  }
}

__device__ __forceinline__ void op_kd_slices(Lane& L, const DevOp& op) {
  int n_step = __builtin_amdgcn_readfirstlane(op.n);
  int fl = __builtin_amdgcn_readfirstlane(op.flags);
  int nh = fl & 0xff, lo = (fl >> 8) & 0xff;
  double len = op.c[0];
  if (nh == 0 && lo == 2) {
    kd_single<2>(L, len, n_step, op.c[1], op.c[2]);
    return;
  }
  if (nh == 0 && lo == 3) {
    kd_single<3>(L, len, n_step, op.c[1], op.c[2]);
    return;
  }
  double kr, ki;
  kick_horner(&op.c[1], nh, lo, L.x, L.y, kr, ki);
  L.px -= 0.5 * kr;
  L.py += 0.5 * ki;
  for (int i = 1; i <= n_step; i++) {
    int st = L.state;
    op_drift(L, len);
    L.state = (st != BMX_ALIVE) ? st : L.state;  // first loss state wins
    kick_horner(&op.c[1], nh, lo, L.x, L.y, kr, ki);
    double h = (i == n_step) ? 0.5 : 1.0;
    L.px -= h * kr;
    L.py += h * ki;
  }
}

// apply_multipole_kicks of track_a_bend (track_a_bend.f90:338-471): kick scaled by (1 + g x).
__device__ __forceinline__ void op_bend_kick(Lane& L, const DevOp& op) {
  double kr, ki;
  kick_eval(op, L.x, L.y, kr, ki);
  double f = 1.0 + op.c[13] * L.x;
  L.px -= f * kr;
  L.py += f * ki;
}

// n = 0 kt kick of a multipole element with ref_orb_offset (multipole_mod.f90:249-255)
__device__ __forceinline__ void op_kick_ref0(Lane& L, const DevOp& op) {
  double kc = op.c[0], ks = op.c[1];
  L.px += kc * L.pz;
  L.py += ks * L.pz;
  L.z -= kc * L.x + ks * L.y;
}

// low_energy_z_correction (bmad/low_level/low_energy_z_correction.f90:22-58)
__device__ __forceinline__ double low_energy_dz(const Lane& L, const LatticeConsts& lc, double ds) {
  double pz = L.pz;
  if (lc.lez_k * (pz * pz) < lc.lez_thr) {
    return ds * pz * (1.0 - lc.lez_a * pz + pz * pz * lc.lez_f) * lc.lez_me2;
  }
  return ds * (L.beta - lc.beta0) / lc.beta0;
}

// One quadrupole body step: quad_mat2_calc (bmad/low_level/quad_mat2_calc.f90:22-81)
// for both planes plus the quadratic z terms (track_a_quadrupole.f90:82-112).
// With w = k L^2 for a plane of strength k: c = sum w^n/(2n)!, s/L = sum w^n/(2n+1)!
// (cosh / sinh for k > 0, cos / sin for k < 0).  For |w| <= 1 the series to
// n = 11 is exact to < 1e-21 and replaces sincos, expm1, sqrt and three
// divisions; stronger quadrupoles take the exact trigonometric form.  The two
// planes have w and -w: with v = w^2 each series splits into even and odd
// parts in v, c(+-w) = E(v) +- w O(v), so both planes cost one pass.
__device__ __forceinline__ void quad_cs_pair(double w, double len, double& cp, double& sp, double& cm, double& sm) {
  const cdouble* q = series(CS_QUAD);
  double v = w * w;
  double ec = sg(q[20]), oc = sg(q[21]), es = sg(q[22]), os = sg(q[23]);
#pragma unroll
  for (int m = 4; m >= 0; m--) {
    ec = ec * v + sg(q[4 * m]);
    oc = oc * v + sg(q[4 * m + 1]);
    es = es * v + sg(q[4 * m + 2]);
    os = os * v + sg(q[4 * m + 3]);
  }
  cp = ec + w * oc;
  cm = ec - w * oc;
  sp = (es + w * os) * len;
  sm = (es - w * os) * len;
}

__device__ __forceinline__ void op_quad(Lane& L, const DevOp& op, const LatticeConsts& lc) {
  double len = op.c[0];
  double rel_p = 1.0 + L.pz;
  double k1 = op.c[1] * L.irp;  // charge_dir b1 / (L rel_p)
  double kx = -k1, ky = k1;     // quad_mat2_calc(-k1, ...) for x, quad_mat2_calc(k1, ...) for y
  double cx, sx, cy, sy;
  double w = k1 * len * len;
  if (fabs(w) <= 1.0) {
    quad_cs_pair(w, len, cy, sy, cx, sx);
  } else {
    double sqrt_k = sqrt(fabs(k1));
    double skl = sqrt_k * len;
    double sn, cs;
    sincos(skl, &sn, &cs);
    double em1 = expm1(skl);
    double ch = 1.0 + em1 * em1 / (2.0 * (em1 + 1.0));
    double sh = 0.5 * (em1 + em1 / (em1 + 1.0));
    double isk = 1.0 / sqrt_k;
    if (k1 > 0.0) { cx = cs; sx = sn * isk; cy = ch; sy = sh * isk; }  // x focusing
    else          { cx = ch; sx = sh * isk; cy = cs; sy = sn * isk; }  // y focusing
  }
  double irp = L.irp, irp2 = irp * irp;
  double zx0 = kx * (-cx * sx + len) / 4.0, zx1 = -kx * sx * sx * irp / 2.0, zx2 = -(cx * sx + len) * irp2 / 4.0;
  double zy0 = ky * (-cy * sy + len) / 4.0, zy1 = -ky * sy * sy * irp / 2.0, zy2 = -(cy * sy + len) * irp2 / 4.0;
  double x = L.x, px = L.px, y = L.y, py = L.py;
  L.z += zx0 * x * x + zx1 * x * px + zx2 * px * px + zy0 * y * y + zy1 * y * py + zy2 * py * py;
  L.x = cx * x + (sx * irp) * px;
  L.px = (kx * sx * rel_p) * x + cx * px;
  L.y = cy * y + (sy * irp) * py;
  L.py = (ky * sy * rel_p) * y + cy * py;
  L.z += low_energy_dz(L, lc, len);
}

// sinc (sim_utils/special_functions/sinc.f90)
__device__ __forceinline__ double d_sinc(double x) { return (fabs(x) < 1e-8) ? 1.0 : sin(x) / x; }

// Short Taylor series for the small angles of the bend body.  Each is written
// as x + x^3 P(x^2), so the leading term enters exactly and the last addition
// rounds once: the result is correctly rounded in nearly every case, like the
// host libm calls the reference makes (those angles enter the exit px through
// a difference of O(angle) terms, so one ulp there is visible downstream).

// asin(s), |s| <= 0.05, to s^13 (truncation < 1e-19 relative).
__device__ __forceinline__ double asin_small(double s) {
  double s2 = s * s;
  return s + s * s2 * poly<6>(series(CS_ASIN), s2);
}

// sin(e), |e| <= 0.2, to e^11 (truncation < 1e-18 relative).
__device__ __forceinline__ double sin_small(double e) {
  double e2 = e * e;
  return e + e * e2 * poly<5>(series(CS_SIN), e2);
}

// cos(e), |e| <= 0.2, to e^12 (truncation < 1e-20).
__device__ __forceinline__ double cos_small(double e) {
  double e2 = e * e;
  return 1.0 + e2 * (-0.5 + e2 * poly<5>(series(CS_COS), e2));
}

// atan(t), |t| <= 0.2, to t^13 (truncation < 1e-11 * t^14, < 2e-21 relative).
__device__ __forceinline__ double atan_small(double t) {
  double t2 = t * t;
  return t + t * t2 * poly<6>(series(CS_ATAN), t2);
}

// x / sin(x) = 1 / sinc(x), |x| <= 0.1, to x^14 (truncation < 1e-20).
__device__ __forceinline__ double inv_sinc_small(double x) {
  double x2 = x * x;
  return 1.0 + x2 * poly<7>(series(CS_ISINC), x2);
}

// The reference's transcendental calls for the bend entry and exit angles
// (track_a_bend.f90:172-177, 205-207) for large angles.  Kept out of line:
// rare, and inlined they would raise the register allocation of the whole
// element loop.
struct BendAngles { double phi_1, sin_plus, cos_plus; };
struct BendExit { double L_p, se; };

__device__ __noinline__ BendAngles bend_entry_exact(double sphi, double angle) {
  BendAngles b;
  b.phi_1 = asin(sphi);
  b.sin_plus = sin(angle + b.phi_1);
  b.cos_plus = cos(angle + b.phi_1);
  return b;
}

__device__ __noinline__ BendExit bend_exit_exact(double phi_1, double L_u, double L_v, double angle, double L_c) {
  double angle_p = 2.0 * (angle + phi_1 - atan2(L_u, -L_v));
  return BendExit{L_c / d_sinc(angle_p / 2.0), sin(phi_1 + angle - angle_p)};
}

// Square root, reciprocal and quotient for operands known to lie well inside
// the normal range (the bend's common path checks that they do): the IEEE
// sequences the compiler emits for sqrt() and '/', without their range
// scaling and special-value fix-ups.  From v_rsq / v_rcp (relative error
// below 2^-22) two Newton steps and an FMA residual correction give the
// correctly rounded result.
__device__ __forceinline__ double sqrt_mid(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

__device__ __forceinline__ double rcp_iter(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  return fma(r, e, r);
}

__device__ __forceinline__ double rcp_mid(double b) {
  double r = rcp_iter(b);
  return fma(fma(-b, r, 1.0), r, r);
}

__device__ __forceinline__ double div_mid(double a, double b) {
  double r = rcp_iter(b);
  double q = a * r;
  return fma(fma(-b, q, a), r, q);
}

// Exact sector-bend body step (track_a_bend.f90:120-209), in the reference's
// order of operations: phi_1 = asin(px/pt), sin/cos(angle + phi_1), the exit
// offset xi, angle_p = 2 (angle + phi_1 - atan2(L_u, -L_v)),
// px = pt sin(phi_1 + angle - angle_p), L_p = L_c / sinc(angle_p / 2).
// This is the general form, with every branch of the reference and the libm
// calls for large angles; a wave runs it when any of its lanes leaves the
// common path of op_bend.  px/pt is taken as px * (1/pt) and beta sl /
// beta_ref as beta (sl / beta_ref), each within an ulp of the reference's
// quotient, so that on the common path both forms give the same bits.
// c: 0 step_len, 1 g, 2 g_tot, 3 angle, 4 cos a, 5 sin a, 6 sinc a, 7 cosc a, 8 sincc a, 9 beta_ref, 10 dg
__device__ __forceinline__ Lane bend_general(Lane L, const double* c, double p0c, double mc2sq, double mal,
                                          double two_mal) {
  double sl = c[0], g = c[1], g_tot = c[2], angle = c[3];
  double cos_a = c[4], sin_a = c[5], sinc_a = c[6], cosc_a = c[7];
  double x = L.x, px = L.px, py = L.py, pz = L.pz;
  double rel_p = 1.0 + pz;
  if (c[10] == 0.0 && fabs(x * g) < 1e-9 && fabs(px) < 1e-9 && fabs(py) < 1e-9 && fabs(pz) < 1e-9) {
    double pc = rel_p * p0c;
    double gam2 = mc2sq / (pc * pc + mc2sq);
    double m56 = sl * (gam2 - (g * sl) * (g * sl) * c[8]);
    L.x = cos_a * x + sl * sinc_a * px + g * sl * sl * cosc_a * pz;
    L.px = -g * sin_a * x + cos_a * px + g * sl * sinc_a * pz;
    L.y = L.y + sl * py;
    L.z = -g * sl * sinc_a * x - g * sl * sl * cosc_a * px + L.z + m56 * pz;
    return L;
  }
  double pt = sqrt(rel_p * rel_p - py * py);
  if (fabs(px) > pt) { L.state = BMX_LOST_PZ; return L; }
  double ipt = 1.0 / pt;
  double g_p = g_tot * ipt;
  double sphi = px * ipt;
  double phi_1, sin_plus, cos_plus;
  if (fabs(sphi) <= 0.05 && fabs(angle) <= 0.15) {
    phi_1 = asin_small(sphi);
    double u = angle + phi_1;
    sin_plus = sin_small(u);
    cos_plus = cos_small(u);
  } else {
    BendAngles b = bend_entry_exact(sphi, angle);
    phi_1 = b.phi_1;
    sin_plus = b.sin_plus;
    cos_plus = b.cos_plus;
  }
  double q = (1.0 + g * x) * sl * sinc_a;
  double alpha = 2.0 * (1.0 + g * x) * sin_plus * sl * sinc_a - g_p * q * q;
  double r = cos_plus * cos_plus + g_p * alpha;
  if (r < 0.0 || (fabs(g_p) < 1e-5 && fabs(cos_plus) < 1e-5)) {
    L.state = BMX_LOST;
    L.x = two_mal;
    return L;
  }
  double rad = sqrt(r);
  double xi;
  if (cos_plus > 0.0) xi = alpha / (rad + cos_plus);
  else xi = (rad - cos_plus) / g_p;
  double xn = x * cos_a - sl * sl * g * cosc_a + xi;
  L.x = xn;
  if (fabs(xn) > mal) { L.state = BMX_LOST; return L; }
  double L_u = xi;
  double L_v = -(sl * sinc_a + x * sin_a);
  double L_c = sqrt(L_v * L_v + L_u * L_u);
  double L_p, se;
  bool small = L_v < 0.0 && fabs(L_u) <= -0.2 * L_v;
  if (small) {
    double half_p = angle + phi_1 - atan_small(L_u / -L_v);  // angle_p / 2
    double e = phi_1 + angle - 2.0 * half_p;                 // phi_1 + angle - angle_p
    small = fabs(half_p) <= 0.1 && fabs(e) <= 0.2;
    if (small) {
      L_p = (fabs(half_p) < 1e-8) ? L_c : L_c * inv_sinc_small(half_p);
      se = sin_small(e);
    }
  }
  if (!small) {
    BendExit b = bend_exit_exact(phi_1, L_u, L_v, angle, L_c);
    L_p = b.L_p;
    se = b.se;
  }
  L.px = pt * se;
  L.y = L.y + py * L_p * ipt;
  L.z = L.z + L.beta * c[13] - rel_p * L_p * ipt;
  return L;
}

// The common path of the bend body: small entry and exit angles (the series
// of asin, sin, cos, atan, 1/sinc), cos(angle + phi_1) > 0, no loss, and the
// square roots and quotients in mid range.  It is the general form's own
// sequence on that path (L_c / sinc(h) for |h| < 1e-8 is L_c exactly, as in
// the reference), without its branches.  The whole wave takes it unless some
// lane leaves it; then the whole wave takes bend_general.
// c: as bend_general, 13 step_len / beta_ref
__device__ __forceinline__ void op_bend(Lane& L, const DevOp& op, const LatticeConsts& lc) {
  double sl = op.c[0], g = op.c[1], g_tot = op.c[2], angle = op.c[3];
  double cos_a = op.c[4], sin_a = op.c[5], sinc_a = op.c[6], cosc_a = op.c[7];
  double x = L.x, px = L.px, py = L.py, pz = L.pz;
  double rel_p = 1.0 + pz;
  double p2 = rel_p * rel_p - py * py;
  double pt = sqrt_mid(p2);
  double ipt = rcp_mid(pt);
  double g_p = g_tot * ipt;
  double sphi = px * ipt;
  double phi_1 = asin_small(sphi);
  double u = angle + phi_1;
  double sin_plus = sin_small(u), cos_plus = cos_small(u);
  double q = (1.0 + g * x) * sl * sinc_a;
  double alpha = 2.0 * (1.0 + g * x) * sin_plus * sl * sinc_a - g_p * q * q;
  double r = cos_plus * cos_plus + g_p * alpha;
  double xi = div_mid(alpha, sqrt_mid(r) + cos_plus);
  double xn = x * cos_a - sl * sl * g * cosc_a + xi;
  double L_u = xi;
  double L_v = -(sl * sinc_a + x * sin_a);
  double L_c2 = L_v * L_v + L_u * L_u;
  double half_p = angle + phi_1 - atan_small(div_mid(L_u, -L_v));
  double e = phi_1 + angle - 2.0 * half_p;
  bool paraxial = (op.c[10] == 0.0) & (fabs(x * g) < 1e-9) & (fabs(px) < 1e-9) & (fabs(py) < 1e-9) & (fabs(pz) < 1e-9);
  bool common = !paraxial & (p2 > 0.25) & (p2 < 4.0) & (fabs(sphi) <= 0.05) & (fabs(angle) <= 0.15) & (r >= 0.25) &
                (r <= 4.0) & (fabs(xn) <= lc.max_aperture_limit) & (L_v < -1e-100) & (L_c2 < 1e200) &
                (fabs(L_u) <= -0.2 * L_v) & (fabs(half_p) <= 0.1) & (fabs(e) <= 0.2);
  if (wave_any(!common)) {
    L = bend_general(L, op.c, lc.p0c, lc.mc2sq, lc.max_aperture_limit, lc.two_max_ap);
    return;
  }
  double L_p = sqrt_mid(L_c2) * inv_sinc_small(half_p);
  double se = sin_small(e);
  L.x = xn;
  L.px = pt * se;
  L.y = L.y + py * L_p * ipt;
  L.z = L.z + L.beta * op.c[13] - rel_p * L_p * ipt;
}

// sbend_body_with_k1_map (track_a_bend.f90:500-659), forward, orientation +1.
// c: 0 step_len, 1 g, 2 g_tot, 3 k1
__device__ void op_bend_k1(Lane& L, const DevOp& op, const LatticeConsts& lc) {
  double length = op.c[0], g = op.c[1], g_tot = op.c[2], k1 = op.c[3];
  double rel_p = 1.0 + L.pz, rel_p2 = rel_p * rel_p;
  double k_x = k1 + g * g_tot;
  double x_c = (g * rel_p - g_tot) / k_x;
  double om_x = sqrt(fabs(k_x) / rel_p), om_y = sqrt(fabs(k1) / rel_p);
  double tau_x = (k_x >= 0.0) ? -1.0 : 1.0, tau_y = (k1 >= 0.0) ? 1.0 : -1.0;
  double arg = om_x * length, s_x, c_x, z2, s_y, c_y;
  if (arg < 1e-6) {
    s_x = (1.0 + tau_x * arg * arg / 6.0) * length;
    c_x = 1.0 + tau_x * arg * arg / 2.0;
    z2 = g * length * length / (2.0 * rel_p);
  } else if (k_x > 0.0) {
    s_x = sin(arg) / om_x;
    c_x = cos(arg);
    z2 = tau_x * g * (1.0 - c_x) / (rel_p * om_x * om_x);
  } else {
    s_x = sinh(arg) / om_x;
    c_x = cosh(arg);
    z2 = tau_x * g * (1.0 - c_x) / (rel_p * om_x * om_x);
  }
  arg = om_y * length;
  if (arg < 1e-6) {
    s_y = (1.0 + tau_y * arg * arg / 6.0) * length;
    c_y = 1.0 + tau_y * arg * arg / 2.0;
  } else if (k1 < 0.0) {
    s_y = sin(arg) / om_y;
    c_y = cos(arg);
  } else {
    s_y = sinh(arg) / om_y;
    c_y = cosh(arg);
  }
  double r0 = L.x - x_c, r1 = L.px, r2 = L.y, r3 = L.py, r4 = L.z;
  double z0 = -g * x_c * length, z1 = -g * s_x;
  double z11 = tau_x * om_x * om_x * (length - c_x * s_x) / 4.0;
  double z12 = -tau_x * om_x * om_x * s_x * s_x / (2.0 * rel_p);
  double z22 = -(length + c_x * s_x) / (4.0 * rel_p2);
  double z33 = tau_y * om_y * om_y * (length - c_y * s_y) / 4.0;
  double z34 = -tau_y * om_y * om_y * s_y * s_y / (2.0 * rel_p);
  double z44 = -(length + c_y * s_y) / (4.0 * rel_p2);
  L.x = c_x * r0 + s_x * r1 / rel_p + x_c;
  L.px = tau_x * om_x * om_x * rel_p * s_x * r0 + c_x * r1;
  L.y = c_y * r2 + s_y * r3 / rel_p;
  L.py = tau_y * om_y * om_y * rel_p * s_y * r2 + c_y * r3;
  L.z = r4 + z0 + z1 * r0 + z2 * r1 + z11 * r0 * r0 + z12 * r0 * r1 + z22 * r1 * r1 + z33 * r2 * r2 + z34 * r2 * r3 +
        z44 * r3 * r3;
  L.z += low_energy_dz(L, lc, length);
}

// hwang_bend_edge_kick (bmad/modules/fringe_mod.f90:240-418), time_dir = +1.
// c: 0 gt, 1 gt2, 2 gs2, 3 k1_tane, 4 fg_factor, 5 g_tot, 6 tan_e
__device__ __forceinline__ void op_hwang(Lane& L, const DevOp& op) {
  double gt = op.c[0], gt2 = op.c[1], gs2 = op.c[2], k1t = op.c[3], fg = op.c[4], g_tot = op.c[5], te = op.c[6];
  double ef = L.irp;
  double v0 = L.x, v1 = L.px, v2 = L.y, v3 = L.py;
  double dx, dpx, dy, dpy, dz;
  if (__builtin_amdgcn_readfirstlane(op.flags) & F_HWANG_ENTER) {
    dx = (-gt2 * v0 * v0 + gs2 * v2 * v2) * ef / 2.0;
    dpx = (gt * g_tot * (1.0 + 2.0 * te * te) * v2 * v2 / 2.0 + gt2 * (v0 * v1 - v2 * v3) + k1t * (v0 * v0 - v2 * v2)) * ef;
    dy = gt2 * v0 * v2 * ef;
    dpy = (fg * v2 - gt2 * v0 * v3 - (g_tot + gt2) * v1 * v2 - 2.0 * k1t * v0 * v2) * ef;
    dz = ef * ef * 0.5 *
         (v2 * v2 * fg + v0 * v0 * v0 * (4.0 * k1t - gt * gt2) / 6.0 + 0.5 * v0 * v2 * v2 * (-4.0 * k1t + gt * gs2) +
          (v0 * v0 * v1 - 2.0 * v0 * v2 * v3) * gt2 - v1 * v2 * v2 * gs2);
  } else {
    dx = (gt2 * v0 * v0 - gs2 * v2 * v2) * ef / 2.0;
    dpx = (gt2 * (v2 * v3 - v0 * v1) + k1t * (v0 * v0 - v2 * v2) - gt * gt2 * (v0 * v0 + v2 * v2) / 2.0) * ef;
    dy = -gt2 * v0 * v2 * ef;
    dpy = (fg * v2 + gt2 * v0 * v3 + (g_tot + gt2) * v1 * v2 + (gt * gs2 - 2.0 * k1t) * v0 * v2) * ef;
    dz = ef * ef * 0.5 *
         (v2 * v2 * fg + v0 * v0 * v0 * (4.0 * k1t - gt * gt2) / 6.0 + 0.5 * v0 * v2 * v2 * (-4.0 * k1t + gt * gs2) -
          (v0 * v0 * v1 - 2.0 * v0 * v2 * v3) * gt2 + v1 * v2 * v2 * gs2);
  }
  L.x = v0 + dx;
  L.px = v1 + (dpx + gt * v0);
  L.y = v2 + dy;
  L.py = v3 + (dpy - gt * v2);
  L.z += dz;
}

// linear_bend_edge_kick (fringe_mod.f90:129-190): c0 = ht_x, c1 = ht_y
__device__ __forceinline__ void op_linear_edge(Lane& L, const DevOp& op) {
  L.px += op.c[0] * L.x;
  L.py += op.c[1] * L.y;
}

// tilt_coords (bmad/code/tilt_coords.f90)
__device__ __forceinline__ void op_tilt(Lane& L, const DevOp& op) {
  double c = op.c[0], s = op.c[1];
  double x = L.x, y = L.y, px = L.px, py = L.py;
  L.x = c * x + s * y;
  L.y = -s * x + c * y;
  L.px = c * px + s * py;
  L.py = -s * px + c * py;
}

// offset_particle for a straight element, set (offset_particle.f90:170-200, 255-282)
// and unset (:410-470); c: 0..2 offsets, 3 L/2, 4..12 W (row-major).
__device__ __forceinline__ void op_offset(Lane& L, const DevOp& op, bool set) {
  double rel_p = 1.0 + L.pz;
  double r0, r1, r2;
  if (set) {
    r0 = L.x - op.c[0];
    r1 = L.y - op.c[1];
    r2 = 0.0 - (op.c[2] + op.c[3]);
  } else {
    r0 = L.x;
    r1 = L.y;
    r2 = 0.0 + op.c[3];  // r(3) - ds_center with ds_center = -L/2
  }
  double s_pos;  // s_body (set) or s_lab (unset)
  if (__builtin_amdgcn_readfirstlane(op.flags) & F_OFFSET_PITCH) {
    const double* w = &op.c[4];
    double n0 = w[0] * r0 + w[1] * r1 + w[2] * r2;
    double n1 = w[3] * r0 + w[4] * r1 + w[5] * r2;
    double n2 = w[6] * r0 + w[7] * r1 + w[8] * r2;
    if (set) { r0 = n0; r1 = n1; r2 = n2 + op.c[3]; }
    else { r0 = n0 + op.c[0]; r1 = n1 + op.c[1]; r2 = n2 + (op.c[2] + op.c[3]); }
    double pzz = rel_p * rel_p - L.px * L.px - L.py * L.py;
    if (pzz <= 0.0) { L.state = BMX_LOST_PZ; return; }
    double p2 = sqrt(pzz);
    double q0 = w[0] * L.px + w[1] * L.py + w[2] * p2;
    double q1 = w[3] * L.px + w[4] * L.py + w[5] * p2;
    L.px = q0;
    L.py = q1;
  } else {
    if (set) { r2 = r2 + op.c[3]; }
    else { r0 = r0 + op.c[0]; r1 = r1 + op.c[1]; r2 = r2 + (op.c[2] + op.c[3]); }
    double pzz = rel_p * rel_p - L.px * L.px - L.py * L.py;
    if (pzz <= 0.0) { L.state = BMX_LOST_PZ; return; }
  }
  L.x = r0;
  L.y = r1;
  s_pos = r2;
  double s_target = set ? 0.0 : 2.0 * op.c[3];
  double ds = s_target - s_pos;
  if (fabs(ds) > 1e-14) op_drift_noref(L, ds);
}

// apply_energy_kick (bmad/low_level/apply_energy_kick.f90:23-74)
__device__ __forceinline__ void energy_kick(Lane& L, double dE, const LatticeConsts& lc) {
  double pc = (1.0 + L.pz) * lc.p0c;
  double beta_old = L.beta;
  double E_old = pc / L.beta;
  double E_new = E_old + dE;
  if (E_new < lc.mc2) {
    L.pz = -1.0;
    L.beta = 0.0;
    L.state = BMX_LOST_PZ;
    return;
  }
  L.pz = L.pz + (1.0 + L.pz) * d_sqrt_one((2.0 * E_old * dE + dE * dE) / (pc * pc));
  double pc_new = lc.p0c * (1.0 + L.pz);
  double beta_new = pc_new / E_new;
  L.z = L.z * beta_new / beta_old;
  L.beta = beta_new;
  refresh_invariants(L, lc);
}

// track_a_rfcavity body (bmad/low_level/track_a_rfcavity.f90:60-140), relative time:
// particle_rf_time = -z/(beta c) (bmad/code/particle_rf_time.f90:89).
// c: 0 L, 1 voltage, 2 phi0 (+ phi0_multipass), 3 rf_frequency, 4 1/(2 f); n = n_slice
__device__ void op_rfcavity(Lane& L, const DevOp& op, const LatticeConsts& lc) {
  double len = op.c[0], volt = op.c[1], freq = op.c[3], half_t = op.c[4];
  int n_slice = __builtin_amdgcn_readfirstlane(op.n);
  double time = -L.z / (L.beta * D_C_LIGHT);
  if (freq != 0.0 && !(time >= -half_t && time < half_t)) time = d_modulo2(time, half_t);
  double phase = d_modulo2(D_TWOPI * (op.c[2] - time * freq), D_PI);
  double dl = len / n_slice;
  for (int i = 0; i <= n_slice; i++) {
    double dE = volt * sin(phase) / n_slice;
    if (i == 0 || i == n_slice) dE = dE / 2.0;
    energy_kick(L, dE, lc);
    if (L.state != BMX_ALIVE) return;
    if (L.pz == -1.0) { L.state = BMX_LOST_PZ; return; }
    if (i != n_slice && dl != 0.0) {
      double z0 = L.z;
      op_drift(L, dl);
      if (L.state != BMX_ALIVE) return;
      phase = phase + D_TWOPI * freq * (L.z - z0) / (D_C_LIGHT * L.beta);
    }
  }
}

// orbit_too_large (bmad/low_level/orbit_too_large.f90:26-120).  A straight-
// line screen of all four conditions runs first; only lanes that fail it take
// the reference's ordered tests, which pick the loss state.
__device__ __forceinline__ bool too_large(Lane& L, const LatticeConsts& lc) {
  double mal = lc.max_aperture_limit;
  double rel_p = 1.0 + L.pz;
  double f = L.px * L.px + L.py * L.py - rel_p * rel_p;
  bool bad = (fabs(L.x) > mal) | (fabs(L.y) > mal) | (rel_p < 0.0) | (f > 0.0);
  if (!bad) return false;
  if (L.x > mal) { L.state = BMX_LOST_POS_X; return true; }
  if (-L.x > mal) { L.state = BMX_LOST_NEG_X; return true; }
  if (L.y > mal) { L.state = BMX_LOST_POS_Y; return true; }
  if (-L.y > mal) { L.state = BMX_LOST_NEG_Y; return true; }
  if (rel_p < 0.0) { L.state = BMX_LOST_PZ; return true; }
  if (fabs(L.px) > fabs(L.py)) L.state = (L.px > 0.0) ? BMX_LOST_POS_X : BMX_LOST_NEG_X;
  else L.state = (L.py > 0.0) ? BMX_LOST_POS_Y : BMX_LOST_NEG_Y;
  return true;
}

// aperture part of check_aperture_limit (bmad/low_level/check_aperture_limit.f90:257-309)
// c, rectangular: 0 x1_lim, 1 x2_lim, 2 y1_lim, 3 y2_lim;
// elliptical: 0 cx, 1 cy, 2 1/wx, 3 1/wy, 4 wx, 5 wy
__device__ __forceinline__ void aperture(Lane& L, const double* c, bool elliptical) {
  if (elliptical) {
    double xp = L.x - c[0], yp = L.y - c[1];
    // screen with the reciprocal widths: it is within a few ulp of the
    // reference's r, so r_fast < 1 - 1e-13 decides r < 1 exactly
    double fx = xp * c[2], fy = yp * c[3];
    if (fx * fx + fy * fy < 1.0 - 1e-13) return;
    double ax = xp / c[4], ay = yp / c[5];
    double r = ax * ax + ay * ay;
    if (r < 1.0) return;
    if (fabs(ax) > fabs(ay)) L.state = (xp > 0.0) ? BMX_LOST_POS_X : BMX_LOST_NEG_X;
    else L.state = (yp > 0.0) ? BMX_LOST_POS_Y : BMX_LOST_NEG_Y;
    return;
  }
  double x1 = c[0], x2 = c[1], y1 = c[2], y2 = c[3], uf = 0.0, f;
  if (x1 != 0.0 && L.x < x1) { f = fabs((L.x - x1) / (x2 - x1)); if (f > uf) { L.state = BMX_LOST_NEG_X; uf = f; } }
  if (x2 != 0.0 && L.x > x2) { f = fabs((L.x - x2) / (x2 - x1)); if (f > uf) { L.state = BMX_LOST_POS_X; uf = f; } }
  if (y1 != 0.0 && L.y < y1) { f = fabs((L.y - y1) / (y2 - y1)); if (f > uf) { L.state = BMX_LOST_NEG_Y; uf = f; } }
  if (y2 != 0.0 && L.y > y2) { f = fabs((L.y - y2) / (y2 - y1)); if (f > uf) { L.state = BMX_LOST_POS_Y; uf = f; } }
}

// Hwang edge kick with e = 0 and no fringe integral (fringe_mod.f90:240-418
// with tan(e) = 0, sec(e) = 1, fint_gap = 0): of the polynomial only the gs2
// and g_tot terms remain, gs2 = g_tot.  gs = g_tot at the entrance, -g_tot at
// the exit (the two edges differ only in those terms' signs).
__device__ __forceinline__ void hwang_simple(Lane& L, double gs) {
  double ef = L.irp, v1 = L.px, v2 = L.y;
  double a = gs * v2 * v2;
  L.x += a * ef * 0.5;
  L.py -= gs * v1 * v2 * ef;
  L.z -= ef * ef * 0.5 * v1 * a;
}

// Sector bend.  OP_BEND: the body step alone.  OP_SBEND_FUSED: the whole
// element of track_a_bend (track_a_bend.f90:57-113) for simple edges and one
// body step: entrance edge, orbit_too_large, body, orbit_too_large, exit
// edge, returning at the first loss as the reference does.
// c: 0 .. 10 as op_bend, 11 entrance-edge g_tot, 12 minus the exit-edge g_tot
__device__ __forceinline__ void op_sbend(Lane& L, const DevOp& op, int flags, bool fused, const LatticeConsts& lc) {
  if (fused) {
    if (flags & F_HWANG_ENTER) hwang_simple(L, op.c[11]);
    if (too_large(L, lc)) return;
  }
  op_bend(L, op, lc);
  if (fused) {
    if (L.state != BMX_ALIVE || too_large(L, lc)) return;
    if (flags & F_EDGE_EXIT) hwang_simple(L, op.c[12]);
  }
}

// Element end (an OP_ELE_END, or folded into the element's last op):
// orbit_too_large, then the exit aperture, for a particle still alive.  A
// straight-line screen decides "alive, inside every limit" for the whole
// wave; only when some lane fails it do the reference's ordered tests run.
// The elliptical screen is the one of aperture(); a NaN coordinate fails
// every screen and takes the ordered tests, as in the reference.
// c: the aperture constants (as aperture()).
__device__ __forceinline__ void ele_end(Lane& L, const double* c, int flags, const LatticeConsts& lc) {
  double mal = lc.max_aperture_limit;
  double rel_p = 1.0 + L.pz;
  double f = L.px * L.px + L.py * L.py - rel_p * rel_p;
  bool inside = (fabs(L.x) <= mal) & (fabs(L.y) <= mal) & (rel_p >= 0.0) & (f <= 0.0);
  if (flags & F_END_APERTURE) {
    if (flags & F_END_ELLIPTICAL) {
      double fx = (L.x - c[0]) * c[2], fy = (L.y - c[1]) * c[3];
      inside &= fx * fx + fy * fy < 1.0 - 1e-13;
    } else {
      inside &= (c[0] == 0.0 || L.x >= c[0]) & (c[1] == 0.0 || L.x <= c[1]) & (c[2] == 0.0 || L.y >= c[2]) &
                (c[3] == 0.0 || L.y <= c[3]);
    }
  }
  bool check = (L.state == BMX_ALIVE) & !inside;
  if (wave_any(check)) {
    if (check && !too_large(L, lc) && (flags & F_END_APERTURE)) aperture(L, c, flags & F_END_ELLIPTICAL);
  }
}

// ---------------------------------------------------------------------------
// one op on one lane
// ---------------------------------------------------------------------------
// The dispatch is a run of independent, wave-uniform if-blocks on a one-hot
// type mask rather than a switch: a switch becomes a branch tree whose join
// blocks copy the whole lane state between registers on every op, while each
// if-block joins back with the lane state in place.  The mask comes through
// readfirstlane, which also keeps the compiler from folding the tests back
// into a switch.  The tests are grouped so that an op passes about half of
// them.
#define TB(t) (1u << (t))
__device__ __forceinline__ void run_op(Lane& L, const DevOp& op, int type, int flags, const LatticeConsts& lc) {
  unsigned tb = (unsigned)__builtin_amdgcn_readfirstlane((int)(1u << type));
  if (tb & (TB(OP_DRIFT) | TB(OP_KICK) | TB(OP_KD_SLICES))) {
    if (tb & TB(OP_DRIFT)) op_drift(L, op.c[0]);
    if (tb & TB(OP_KICK)) op_kick(L, op);
    if (tb & TB(OP_KD_SLICES)) op_kd_slices(L, op);
  }
  if (tb & (TB(OP_QUAD) | TB(OP_BEND) | TB(OP_SBEND_FUSED))) {
    if (tb & TB(OP_QUAD)) op_quad(L, op, lc);
    if (tb & (TB(OP_BEND) | TB(OP_SBEND_FUSED))) op_sbend(L, op, flags, (tb & TB(OP_SBEND_FUSED)) != 0, lc);
  }
  if (tb & (TB(OP_HWANG) | TB(OP_RFCAVITY))) {
    if (tb & TB(OP_HWANG)) op_hwang(L, op);
    if (tb & TB(OP_RFCAVITY)) op_rfcavity(L, op, lc);
  }
  constexpr unsigned RARE = TB(OP_KICK_REF0) | TB(OP_BEND_K1) | TB(OP_BEND_KICK) | TB(OP_LINEAR_EDGE) | TB(OP_TILT) |
                            TB(OP_OFFSET_SET) | TB(OP_OFFSET_UNSET) | TB(OP_TOO_LARGE) | TB(OP_APERTURE_IN);
  if (tb & RARE) {
    if (tb & TB(OP_KICK_REF0)) op_kick_ref0(L, op);
    if (tb & TB(OP_BEND_K1)) op_bend_k1(L, op, lc);
    if (tb & TB(OP_BEND_KICK)) op_bend_kick(L, op);
    if (tb & TB(OP_LINEAR_EDGE)) op_linear_edge(L, op);
    if (tb & TB(OP_TILT)) op_tilt(L, op);
    if (tb & TB(OP_OFFSET_SET)) op_offset(L, op, true);
    if (tb & TB(OP_OFFSET_UNSET)) op_offset(L, op, false);
    if (tb & TB(OP_TOO_LARGE)) too_large(L, lc);
    if (tb & TB(OP_APERTURE_IN)) aperture(L, op.c, flags & F_APERTURE_ELLIPTICAL);
  }
  if (flags & F_SHIFT_OUT) {  // offset_particle unset, pure translation (op_offset without pitch)
    uniform_block();
    double rel_p = 1.0 + L.pz;
    double pzz = rel_p * rel_p - L.px * L.px - L.py * L.py;
    if (L.state == BMX_ALIVE) {
      if (pzz <= 0.0) {
        L.state = BMX_LOST_PZ;
      } else {
        L.x += op.c[SHIFT_C];
        L.y += op.c[SHIFT_C + 1];
      }
    }
  }
  if (flags & F_END) ele_end(L, op.c + END_C, flags, lc);
}
#undef TB

// Element programs are staged through LDS: the workgroup copies OPS_CHUNK
// records at a time and every wave reads them at wave-uniform LDS
// addresses, instead of each CU fetching every record from L2 through its
// scalar cache on the dependent path of the element loop.  256 records are
// 32 KiB: four workgroups of 256 lanes (occupancy 4) fit in a CU's 160 KiB.
#ifndef BMX_OPS_CHUNK
#define BMX_OPS_CHUNK 256
#endif
constexpr int OPS_CHUNK = BMX_OPS_CHUNK;

// Particles per lane.  Each lane carries NP particles through the element
// program together, so the op dispatch, the LDS reads of the op constants
// and the loop control are paid once for NP particles, and the NP particle
// updates of an op sit in one basic block where they can be interleaved.
#ifndef BMX_NP
#define BMX_NP 1
#endif

// The particles of one lane.  live[j]: still tracked.  A particle is lost at
// the op whose step set its state: on_loss() gets its coordinates at that
// moment, then it is parked at the reference orbit (coordinates zero, alive
// state) so that it keeps taking the ordinary paths of every op while the
// lane's other particles go on; its results are never read again.
template <int NP>
struct Pack {
  Lane L[NP];
  bool live[NP];
  int ie_lost[NP], lost_turn[NP];
  __device__ bool any() const {
    bool a = false;
    for (int j = 0; j < NP; j++) a |= live[j];
    return a;
  }
};

__device__ __forceinline__ void park(Lane& L, const LatticeConsts& lc) {
  L.x = 0.0; L.px = 0.0; L.y = 0.0; L.py = 0.0; L.z = 0.0; L.pz = 0.0;
  L.beta = lc.beta0;
  L.irp = 1.0;
  L.bb = 0.0;
  L.state = BMX_ALIVE;
}

// Walk of `turns` turns over ops [lo1, hi1) then [lo2, hi2) for one lane's
// particles.  The whole workgroup calls it together (it holds barriers); the
// element program comes through LDS one chunk at a time.  Lanes without a
// live particle sit out each chunk; a lane whose last particle is lost
// mid-chunk carries the parked particle to the end of the chunk (an exit
// test per op costs more than that).
template <int NP, class OnLoss>
__device__ __forceinline__ void walk(Pack<NP>& P, const DevOp* __restrict__ ops, int lo1, int hi1, int lo2, int hi2,
                                     int turns, const LatticeConsts& lc, DevOp* sops, OnLoss on_loss) {
  for (int t = 0; t < turns; t++) {
    for (int seg = 0; seg < 2; seg++) {
      int lo = seg ? lo2 : lo1, hi = seg ? hi2 : hi1;
      for (int c = lo; c < hi; c += OPS_CHUNK) {
        int m = min(OPS_CHUNK, hi - c);
        // barrier: the previous chunk is consumed; leave together once the workgroup has no live particle
        if (!__syncthreads_or(P.any())) return;
        const uint4* src = reinterpret_cast<const uint4*>(ops + c);
        uint4* dst = reinterpret_cast<uint4*>(sops);
        constexpr int W = (int)(sizeof(DevOp) / sizeof(uint4));
        for (int e = threadIdx.x; e < m * W; e += blockDim.x) dst[e] = src[e];
        __syncthreads();
        if (P.any()) {
          for (int k = 0; k < m; k++) {
            const DevOp& op = sops[k];
            int type = __builtin_amdgcn_readfirstlane(op.type);
            int flags = __builtin_amdgcn_readfirstlane(op.flags);
#pragma unroll
            for (int j = 0; j < NP; j++) run_op(P.L[j], op, type, flags, lc);
#pragma unroll
            for (int j = 0; j < NP; j++) {
              if (P.L[j].state != BMX_ALIVE) {
                if (P.live[j]) {
                  P.live[j] = false;
                  P.ie_lost[j] = op.ie;
                  P.lost_turn[j] = t;
                  on_loss(j, P.L[j], t, op.ie);
                }
                park(P.L[j], lc);
              }
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
struct SoA {
  double *x, *px, *y, *py, *z, *pz, *beta;
  int32_t *state, *ix_lost, *turn_lost;
};

// Waves per SIMD the register allocation targets (occupancy); the lane
// program is FP64-latency bound, so occupancy is traded against spills.
#ifndef BMX_WAVES_PER_EU
#define BMX_WAVES_PER_EU (BMX_NP == 1 ? 4 : 2)
#endif

// NP particles of a bunch per lane: particle j of lane `tid` in workgroup
// `blk` is base + j * 256 + tid (coalesced per j).  Load, `n_turns` turns
// over [lo1, hi1) then [lo2, hi2), store.  Particles already lost on entry are
// left untouched; a particle lost on the way is stored at its loss.
template <int NP>
__device__ __forceinline__ void track_lanes(int64_t base, int64_t end, const DevOp* __restrict__ ops, int lo1,
                                            int hi1, int lo2, int hi2, const LatticeConsts& lc, int n_turns, SoA& p,
                                            unsigned long long* __restrict__ n_lost, DevOp* sops) {
  Pack<NP> P;
  int64_t idx[NP];
  bool entered[NP];
#pragma unroll
  for (int j = 0; j < NP; j++) {
    int64_t i = base + (int64_t)j * blockDim.x + threadIdx.x;
    idx[j] = i;
    entered[j] = i < end && p.state[i] == BMX_ALIVE;
    P.live[j] = false;
    P.ie_lost[j] = -1;
    P.lost_turn[j] = 0;
    park(P.L[j], lc);
    if (entered[j]) {
      Lane& L = P.L[j];
      L.x = p.x[i]; L.px = p.px[i]; L.y = p.y[i]; L.py = p.py[i]; L.z = p.z[i]; L.pz = p.pz[i]; L.beta = p.beta[i];
      refresh_invariants(L, lc);
      P.live[j] = !too_large(L, lc);
      if (!P.live[j]) {  // lost before the first element
        p.x[i] = L.x; p.px[i] = L.px; p.y[i] = L.y; p.py[i] = L.py; p.z[i] = L.z; p.pz[i] = L.pz; p.beta[i] = L.beta;
        p.state[i] = L.state;
        if (p.ix_lost) p.ix_lost[i] = ops[lo1].ie;
        if (p.turn_lost) p.turn_lost[i] = 0;
        atomicAdd(n_lost, 1ull);
        park(L, lc);
      }
    }
  }
  walk<NP>(P, ops, lo1, hi1, lo2, hi2, n_turns, lc, sops, [&](int j, const Lane& L, int t, int ie) {
    int64_t i = idx[j];
    p.x[i] = L.x; p.px[i] = L.px; p.y[i] = L.y; p.py[i] = L.py; p.z[i] = L.z; p.pz[i] = L.pz; p.beta[i] = L.beta;
    p.state[i] = L.state;
    if (p.ix_lost) p.ix_lost[i] = ie;
    if (p.turn_lost) p.turn_lost[i] = t;
    atomicAdd(n_lost, 1ull);
  });
#pragma unroll
  for (int j = 0; j < NP; j++) {
    if (!P.live[j]) continue;
    int64_t i = idx[j];
    const Lane& L = P.L[j];
    p.x[i] = L.x; p.px[i] = L.px; p.y[i] = L.y; p.py[i] = L.py; p.z[i] = L.z; p.pz[i] = L.pz; p.beta[i] = L.beta;
  }
}

// Multi-turn bunch tracking: segments [lo1, hi1) then [lo2, hi2) per turn.
__global__ __launch_bounds__(256, BMX_WAVES_PER_EU) void bmx_track_bunch_kernel(
    const DevOp* __restrict__ ops, int lo1, int hi1, int lo2, int hi2, LatticeConsts lc, int n_turns, SoA p,
    int64_t n, unsigned long long* __restrict__ n_lost) {
  __shared__ DevOp sops[OPS_CHUNK];
  int64_t base = (int64_t)blockIdx.x * blockDim.x * BMX_NP;
  track_lanes<BMX_NP>(base, n, ops, lo1, hi1, lo2, hi2, lc, n_turns, p, n_lost, sops);
}

// Particles of several variants in one launch: particles [off[v], off[v+1])
// are tracked through variant v (blockIdx.y = v, whole turns).
__global__ __launch_bounds__(256, BMX_WAVES_PER_EU) void bmx_track_multi_kernel(
    const DevOp* __restrict__ ops, const int32_t* __restrict__ var_op_begin, const int64_t* __restrict__ off,
    LatticeConsts lc, int n_turns, SoA p, unsigned long long* __restrict__ n_lost) {
  __shared__ DevOp sops[OPS_CHUNK];
  int v = blockIdx.y;
  int64_t base = off[v] + (int64_t)blockIdx.x * blockDim.x * BMX_NP;
  track_lanes<BMX_NP>(base, off[v + 1], ops, var_op_begin[v], var_op_begin[v + 1], 0, 0, lc, n_turns, p, n_lost,
                      sops);
}

// Batched grid tracking in turn blocks with lane compaction.  Particles of
// variant v occupy slots [v * n_ic, v * n_ic + count[v]) of a structure-of-
// arrays store; blockIdx.y = variant, so every wave walks one variant's
// program.  A block tracks up to `turns` turns; survivors are appended to the
// destination store through one atomic per wave (ballot + lane prefix), lost
// particles write their survival turn.  The next block launches only over the
// survivors, so waves stay full as the grid loses particles.
struct GridDev {
  int n_angle, n_amp, n_pz, n_ic;
  double amp_max_x, amp_ratio_y, pz_min, pz_max, angle_min, angle_max;
};

struct GridSoA {
  double *x, *px, *y, *py, *z, *pz, *beta;
  int32_t* ic;
};

__global__ __launch_bounds__(256) void bmx_grid_init_kernel(GridDev g, LatticeConsts lc, GridSoA s,
                                                           const double* __restrict__ ref_orbit,
                                                           int32_t* __restrict__ survival, int n_turns) {
  int v = blockIdx.y;
  int ic = blockIdx.x * blockDim.x + threadIdx.x;
  if (ic >= g.n_ic) return;
  int i_amp = ic % g.n_amp;
  int i_ang = (ic / g.n_amp) % g.n_angle;
  int i_pz = ic / (g.n_amp * g.n_angle);
  double amp = g.amp_max_x * (double)(i_amp + 1) / (double)g.n_amp;
  double th = (g.n_angle > 1) ? g.angle_min + (g.angle_max - g.angle_min) * (double)i_ang / (double)(g.n_angle - 1)
                              : g.angle_min;
  double pz0 = (g.n_pz > 1) ? g.pz_min + (g.pz_max - g.pz_min) * (double)i_pz / (double)(g.n_pz - 1) : g.pz_min;
  double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, pz0};
  if (ref_orbit) {  // offsets are taken from the reference orbit of this variant and pz slice
    const double* o = ref_orbit + ((int64_t)v * g.n_pz + i_pz) * 6;
    for (int k = 0; k < 6; k++) r[k] = o[k];
  }
  int64_t k = (int64_t)v * g.n_ic + ic;
  s.x[k] = r[0] + amp * cos(th);
  s.y[k] = r[2] + amp * g.amp_ratio_y * sin(th);
  s.px[k] = r[1]; s.py[k] = r[3]; s.z[k] = r[4]; s.pz[k] = r[5];
  double pc = lc.p0c * (1.0 + r[5]);
  s.beta[k] = (r[5] == 0.0) ? lc.beta0 : pc / sqrt(pc * pc + lc.mc2sq);
  s.ic[k] = ic;
  survival[k] = n_turns;
}

// Explicit initial conditions already placed in slots [v * stride, v * stride + count[v]):
// beta from pz (init_coord), slot ids, survival = n_turns.
__global__ __launch_bounds__(256) void bmx_ics_init_kernel(LatticeConsts lc, GridSoA s, const int32_t* __restrict__ count,
                                                          int stride, int32_t* __restrict__ survival, int n_turns) {
  int v = blockIdx.y;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count[v]) return;
  int64_t k = (int64_t)v * stride + i;
  double pz = s.pz[k];
  double pc = lc.p0c * (1.0 + pz);
  s.beta[k] = (pz == 0.0) ? lc.beta0 : pc / sqrt(pc * pc + lc.mc2sq);
  s.ic[k] = i;
  survival[k] = n_turns;
}

__global__ __launch_bounds__(256, BMX_WAVES_PER_EU) void bmx_grid_block_kernel(
    const DevOp* __restrict__ ops, const int32_t* __restrict__ var_op_begin, const int32_t* __restrict__ var_n_ele,
    LatticeConsts lc, int n_ic, GridSoA src, GridSoA dst, const int32_t* __restrict__ count_in,
    int32_t* __restrict__ count_out, const int32_t* __restrict__ d_turn0, int turns_per_block, int n_turns,
    int32_t* __restrict__ survival, unsigned long long* __restrict__ pet) {
  __shared__ DevOp sops[OPS_CHUNK];
  // the block's first turn lives in device memory so that captured graphs of
  // several blocks replay unchanged; every thread reads the same value, so a
  // finished scan leaves the whole grid together
  const int turn0 = *d_turn0;
  const int turns = min(turns_per_block, n_turns - turn0);
  if (turns <= 0) return;
  constexpr int NP = BMX_NP;
  int v = blockIdx.y;
  int64_t base = (int64_t)v * n_ic;
  int n_in = count_in[v];
  Pack<NP> P;
  int slot[NP], ic[NP];
  bool active[NP];
#pragma unroll
  for (int j = 0; j < NP; j++) {
    slot[j] = (blockIdx.x * NP + j) * blockDim.x + threadIdx.x;
    active[j] = slot[j] < n_in;
    ic[j] = 0;
    P.live[j] = false;
    P.ie_lost[j] = -1;
    P.lost_turn[j] = 0;
    park(P.L[j], lc);
    if (active[j]) {
      int64_t k = base + slot[j];
      Lane& L = P.L[j];
      L.x = src.x[k]; L.px = src.px[k]; L.y = src.y[k]; L.py = src.py[k]; L.z = src.z[k]; L.pz = src.pz[k];
      L.beta = src.beta[k];
      ic[j] = src.ic[k];
      refresh_invariants(L, lc);
      P.live[j] = !(turn0 == 0 && too_large(L, lc));
      if (!P.live[j]) {
        survival[base + ic[j]] = turn0;
        park(L, lc);
      }
    }
  }
  int lo = var_op_begin[v], hi = var_op_begin[v + 1];
  walk<NP>(P, ops, lo, hi, 0, 0, turns, lc, sops,
           [&](int j, const Lane&, int t, int) { survival[base + ic[j]] = turn0 + t; });
  // live particle-element-turns of this lane
  unsigned long long my_pet = 0;
#pragma unroll
  for (int j = 0; j < NP; j++)
    if (active[j]) {
      unsigned long long tr = P.live[j] ? (unsigned long long)turns : (unsigned long long)P.lost_turn[j];
      my_pet += tr * (unsigned long long)var_n_ele[v] + (P.live[j] ? 0ull : (unsigned long long)(P.ie_lost[j] + 1));
    }
  // wave-level compaction of survivors, one ballot per particle slot
  int lane = __lane_id();
#pragma unroll
  for (int j = 0; j < NP; j++) {
    bool keep = active[j] && P.live[j];
    unsigned long long mask = __ballot(keep);
    unsigned long long below = (lane == 0) ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    int n_alive = __popcll(mask);
    int first = __ffsll((long long)mask) - 1;
    int slot0 = 0;
    if (n_alive > 0 && lane == first) slot0 = atomicAdd(&count_out[v], n_alive);
    slot0 = __shfl(slot0, first < 0 ? 0 : first);
    if (keep) {
      int64_t k = base + slot0 + __popcll(below);
      const Lane& L = P.L[j];
      dst.x[k] = L.x; dst.px[k] = L.px; dst.y[k] = L.y; dst.py[k] = L.py; dst.z[k] = L.z; dst.pz[k] = L.pz;
      dst.beta[k] = L.beta;
      dst.ic[k] = ic[j];
    }
  }
  for (int off = 32; off > 0; off >>= 1) my_pet += __shfl_xor(my_pet, off);
  if (lane == 0 && my_pet) atomicAdd(pet, my_pet);
}

// Step bookkeeping between turn blocks: clears the count buffer the next
// block appends survivors to and advances the block's first turn.  A kernel,
// not a memset, so that it can sit in a captured graph.
__global__ __launch_bounds__(256) void bmx_grid_advance_kernel(int32_t* __restrict__ count_next, int nv,
                                                              int32_t* __restrict__ d_turn0, int turns) {
  for (int v = threadIdx.x; v < nv; v += blockDim.x) count_next[v] = 0;
  if (threadIdx.x == 0) *d_turn0 += turns;
}

// Bunch moments over live particles (the sums of ltt_calc_bunch_data,
// bsim/modules/lt_tracking_mod.f90:2391-2402): out[0] = n_live,
// out[1..6] = sum vec_i, out[7..27] = sum vec_i vec_j (i <= j, row-major upper
// triangle).  Two passes with a fixed reduction order, so the sums do not
// depend on scheduling: per-block partials, then one block over the partials.
constexpr int N_MOM = 28;

__device__ __forceinline__ void block_reduce_moments(double (&a)[N_MOM], double* red) {
  // red: N_MOM x 256 doubles of LDS
  int t = threadIdx.x;
  for (int k = 0; k < N_MOM; k++) red[k * 256 + t] = a[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < N_MOM; k++) red[k * 256 + t] += red[k * 256 + t + w];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void bmx_moments_partial_kernel(SoA p, int64_t n, double* __restrict__ partial) {
  extern __shared__ double red[];
  double a[N_MOM];
  for (int k = 0; k < N_MOM; k++) a[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    if (p.state[i] != BMX_ALIVE) continue;
    double v[6] = {p.x[i], p.px[i], p.y[i], p.py[i], p.z[i], p.pz[i]};
    a[0] += 1.0;
    int k = 7;
    for (int r = 0; r < 6; r++) {
      a[1 + r] += v[r];
      for (int c = r; c < 6; c++) a[k++] += v[r] * v[c];
    }
  }
  block_reduce_moments(a, red);
  if (threadIdx.x < N_MOM) partial[(int64_t)blockIdx.x * N_MOM + threadIdx.x] = red[threadIdx.x * 256];
}

__global__ __launch_bounds__(256) void bmx_moments_final_kernel(const double* __restrict__ partial, int n_part,
                                                               double* __restrict__ out) {
  extern __shared__ double red[];
  double a[N_MOM];
  for (int k = 0; k < N_MOM; k++) a[k] = 0.0;
  for (int b = threadIdx.x; b < n_part; b += 256)
    for (int k = 0; k < N_MOM; k++) a[k] += partial[(int64_t)b * N_MOM + k];
  block_reduce_moments(a, red);
  if (threadIdx.x < N_MOM) out[threadIdx.x] = red[threadIdx.x * 256];
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct bmx_ctx {
  int device;
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  std::string err;
  bmx_schedule sched;  // turn-block schedule used when a call passes turns_per_block <= 0
};

struct bmx_lattice {
  int device;
  int n_variants;
  std::vector<int32_t> var_op_begin;            // n_variants + 1
  std::vector<std::vector<int32_t>> ele_op_begin;  // per variant, n_ele + 1 (absolute op indices)
  std::vector<int32_t> var_n_ele;
  DevOp* d_ops;
  int32_t* d_var_op_begin;
  int32_t* d_var_n_ele;
  int64_t n_ops;
  LatticeConsts lc;
};

static int set_err(bmx_ctx* ctx, const std::string& s) {
  if (ctx) ctx->err = s;
  return -1;
}

#define HIP_TRY(ctx, call)                                                              \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return set_err((ctx), std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Shared with bmx_mlp.hip (hidden: not part of the exported ABI).
extern "C" __attribute__((visibility("hidden"))) hipStream_t bmx_ctx_stream(const bmx_ctx* ctx) { return ctx->stream; }
// defined in bmx_mlp.hip: frees the context's GEMM workspace
extern "C" void bmx_mlp_release(const bmx_ctx* ctx);
extern "C" __attribute__((visibility("hidden"))) int32_t bmx_ctx_set_error(bmx_ctx* ctx, const char* msg) {
  return set_err(ctx, msg ? msg : "error");
}

extern "C" {

int32_t bmx_abi_version(void) { return BMX_ABI_VERSION; }

int32_t bmx_sizeof(int32_t which) {
  switch (which) {
    case 0: return (int32_t)sizeof(bmx_ele_rec);
    case 1: return (int32_t)sizeof(bmx_particles);
    case 2: return (int32_t)sizeof(bmx_stats);
    case 3: return (int32_t)sizeof(bmx_grid);
    case 4: return (int32_t)sizeof(bmx_config);
    case 5: return (int32_t)sizeof(bmx_schedule);
  }
  return -1;
}

int32_t bmx_ctx_create(bmx_ctx** out, int32_t device, uint32_t flags) {
  (void)flags;
  if (!out) return -1;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return -2;
  if (device < 0 || device >= n) return -3;
  if (hipSetDevice(device) != hipSuccess) return -4;
  bmx_ctx* c = new bmx_ctx();
  c->device = device;
  c->sched.turns_per_block = 32;
  c->sched.graph_depth = 1;
  c->sched.probe_pet_per_s = 0.0;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return -5;
  }
  *out = c;
  return 0;
}

void bmx_ctx_destroy(bmx_ctx* ctx) {
  if (!ctx) return;
  bmx_mlp_release(ctx);
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* bmx_last_error(const bmx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int32_t bmx_lattice_upload(bmx_ctx* ctx, const bmx_ele_rec* eles, int32_t n_ele_total, const int32_t* variant_offsets,
                           int32_t n_variants, const bmx_ref* ref, const bmx_config* cfg, bmx_lattice** out) {
  if (!ctx || !eles || !ref || !out || n_ele_total <= 0 || n_variants <= 0) return set_err(ctx, "invalid arguments");
  *out = nullptr;
  bmx_config c = cfg ? *cfg : bmx_config{1, 0, 1e3};
  std::vector<int32_t> offs;
  if (variant_offsets) offs.assign(variant_offsets, variant_offsets + n_variants + 1);
  else if (n_variants == 1) offs = {0, n_ele_total};
  else return set_err(ctx, "variant_offsets is required for n_variants > 1");
  for (int v = 0; v < n_variants; v++)
    if (offs[v] < 0 || offs[v + 1] <= offs[v] || offs[v + 1] > n_ele_total) return set_err(ctx, "bad variant_offsets");
  bmx_lattice* lat = new bmx_lattice();
  lat->device = ctx->device;
  lat->n_variants = n_variants;
  std::vector<DevOp> prog;
  for (int v = 0; v < n_variants; v++) {
    lat->var_op_begin.push_back((int32_t)prog.size());
    std::vector<int32_t> eb;
    LatticeConsts lc;
    std::memset(&lc, 0, sizeof(lc));
    std::string err;
    if (!compile_variant(eles + offs[v], offs[v + 1] - offs[v], *ref, c, prog, eb, &lc, &err)) {
      delete lat;
      return set_err(ctx, "variant " + std::to_string(v) + ": " + err);
    }
    lc.max_aperture_limit = c.max_aperture_limit;
    lc.two_max_ap = 2.0 * c.max_aperture_limit;
    lc.aperture_limit_on = c.aperture_limit_on;
    if (v == 0) lat->lc = lc;
    else if (lc.p0c != lat->lc.p0c || lc.mc2 != lat->lc.mc2) {
      delete lat;
      return set_err(ctx, "all variants must share the reference energy and species");
    }
    lat->ele_op_begin.push_back(eb);
    lat->var_n_ele.push_back(offs[v + 1] - offs[v]);
  }
  lat->var_op_begin.push_back((int32_t)prog.size());
  lat->n_ops = (int64_t)prog.size();
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (hipMalloc(&lat->d_ops, prog.size() * sizeof(DevOp)) != hipSuccess ||
      hipMalloc(&lat->d_var_op_begin, lat->var_op_begin.size() * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&lat->d_var_n_ele, lat->var_n_ele.size() * sizeof(int32_t)) != hipSuccess) {
    delete lat;
    return set_err(ctx, "hipMalloc of the element program failed");
  }
  HIP_TRY(ctx, hipMemcpy(lat->d_ops, prog.data(), prog.size() * sizeof(DevOp), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(lat->d_var_op_begin, lat->var_op_begin.data(), lat->var_op_begin.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(lat->d_var_n_ele, lat->var_n_ele.data(), lat->var_n_ele.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
  *out = lat;
  return 0;
}

void bmx_lattice_destroy(bmx_lattice* lat) {
  if (!lat) return;
  hipSetDevice(lat->device);
  hipFree(lat->d_ops);
  hipFree(lat->d_var_op_begin);
  hipFree(lat->d_var_n_ele);
  delete lat;
}

}  // extern "C"

// Device views of a bmx_particles: host arrays are uploaded into buffers the
// call owns; results go back and the buffers are freed in finish_particles.
struct Staged {
  SoA s;
  void* dp[10];
  void* hp[10];
  size_t by[10];
  std::vector<void*> owned;
  unsigned long long* d_lost = nullptr;
};

static int stage_particles(bmx_ctx* ctx, const bmx_particles* p, Staged& st) {
  int64_t n = p->n;
  size_t bd = (size_t)n * sizeof(double), bi = (size_t)n * sizeof(int32_t);
  void* hp[10] = {p->x, p->px, p->y, p->py, p->z, p->pz, p->beta, p->state, p->ix_lost, p->turn_lost};
  size_t by[10] = {bd, bd, bd, bd, bd, bd, bd, bi, bi, bi};
  int rc = 0;
  for (int k = 0; k < 10; k++) {
    st.hp[k] = hp[k];
    st.by[k] = by[k];
    st.dp[k] = nullptr;
    if (rc || !hp[k]) continue;
    if (p->on_device) { st.dp[k] = hp[k]; continue; }
    if (hipMalloc(&st.dp[k], by[k]) != hipSuccess) { rc = -1; continue; }
    st.owned.push_back(st.dp[k]);
    if (hipMemcpyAsync(st.dp[k], hp[k], by[k], hipMemcpyHostToDevice, ctx->stream) != hipSuccess) rc = -1;
  }
  if (rc == 0 && hipMalloc(&st.d_lost, sizeof(unsigned long long)) != hipSuccess) rc = -1;
  if (rc != 0) {
    for (void* q : st.owned) (void)hipFree(q);
    if (st.d_lost) (void)hipFree(st.d_lost);
    return set_err(ctx, "device allocation or upload of the particle arrays failed");
  }
  (void)hipMemsetAsync(st.d_lost, 0, sizeof(unsigned long long), ctx->stream);
  st.s.x = (double*)st.dp[0]; st.s.px = (double*)st.dp[1]; st.s.y = (double*)st.dp[2]; st.s.py = (double*)st.dp[3];
  st.s.z = (double*)st.dp[4]; st.s.pz = (double*)st.dp[5]; st.s.beta = (double*)st.dp[6];
  st.s.state = (int32_t*)st.dp[7]; st.s.ix_lost = (int32_t*)st.dp[8]; st.s.turn_lost = (int32_t*)st.dp[9];
  return 0;
}

static int finish_particles(bmx_ctx* ctx, const bmx_particles* p, Staged& st, hipError_t le, bmx_stats* out) {
  unsigned long long h_lost = 0;
  (void)hipMemcpyAsync(&h_lost, st.d_lost, sizeof(h_lost), hipMemcpyDeviceToHost, ctx->stream);
  if (!p->on_device)
    for (int k = 0; k < 10; k++)
      if (st.hp[k]) (void)hipMemcpyAsync(st.hp[k], st.dp[k], st.by[k], hipMemcpyDeviceToHost, ctx->stream);
  hipError_t se = hipStreamSynchronize(ctx->stream);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  for (void* q : st.owned) (void)hipFree(q);
  (void)hipFree(st.d_lost);
  if (le != hipSuccess) return set_err(ctx, std::string("kernel launch: ") + hipGetErrorString(le));
  if (se != hipSuccess) return set_err(ctx, std::string("stream sync: ") + hipGetErrorString(se));
  if (out) {
    out->n_lost = (int64_t)h_lost;
    out->kernel_ms = ms;
  }
  return 0;
}

static bool particles_ok(const bmx_particles* p) {
  return p->x && p->px && p->y && p->py && p->z && p->pz && p->beta && p->state;
}

extern "C" {

int32_t bmx_track_bunch(bmx_ctx* ctx, const bmx_lattice* lat, int32_t ix_start, int32_t ix_end, int32_t n_turns,
                        int32_t direction, bmx_particles* p, bmx_stats* out) {
  if (!ctx || !lat || !p) return set_err(ctx, "invalid arguments");
  if (direction != 1) return set_err(ctx, "only direction = +1 is supported on the device path");
  if (n_turns < 0) return set_err(ctx, "n_turns < 0");
  if (!particles_ok(p)) return set_err(ctx, "particle arrays x..pz, beta, state are required");
  int n_ele = lat->var_n_ele[0];
  if (ix_start < 0 || ix_start > n_ele || ix_end < 0 || ix_end > n_ele) return set_err(ctx, "ix_start/ix_end out of range");
  const std::vector<int32_t>& eb = lat->ele_op_begin[0];
  // Bmad element numbering: elements ix_start+1 .. ix_end are array entries ix_start .. ix_end-1.
  int lo1, hi1, lo2 = 0, hi2 = 0;
  if (ix_end > ix_start) { lo1 = eb[ix_start]; hi1 = eb[ix_end]; }
  else { lo1 = eb[ix_start]; hi1 = eb[n_ele]; lo2 = eb[0]; hi2 = eb[ix_end]; }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int64_t n = p->n;
  if (out) std::memset(out, 0, sizeof(*out));
  if (n <= 0 || n_turns == 0) return 0;
  Staged st;
  if (stage_particles(ctx, p, st) != 0) return -1;
  const int block = 256;
  int64_t grid = (n + block * BMX_NP - 1) / (block * BMX_NP);
  (void)hipEventRecord(ctx->ev0, ctx->stream);
  hipLaunchKernelGGL(bmx_track_bunch_kernel, dim3((unsigned)grid), dim3(block), 0, ctx->stream, lat->d_ops, lo1, hi1,
                     lo2, hi2, lat->lc, n_turns, st.s, n, st.d_lost);
  hipError_t le = hipGetLastError();
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  return finish_particles(ctx, p, st, le, out);
}

int32_t bmx_track_multi(bmx_ctx* ctx, const bmx_lattice* lat, const int64_t* part_offsets, int32_t n_turns,
                        bmx_particles* p, bmx_stats* out) {
  if (!ctx || !lat || !p || !part_offsets) return set_err(ctx, "invalid arguments");
  if (n_turns < 0) return set_err(ctx, "n_turns < 0");
  if (!particles_ok(p)) return set_err(ctx, "particle arrays x..pz, beta, state are required");
  int nv = lat->n_variants;
  if (part_offsets[0] != 0 || part_offsets[nv] != p->n) return set_err(ctx, "part_offsets must run from 0 to p->n");
  int64_t max_count = 0;
  for (int v = 0; v < nv; v++) {
    if (part_offsets[v + 1] < part_offsets[v]) return set_err(ctx, "part_offsets must be non-decreasing");
    max_count = std::max(max_count, part_offsets[v + 1] - part_offsets[v]);
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (out) std::memset(out, 0, sizeof(*out));
  if (p->n <= 0 || n_turns == 0) return 0;
  Staged st;
  if (stage_particles(ctx, p, st) != 0) return -1;
  int64_t* d_off = nullptr;
  if (hipMalloc(&d_off, (nv + 1) * sizeof(int64_t)) != hipSuccess) {
    finish_particles(ctx, p, st, hipSuccess, nullptr);
    return set_err(ctx, "hipMalloc of the variant offsets failed");
  }
  (void)hipMemcpyAsync(d_off, part_offsets, (nv + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream);
  const int block = 256;
  (void)hipEventRecord(ctx->ev0, ctx->stream);
  hipLaunchKernelGGL(bmx_track_multi_kernel,
                     dim3((unsigned)((max_count + block * BMX_NP - 1) / (block * BMX_NP)), (unsigned)nv),
                     dim3(block), 0, ctx->stream, lat->d_ops, lat->d_var_op_begin, d_off, lat->lc, n_turns, st.s,
                     st.d_lost);
  hipError_t le = hipGetLastError();
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  int rc = finish_particles(ctx, p, st, le, out);
  (void)hipFree(d_off);
  return rc;
}

}  // extern "C"

// Survival tracking of per-variant particle stores in turn blocks with lane
// compaction (bmx_grid_block_kernel), shared by the polar grid and explicit
// initial conditions.  The caller fills store 0 and the survival array through
// `init` (a launcher run on the stream after the buffers exist).
struct SurvivalRun {
  int nv = 0, stride = 0;
  GridSoA st[2];
  int32_t* d_surv = nullptr;
  int32_t* d_count = nullptr;
  int32_t* d_turn0 = nullptr;
  unsigned long long* d_pet = nullptr;
  std::vector<void*> owned;
  ~SurvivalRun() {
    for (void* q : owned) (void)hipFree(q);
  }
  bool alloc(int n_var, int n_stride) {
    nv = n_var;
    stride = n_stride;
    int64_t total = (int64_t)nv * stride;
    auto dalloc = [&](size_t bytes) -> void* {
      void* p = nullptr;
      if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
      owned.push_back(p);
      return p;
    };
    for (int b = 0; b < 2; b++) {
      double** f[7] = {&st[b].x, &st[b].px, &st[b].y, &st[b].py, &st[b].z, &st[b].pz, &st[b].beta};
      for (int k = 0; k < 7; k++)
        if (!(*f[k] = (double*)dalloc(total * sizeof(double)))) return false;
      if (!(st[b].ic = (int32_t*)dalloc(total * sizeof(int32_t)))) return false;
    }
    return (d_surv = (int32_t*)dalloc(total * sizeof(int32_t))) && (d_count = (int32_t*)dalloc(2 * nv * sizeof(int32_t))) &&
           (d_turn0 = (int32_t*)dalloc(sizeof(int32_t))) && (d_pet = (unsigned long long*)dalloc(sizeof(unsigned long long)));
  }
};

// Runs the turn blocks; counts[v] = initial particles of variant v.  Fills the
// host survival array (nv * stride), PET and timing.  Blocks alternate the
// parity p: store p -> store 1-p, counts p -> counts 1-p, then the advance
// kernel clears counts p and moves the first turn on.  graph_depth > 1
// captures that many (an even number of) steps into one hipGraph, replayed
// until the scan ends; the host reads the live counts after every replay and
// re-captures with a smaller launch once the live particles fall below 3/4 of
// the captured size.  graph_depth <= 1 launches and synchronises block by block.
static int run_survival(bmx_ctx* ctx, const bmx_lattice* lat, SurvivalRun& R, std::vector<int32_t> h_count,
                        int n_turns, int tpb, int graph_depth, int32_t* h_surv, bmx_stats* out,
                        const std::function<void()>& init) {
  const int block = 256;
  int nv = R.nv;
  if (graph_depth > 1 && (graph_depth & 1)) graph_depth++;
  (void)hipMemcpyAsync(R.d_count, h_count.data(), nv * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
  (void)hipMemsetAsync(R.d_count + nv, 0, nv * sizeof(int32_t), ctx->stream);
  (void)hipMemsetAsync(R.d_pet, 0, sizeof(unsigned long long), ctx->stream);
  (void)hipMemsetAsync(R.d_turn0, 0, sizeof(int32_t), ctx->stream);
  (void)hipEventRecord(ctx->ev0, ctx->stream);
  init();
  int cur = 0, turn0 = 0, max_live = 0;
  for (int v = 0; v < nv; v++) max_live = std::max(max_live, h_count[v]);
  auto step = [&](int p, int live) {
    dim3 gd((unsigned)((live + block * BMX_NP - 1) / (block * BMX_NP)), (unsigned)nv);
    hipLaunchKernelGGL(bmx_grid_block_kernel, gd, dim3(block), 0, ctx->stream, lat->d_ops, lat->d_var_op_begin,
                       lat->d_var_n_ele, lat->lc, R.stride, R.st[p], R.st[1 - p], R.d_count + p * nv,
                       R.d_count + (1 - p) * nv, R.d_turn0, tpb, n_turns, R.d_surv, R.d_pet);
    hipLaunchKernelGGL(bmx_grid_advance_kernel, dim3(1), dim3(256), 0, ctx->stream, R.d_count + p * nv, nv, R.d_turn0,
                       tpb);
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_live = 0;
  hipError_t le = hipGetLastError();
  while (le == hipSuccess && turn0 < n_turns && max_live > 0) {
    if (graph_depth > 1) {
      if (!exec || max_live < (graph_live * 3) / 4) {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        le = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (le != hipSuccess) break;
        for (int k = 0; k < graph_depth; k++) step((cur + k) & 1, max_live);
        le = hipStreamEndCapture(ctx->stream, &graph);
        if (le == hipSuccess) le = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (le != hipSuccess) break;
        graph_live = max_live;
      }
      le = hipGraphLaunch(exec, ctx->stream);
      if (le != hipSuccess) break;
      // an even number of steps: the next block reads counts `cur` again
    } else {
      step(cur, max_live);
      le = hipGetLastError();
      if (le != hipSuccess) break;
      cur = 1 - cur;
    }
    (void)hipMemcpyAsync(h_count.data(), R.d_count + cur * nv, nv * sizeof(int32_t), hipMemcpyDeviceToHost,
                         ctx->stream);
    (void)hipMemcpyAsync(&turn0, R.d_turn0, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) break;
    max_live = 0;
    for (int v = 0; v < nv; v++) max_live = std::max(max_live, h_count[v]);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  unsigned long long h_pet = 0;
  int64_t total = (int64_t)nv * R.stride;
  (void)hipMemcpyAsync(&h_pet, R.d_pet, sizeof(h_pet), hipMemcpyDeviceToHost, ctx->stream);
  (void)hipMemcpyAsync(h_surv, R.d_surv, total * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
  hipError_t se = hipStreamSynchronize(ctx->stream);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  if (le != hipSuccess) return set_err(ctx, std::string("turn blocks: ") + hipGetErrorString(le));
  if (se != hipSuccess) return set_err(ctx, std::string("stream sync: ") + hipGetErrorString(se));
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->kernel_ms = ms;
    out->pet = (double)h_pet;
  }
  return 0;
}

extern "C" {

int32_t bmx_track_grid(bmx_ctx* ctx, const bmx_lattice* lat, const bmx_grid* grid, const double* ref_orbit,
                       int32_t n_turns, int32_t* survival_turn, double* da_area, bmx_stats* out) {
  if (!ctx || !lat || !grid || !survival_turn) return set_err(ctx, "invalid arguments");
  if (grid->n_angle <= 0 || grid->n_amp <= 0 || grid->n_pz <= 0 || n_turns < 0) return set_err(ctx, "bad grid");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  GridDev g;
  g.n_angle = grid->n_angle; g.n_amp = grid->n_amp; g.n_pz = grid->n_pz;
  g.n_ic = grid->n_angle * grid->n_amp * grid->n_pz;
  g.amp_max_x = grid->amp_max_x; g.amp_ratio_y = grid->amp_ratio_y;
  g.pz_min = grid->pz_min; g.pz_max = grid->pz_max;
  g.angle_min = grid->angle_min; g.angle_max = grid->angle_max;
  int nv = lat->n_variants;
  int tpb = grid->turns_per_block > 0 ? grid->turns_per_block : ctx->sched.turns_per_block;
  int64_t total = (int64_t)g.n_ic * nv;
  SurvivalRun R;
  double* d_ref = nullptr;
  if (!R.alloc(nv, g.n_ic)) return set_err(ctx, "device allocation for the grid failed");
  if (ref_orbit) {
    size_t bytes = (size_t)nv * g.n_pz * 6 * sizeof(double);
    if (hipMalloc(&d_ref, bytes) != hipSuccess) return set_err(ctx, "device allocation for the grid failed");
    R.owned.push_back(d_ref);
    (void)hipMemcpyAsync(d_ref, ref_orbit, bytes, hipMemcpyHostToDevice, ctx->stream);
  }
  std::vector<int32_t> h_count(nv, g.n_ic);
  const int block = 256;
  int rc = run_survival(ctx, lat, R, h_count, n_turns, tpb, ctx->sched.graph_depth, survival_turn, out, [&]() {
    hipLaunchKernelGGL(bmx_grid_init_kernel, dim3((unsigned)((g.n_ic + block - 1) / block), (unsigned)nv), dim3(block),
                       0, ctx->stream, g, lat->lc, R.st[0], d_ref, R.d_surv, n_turns);
  });
  if (rc != 0) return rc;
  // DA area per variant: on the middle pz slice, the boundary point of each
  // angle is the (x, y) offset of the last IC of the contiguous surviving run
  // from the smallest amplitude; area = the polygon through the origin and
  // those points in angle order, 1/2 sum (x_i y_{i+1} - x_{i+1} y_i).
  if (da_area) {
    int ipz = g.n_pz / 2;
    double dth = (g.n_angle > 1) ? (g.angle_max - g.angle_min) / (g.n_angle - 1) : 0.0;
    for (int v = 0; v < nv; v++) {
      std::vector<double> bx(g.n_angle, 0.0), by(g.n_angle, 0.0);
      for (int ia = 0; ia < g.n_angle; ia++) {
        int last = -1;
        for (int k = 0; k < g.n_amp; k++) {
          int ic = (ipz * g.n_angle + ia) * g.n_amp + k;
          if (survival_turn[(int64_t)v * g.n_ic + ic] >= n_turns) last = k;
          else break;
        }
        double amp = (last < 0) ? 0.0 : g.amp_max_x * (double)(last + 1) / g.n_amp;
        double th = g.angle_min + dth * ia;
        bx[ia] = amp * cos(th);
        by[ia] = amp * g.amp_ratio_y * sin(th);
      }
      double a = 0;
      for (int ia = 0; ia + 1 < g.n_angle; ia++) a += 0.5 * (bx[ia] * by[ia + 1] - bx[ia + 1] * by[ia]);
      da_area[v] = a;
    }
  }
  if (out) {
    int64_t live = 0;
    for (int64_t k = 0; k < total; k++) live += (survival_turn[k] >= n_turns);
    out->n_live = live;
    out->n_lost = total - live;
  }
  return 0;
}

int32_t bmx_track_ics(bmx_ctx* ctx, const bmx_lattice* lat, const double* ics, const int64_t* ic_offsets,
                      int32_t n_turns, int32_t turns_per_block, int32_t* survival_turn, bmx_stats* out) {
  if (!ctx || !lat || !ics || !ic_offsets || !survival_turn || n_turns < 0) return set_err(ctx, "invalid arguments");
  int nv = lat->n_variants;
  if (ic_offsets[0] != 0) return set_err(ctx, "ic_offsets[0] must be 0");
  int64_t n = ic_offsets[nv], stride = 0;
  for (int v = 0; v < nv; v++) {
    if (ic_offsets[v + 1] < ic_offsets[v]) return set_err(ctx, "ic_offsets must be non-decreasing");
    stride = std::max(stride, ic_offsets[v + 1] - ic_offsets[v]);
  }
  if (stride > INT32_MAX / 2) return set_err(ctx, "too many initial conditions per variant");
  if (out) std::memset(out, 0, sizeof(*out));
  if (n == 0) return 0;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  SurvivalRun R;
  if (!R.alloc(nv, (int)stride)) return set_err(ctx, "device allocation for the initial conditions failed");
  // padded host staging of x..pz, [v * stride + i]
  int64_t total = (int64_t)nv * stride;
  std::vector<double> pad(total);
  double* dst[6] = {R.st[0].x, R.st[0].px, R.st[0].y, R.st[0].py, R.st[0].z, R.st[0].pz};
  for (int c = 0; c < 6; c++) {
    std::fill(pad.begin(), pad.end(), 0.0);
    for (int v = 0; v < nv; v++)
      std::memcpy(&pad[(size_t)v * stride], ics + (size_t)c * n + ic_offsets[v],
                  (size_t)(ic_offsets[v + 1] - ic_offsets[v]) * sizeof(double));
    if (hipMemcpy(dst[c], pad.data(), total * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return set_err(ctx, "upload of the initial conditions failed");
  }
  std::vector<int32_t> h_count(nv);
  for (int v = 0; v < nv; v++) h_count[v] = (int32_t)(ic_offsets[v + 1] - ic_offsets[v]);
  std::vector<int32_t> h_surv(total);
  const int block = 256;
  int rc = run_survival(ctx, lat, R, h_count, n_turns, turns_per_block > 0 ? turns_per_block : ctx->sched.turns_per_block,
                        ctx->sched.graph_depth, h_surv.data(), out,
                        [&]() {
                          hipLaunchKernelGGL(bmx_ics_init_kernel,
                                             dim3((unsigned)((stride + block - 1) / block), (unsigned)nv), dim3(block),
                                             0, ctx->stream, lat->lc, R.st[0], R.d_count, (int)stride, R.d_surv, n_turns);
                        });
  if (rc != 0) return rc;
  int64_t live = 0;
  for (int v = 0; v < nv; v++)
    for (int64_t i = ic_offsets[v]; i < ic_offsets[v + 1]; i++) {
      survival_turn[i] = h_surv[(size_t)v * stride + (i - ic_offsets[v])];
      live += survival_turn[i] >= n_turns;
    }
  if (out) {
    out->n_live = live;
    out->n_lost = n - live;
  }
  return 0;
}



int32_t bmx_ctx_set_schedule(bmx_ctx* ctx, const bmx_schedule* s) {
  if (!ctx || !s || s->turns_per_block <= 0 || s->graph_depth < 0) return set_err(ctx, "invalid schedule");
  ctx->sched = *s;
  return 0;
}

int32_t bmx_ctx_get_schedule(const bmx_ctx* ctx, bmx_schedule* s) {
  if (!ctx || !s) return -1;
  *s = ctx->sched;
  return 0;
}

int32_t bmx_tune_grid(bmx_ctx* ctx, const bmx_lattice* lat, const bmx_grid* grid, const double* ref_orbit,
                      int32_t n_turns, bmx_schedule* best) {
  if (!ctx || !lat || !grid || !best || n_turns <= 0 || grid->n_angle <= 0 || grid->n_amp <= 0 || grid->n_pz <= 0)
    return set_err(ctx, "invalid arguments");
  // sparse sample of the grid: every 4th angle and amplitude, every pz slice, every variant
  int nv = lat->n_variants;
  std::vector<std::array<double, 3>> per;  // (x offset, y offset, pz slice)
  double dth = (grid->n_angle > 1) ? (grid->angle_max - grid->angle_min) / (grid->n_angle - 1) : 0.0;
  for (int ip = 0; ip < grid->n_pz; ip++)
    for (int ia = 0; ia < grid->n_angle; ia += 4)
      for (int k = 0; k < grid->n_amp; k += 4) {
        double amp = grid->amp_max_x * (double)(k + 1) / grid->n_amp, th = grid->angle_min + dth * ia;
        per.push_back({amp * cos(th), amp * grid->amp_ratio_y * sin(th), (double)ip});
      }
  int64_t m = (int64_t)per.size(), n = m * nv;
  std::vector<double> ics(6 * n);
  std::vector<int64_t> off(nv + 1);
  for (int v = 0; v <= nv; v++) off[v] = (int64_t)v * m;
  for (int v = 0; v < nv; v++)
    for (int64_t j = 0; j < m; j++) {
      int ip = (int)per[j][2];
      double pz0 = (grid->n_pz > 1) ? grid->pz_min + (grid->pz_max - grid->pz_min) * ip / (grid->n_pz - 1) : grid->pz_min;
      double r[6] = {0, 0, 0, 0, 0, pz0};
      if (ref_orbit)
        for (int c = 0; c < 6; c++) r[c] = ref_orbit[((int64_t)v * grid->n_pz + ip) * 6 + c];
      r[0] += per[j][0];
      r[2] += per[j][1];
      for (int c = 0; c < 6; c++) ics[(size_t)c * n + off[v] + j] = r[c];
    }
  std::vector<int32_t> surv(n);
  const int tpbs[6] = {4, 8, 16, 32, 64, 128};
  const int depths[2] = {1, 8};
  int probe = std::min(n_turns, 64);
  bmx_schedule saved = ctx->sched, pick = saved;
  double best_rate = -1;
  for (int d : depths)
    for (int t : tpbs) {
      if (t > probe && t != tpbs[0]) continue;
      ctx->sched.turns_per_block = t;
      ctx->sched.graph_depth = d;
      bmx_stats st;
      auto t0 = std::chrono::steady_clock::now();
      int rc = bmx_track_ics(ctx, lat, ics.data(), off.data(), probe, t, surv.data(), &st);
      double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (rc != 0) {
        ctx->sched = saved;
        return rc;
      }
      double rate = st.pet / std::max(wall, 1e-9);
      if (rate > best_rate) {
        best_rate = rate;
        pick.turns_per_block = t;
        pick.graph_depth = d;
        pick.probe_pet_per_s = rate;
      }
    }
  ctx->sched = pick;
  *best = pick;
  return 0;
}

int32_t bmx_bunch_moments(bmx_ctx* ctx, const bmx_particles* p, double* moments) {
  if (!ctx || !p || !moments || !particles_ok(p)) return set_err(ctx, "invalid arguments");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  for (int k = 0; k < N_MOM; k++) moments[k] = 0.0;
  if (p->n <= 0) return 0;
  Staged st;
  bmx_particles q = *p;
  q.ix_lost = nullptr;
  q.turn_lost = nullptr;
  if (stage_particles(ctx, &q, st) != 0) return -1;
  const int blocks = 1024;
  double *d_part = nullptr, *d_out = nullptr;
  if (hipMalloc(&d_part, (size_t)blocks * N_MOM * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_out, N_MOM * sizeof(double)) != hipSuccess) {
    if (d_part) (void)hipFree(d_part);
    st.hp[0] = nullptr;  // nothing to copy back
    q.on_device = 1;
    finish_particles(ctx, &q, st, hipSuccess, nullptr);
    return set_err(ctx, "hipMalloc for the moment reduction failed");
  }
  size_t lds = (size_t)N_MOM * 256 * sizeof(double);
  hipLaunchKernelGGL(bmx_moments_partial_kernel, dim3(blocks), dim3(256), lds, ctx->stream, st.s, p->n, d_part);
  hipLaunchKernelGGL(bmx_moments_final_kernel, dim3(1), dim3(256), lds, ctx->stream, d_part, blocks, d_out);
  hipError_t le = hipGetLastError();
  (void)hipMemcpyAsync(moments, d_out, N_MOM * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  // no coordinates change: skip the copy-back of host arrays
  q.on_device = 1;
  int rc = finish_particles(ctx, &q, st, le, nullptr);
  (void)hipFree(d_part);
  (void)hipFree(d_out);
  return rc;
}

// closed_orbit_calc (bmad/code/closed_orbit_calc.f90:81-...) in the form the
// batched tracker serves: Newton iteration x <- x + (I - M)^-1 (T(x) - x) on
// the one-turn map T of every variant at once, with M from one-sided finite
// differences (one launch tracks 1 + n_dim probes per variant).  n_dim = 4
// keeps pz (and z) of the guess fixed; n_dim = 6 closes all six (rf on).
int32_t bmx_closed_orbit(bmx_ctx* ctx, const bmx_lattice* lat, int32_t n_dim, double* orbit, int32_t* status,
                         bmx_stats* out) {
  if (!ctx || !lat || !orbit || !status || (n_dim != 4 && n_dim != 6)) return set_err(ctx, "invalid arguments");
  int nv = lat->n_variants;
  int np = 1 + n_dim;
  int64_t n = (int64_t)nv * np;
  std::vector<int64_t> off(nv + 1);
  for (int v = 0; v <= nv; v++) off[v] = (int64_t)v * np;
  std::vector<double> c(6 * n), beta(n);
  std::vector<int32_t> state(n), ixl(n), tl(n);
  std::vector<int> done(nv, 0);
  for (int v = 0; v < nv; v++) status[v] = 1;
  const double h = 1e-8;
  double mc2 = lat->lc.mc2, p0c = lat->lc.p0c;
  if (out) std::memset(out, 0, sizeof(*out));
  for (int it = 0; it < 40; it++) {
    for (int v = 0; v < nv; v++)
      for (int j = 0; j < np; j++) {
        int64_t k = off[v] + j;
        for (int d = 0; d < 6; d++) c[(size_t)d * n + k] = orbit[v * 6 + d] + ((j > 0 && d == j - 1) ? h : 0.0);
        double pz = c[(size_t)5 * n + k], pc = p0c * (1.0 + pz);
        beta[k] = pc / std::sqrt(pc * pc + mc2 * mc2);
        state[k] = BMX_ALIVE;
      }
    bmx_particles p;
    std::memset(&p, 0, sizeof(p));
    p.x = &c[0]; p.px = &c[(size_t)n]; p.y = &c[(size_t)2 * n]; p.py = &c[(size_t)3 * n];
    p.z = &c[(size_t)4 * n]; p.pz = &c[(size_t)5 * n];
    p.beta = beta.data(); p.state = state.data(); p.ix_lost = ixl.data(); p.turn_lost = tl.data();
    p.n = n;
    p.on_device = 0;
    bmx_stats st;
    if (bmx_track_multi(ctx, lat, off.data(), 1, &p, &st) != 0) return -1;
    if (out) out->kernel_ms += st.kernel_ms;
    int n_open = 0;
    for (int v = 0; v < nv; v++) {
      if (done[v]) continue;
      bool lost = false;
      for (int j = 0; j < np; j++) lost |= state[off[v] + j] != BMX_ALIVE;
      if (lost) { status[v] = 2; done[v] = 1; continue; }
      // residual r = T(x) - x and Jacobian columns over the first n_dim coordinates
      double r[6], A[6][7];
      double rmax = 0;
      for (int d = 0; d < n_dim; d++) {
        r[d] = c[(size_t)d * n + off[v]] - orbit[v * 6 + d];
        rmax = std::max(rmax, std::fabs(r[d]));
      }
      if (rmax < 1e-15) { status[v] = 0; done[v] = 1; continue; }
      for (int d = 0; d < n_dim; d++) {
        for (int j = 0; j < n_dim; j++) {
          double mdj = (c[(size_t)d * n + off[v] + 1 + j] - c[(size_t)d * n + off[v]]) / h;
          A[d][j] = (d == j ? 1.0 : 0.0) - mdj;  // (I - M)
        }
        A[d][n_dim] = r[d];
      }
      // Gaussian elimination with partial pivoting: (I - M) dx = r
      bool sing = false;
      for (int col = 0; col < n_dim && !sing; col++) {
        int piv = col;
        for (int row = col + 1; row < n_dim; row++)
          if (std::fabs(A[row][col]) > std::fabs(A[piv][col])) piv = row;
        if (std::fabs(A[piv][col]) < 1e-14) { sing = true; break; }
        for (int k = 0; k <= n_dim; k++) std::swap(A[col][k], A[piv][k]);
        for (int row = col + 1; row < n_dim; row++) {
          double f = A[row][col] / A[col][col];
          for (int k = col; k <= n_dim; k++) A[row][k] -= f * A[col][k];
        }
      }
      if (sing) { status[v] = 3; done[v] = 1; continue; }
      double dx[6];
      for (int row = n_dim - 1; row >= 0; row--) {
        double acc = A[row][n_dim];
        for (int k = row + 1; k < n_dim; k++) acc -= A[row][k] * dx[k];
        dx[row] = acc / A[row][row];
      }
      double dmax = 0;
      for (int d = 0; d < n_dim; d++) {
        orbit[v * 6 + d] += dx[d];
        dmax = std::max(dmax, std::fabs(dx[d]));
      }
      if (dmax < 1e-15) { status[v] = 0; done[v] = 1; continue; }
      n_open++;
    }
    if (n_open == 0) break;
  }
  return 0;
}

}  // extern "C"
