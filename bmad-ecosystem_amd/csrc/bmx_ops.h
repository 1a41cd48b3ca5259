// bmx_ops.h -- device side of the tracker, part 1: the per-lane particle
// state, the series and special functions, one function per element op and
// the op dispatch (run_op).  Included by bmx_track.hip and, as text, by the
// runtime-compiled cell kernels (bmx_jit.cpp); it needs bmx.h and
// bmx_internal.h, with `using namespace bmx` in effect.
#pragma once


#define D_C_LIGHT 2.99792458e8
#define D_TWOPI 6.283185307179586476925286766559
#define D_PI 3.141592653589793238462643383279

// ---------------------------------------------------------------------------
// per-lane particle state
// ---------------------------------------------------------------------------
// BMX_ABS_TIME 0 (runtime-compiled cell kernels of a lattice without
// absolute-time rf) leaves out the particle's time offset.
#ifndef BMX_ABS_TIME
#define BMX_ABS_TIME 1
#endif
// BMX_RAD 0 (runtime-compiled cell kernels of a lattice without radiation)
// leaves out the radiation op and the particle index it draws with.
#ifndef BMX_RAD
#define BMX_RAD 1
#endif
struct Lane {
  double x, px, y, py, z, pz, beta;
  double irp;  // 1 / (1 + pz)
  double bb;   // beta/beta0 - 1 as track_a_drift forms it (track_a_drift.f90:68-69)
#if BMX_ABS_TIME
  // t - t_ref + z / (beta c): the particle time t less its value on the
  // reference relation t = t_ref(s) - z / (beta c), which every map of the
  // path conserves (apply_energy_kick scales z with beta); zero for a
  // particle set up by init_coord (init_coord.f90:222)
  double dt;
#endif
#if BMX_RAD
  uint32_t pid;  // the particle's index in its store: the radiation stream's counter
#endif
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
typedef __attribute__((address_space(4))) const bmx_f64 cdouble;
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
__constant__ bmx_f64 kSeries[CS_END] = {
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

__device__ __forceinline__ bmx_f64 sg(bmx_f64 c) {
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

// The rare paths (a lane off an op's common path, a loss, an exact-form
// fallback) are marked unlikely, so that their code is laid out after the
// hot code of the element loop and the loop stays compact in the
// instruction cache.

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
// For u < 1e-4 (angles below 10 mrad) S is summed as its binomial series,
// which needs no division or square root; larger angles take the exact form.
__device__ __forceinline__ void drift_apply(Lane& L, double len, double pxr, double pyr, double S) {
  double lps = len * (1.0 + S);
  L.x += lps * pxr;
  L.y += lps * pyr;
  L.z += len * (L.bb - S);
}

// Every lane sums the series to u^3, exact to 3e-25 for u < 1e-6 (angles
// below 1 mrad, the usual case).  Only a wave holding a lane at u >= 1e-6
// takes, for those lanes, the series to u^5 (truncation < 3e-25 for
// u < 1e-4), the exact form, or the lost_pz test (u >= 1, which leaves the
// particle where it is).
__device__ __forceinline__ void drift_c(Lane& L, double len, double d0, double d1, double d2) {
  double pxr = L.px * L.irp, pyr = L.py * L.irp;
  double u = pxr * pxr + pyr * pyr;
  double S = u * (0.5 + u * (0.375 + u * d0));
  bool wide = u >= 1e-6;
  if (__builtin_expect(wave_any(wide), 0)) {
    bool move = true;
    if (wide) {
      if (u < 1e-4) {
        S = u * (0.5 + u * (0.375 + u * (d0 + u * (d1 + u * d2))));
      } else if (u >= 1.0) {
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

// track_a_drift with direction = -1 (bmad/low_level/track_a_drift.f90:25-116,
// backward branch): x and y as forward, dz = L ((-beta / beta_ref - 1) - (1 / ps - 1))
__device__ __forceinline__ void drift_back(Lane& L, double len, double beta_ref) {
  double pxr = L.px * L.irp, pyr = L.py * L.irp;
  double pxy2 = pxr * pxr + pyr * pyr;
  if (pxy2 >= 1.0) { L.state = BMX_LOST_PZ; return; }
  double ps = sqrt(1.0 - pxy2);
  L.x += len * pxr / ps;
  L.y += len * pyr / ps;
  L.z += len * ((-L.beta / beta_ref - 1.0) - (1.0 / ps - 1.0));
}
__device__ __forceinline__ void op_drift_back(Lane& L, const DevOp& op) { drift_back(L, op.c[0], op.c[1]); }

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
__device__ __forceinline__ void kick_horner(const bmx_f64* c, int nh, int lo, double x, double y, double& kr,
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

// nh = flags & 0xff, lo = n of the op (wave-uniform; constants in a cell kernel)
__device__ __forceinline__ void kick_eval(const DevOp& op, int nh, int lo, double x, double y, double& kr, double& ki) {
  kick_horner(op.c, nh, lo, x, y, kr, ki);
}

__device__ __forceinline__ void op_kick(Lane& L, const DevOp& op, int nh, int lo) {
  double kr, ki;
  kick_eval(op, nh, lo, L.x, L.y, kr, ki);
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

__device__ __forceinline__ void op_kd_slices(Lane& L, const DevOp& op, int n_step, int fl) {
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
__device__ __forceinline__ void op_bend_kick(Lane& L, const DevOp& op, int nh, int lo) {
  double kr, ki;
  kick_eval(op, nh, lo, L.x, L.y, kr, ki);
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

// c: 0 step length L, 1 k1 (1+pz) = b1 / L, 2 k1 (1+pz) L^2
__device__ __forceinline__ void op_quad(Lane& L, const DevOp& op, const LatticeConsts& lc) {
  double len = op.c[0];
  double rel_p = 1.0 + L.pz;
  double irp = L.irp;
  double k1 = op.c[1] * irp;  // charge_dir b1 / (L rel_p)
  double kx = -k1, ky = k1;   // quad_mat2_calc(-k1, ...) for x, quad_mat2_calc(k1, ...) for y
  double cx, sx, cy, sy;
  double w = op.c[2] * irp;   // k1 L^2
  if (__builtin_expect(fabs(w) <= 1.0, 1)) {
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
  // z: per plane the reference's three terms c1 x^2 + c2 x px + c3 px^2 with
  // c1 = k (L - c s)/4, c2 = -k s^2 / (2 (1+pz)), c3 = -(L + c s) / (4 (1+pz)^2)
  // (track_a_quadrupole.f90:82-112), written with p = px / (1+pz) as
  // [x (k (L - c s) x - 2 k s^2 p) - (L + c s) p^2] / 4
  double x = L.x, px = L.px, y = L.y, py = L.py;
  double pxr = px * irp, pyr = py * irp;
  double ax = cx * sx, ay = cy * sy;
  double kxs = kx * sx, kys = ky * sy;
  double zx = x * (kx * (len - ax) * x - 2.0 * kxs * sx * pxr) - (len + ax) * pxr * pxr;
  double zy = y * (ky * (len - ay) * y - 2.0 * kys * sy * pyr) - (len + ay) * pyr * pyr;
  L.z += 0.25 * (zx + zy);
  L.x = cx * x + sx * pxr;
  L.px = (kxs * rel_p) * x + cx * px;
  L.y = cy * y + sy * pyr;
  L.py = (kys * rel_p) * y + cy * py;
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
// correctly rounded result.  The transfer-matrix build (BMX_TANGENT) takes
// the plain operations, whose derivatives its number type carries.
#if BMX_TANGENT
__device__ __forceinline__ double sqrt_mid(double x) { return sqrt(x); }
__device__ __forceinline__ double rcp_mid(double b) { return 1.0 / b; }
__device__ __forceinline__ double div_mid(double a, double b) { return a / b; }
#else
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
#endif

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
__device__ __forceinline__ Lane bend_general(Lane L, const bmx_f64* c, double p0c, double mc2sq, double mal,
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
  if (__builtin_expect(wave_any(!common), 0)) {
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
  L.z = r4 + op.c[4] * z0 + z1 * r0 + z2 * r1 + z11 * r0 * r0 + z12 * r0 * r1 + z22 * r1 * r1 + z33 * r2 * r2 + z34 * r2 * r3 +
        z44 * r3 * r3;
  L.z += low_energy_dz(L, lc, length);
}

// hwang_bend_edge_kick (bmad/modules/fringe_mod.f90:240-418), time_dir = +1.
// c: 0 gt, 1 gt2, 2 gs2, 3 k1_tane, 4 fg_factor, 5 g_tot, 6 tan_e
__device__ __forceinline__ void op_hwang(Lane& L, const DevOp& op, int flags) {
  double gt = op.c[0], gt2 = op.c[1], gs2 = op.c[2], k1t = op.c[3], fg = op.c[4], g_tot = op.c[5], te = op.c[6];
  double ef = L.irp;
  double v0 = L.x, v1 = L.px, v2 = L.y, v3 = L.py;
  double dx, dpx, dy, dpy, dz;
  if (flags & F_HWANG_ENTER) {
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

// hard_multipole_edge_kick (bmad/modules/fringe_mod.f90:691-864) of a
// quadrupole (n = 1, b = k1) or sextupole (n = 2, b = k2) in the element frame
// (a = 0, charge_dir = 1).  The reference's complex expressions with
// cab = c0 / (1+pz) real, poly_n1 = z^n = (ar, ai), poly = z^(n+1) = (br, bi),
// z = x + i y, written out: with xny = x - i cn y (for fx) and y + i cn x
// (for fy), Re(cab w) = cab Re(w), and the derivatives of the reference's
// dpoly and dxny.  The second derivatives only feed its matrix.
// c: 0 c0 = -+ b / (4 (n+2)) (negative at the first track edge), 1 cn = (n+3)/(n+1)
__device__ __forceinline__ void op_hard_mult_edge(Lane& L, const DevOp& op, int n) {
  double x = L.x, y = L.y, px = L.px, py = L.py;
  double c = op.c[0] * L.irp, cn = op.c[1];
  double ar = x, ai = y;
  if (n == 2) {
    ar = x * x - y * y;
    ai = x * y + y * x;
  }
  double br = ar * x - ai * y, bi = ar * y + ai * x;
  double m = (double)(n + 1);
  double fx = c * (br * x + cn * (bi * y));
  double dfx_dx = c * (m * (ar * x + cn * (ai * y)) + br);
  double dfx_dy = c * (cn * bi - m * (ai * x - cn * (ar * y)));
  double fy = c * (br * y - cn * (bi * x));
  double dfy_dx = c * (m * (ar * y - cn * (ai * x)) - cn * bi);
  double dfy_dy = c * (br - m * (cn * (ar * x) + ai * y));
  double denom = (1.0 - dfx_dx) * (1.0 - dfy_dy) - dfx_dy * dfy_dx;
  L.x = x - fx;
  L.px = px + ((1.0 - dfy_dy - denom) * px + dfy_dx * py) / denom;
  L.y = y - fy;
  L.py = py + (dfx_dy * px + (1.0 - dfx_dx - denom) * py) / denom;
  L.z += (L.px * fx + L.py * fy) * L.irp;
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
__device__ __forceinline__ void op_offset(Lane& L, const DevOp& op, bool set, int flags) {
  double rel_p = 1.0 + L.pz;
  if (flags & F_OFFSET_AFFINE) {
    // a misaligned bend (offset_particle.f90:208-235, :409-440): its chain of
    // arc shifts, offsets and rotations composed on the host, acting on
    // (x, y, 0) and the momentum (px, py, pz_s) with pz_s = sqrt((1+pz)^2 - px^2 - py^2)
    const bmx_f64* w = &op.c[4];
    double pzz = rel_p * rel_p - L.px * L.px - L.py * L.py;
    if (pzz <= 0.0) { L.state = BMX_LOST_PZ; return; }
    double x = L.x, y = L.y;
    double r0 = w[0] * x + w[1] * y + op.c[0];
    double r1 = w[3] * x + w[4] * y + op.c[1];
    double r2 = w[6] * x + w[7] * y + op.c[2];
    double p2 = sqrt(pzz);
    double q0 = w[0] * L.px + w[1] * L.py + w[2] * p2;
    double q1 = w[3] * L.px + w[4] * L.py + w[5] * p2;
    L.px = q0;
    L.py = q1;
    L.x = r0;
    L.y = r1;
    double ds = (set ? 0.0 : 2.0 * op.c[3]) - r2;
    if (fabs(ds) > 1e-14) op_drift_noref(L, ds);
    return;
  }
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
  if (flags & F_OFFSET_PITCH) {
    const bmx_f64* w = &op.c[4];
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

// ab_multipole_kicks of electric multipoles (bmad/modules/multipole_mod.f90:
// 82-174, with ab_multipole_kick's electric branch :338-345): the ab kick of
// coefficients proportional to 1/beta (the rest folded into the op), then
// the momentum change it implies: (1+pz)^2 - px^2 - py^2 is kept, beta follows
// the new pz, and z scales with beta.  Coefficients as op_kick.
__device__ __forceinline__ void op_ekick(Lane& L, const DevOp& op, int nh, int lo, const LatticeConsts& lc) {
  double kr, ki;
  kick_horner(op.c, nh, lo, L.x, L.y, kr, ki);
  double ib = 1.0 / L.beta;
  double kx = -kr * ib, ky = ki * ib;
  double px0 = L.px, py0 = L.py, pz0 = L.pz;
  L.px = px0 + kx;
  L.py = py0 + ky;
  double rel0 = 1.0 + pz0;
  double alpha = (kx * (2.0 * px0 + kx) + ky * (2.0 * py0 + ky)) / (rel0 * rel0);
  if (alpha < -1.0) {
    L.state = BMX_LOST_PZ;
    return;
  }
  L.pz = pz0 + rel0 * d_sqrt_one(alpha);
  double rel = 1.0 + L.pz, mc = lc.mc2 / lc.p0c;
  double beta_new = rel / sqrt(rel * rel + mc * mc);
  L.z = L.z * beta_new / L.beta;
  L.beta = beta_new;
  refresh_invariants(L, lc);
}

// electric_longitudinal_fringe (bmad/low_level/apply_element_edge_kick.f90:
// 283-329), bmad_standard field: apply_energy_kick(-ff phi) with phi the
// scalar potential of the electric multipoles.  The op holds the
// coefficients C of -ff phi = Re sum C_n (x + i y)^n, in op_kick's layout.
__device__ __forceinline__ void op_efringe(Lane& L, const DevOp& op, int nh, int lo, const LatticeConsts& lc) {
  double kr, ki;
  kick_horner(op.c, nh, lo, L.x, L.y, kr, ki);
  energy_kick(L, kr, lc);
}

// bend_apply_multipole_kicks, electric part (bmad/low_level/track_a_bend.f90:
// 338-471, bmad_standard fields): the kick of the field E = (Re K, -Im K),
// K = sum c_n (x + i y)^n, scaled by (1 + g x) / (ps beta) with
// ps = sqrt((1+pz)^2 - px^2 - py^2) / (1+pz) (step length, coefficient and
// charge / p0c folded into c), then pz, beta and z as the reference updates
// them.  c: kick coefficients as op_kick, 13 g.
__device__ __forceinline__ void op_bend_ekick(Lane& L, const DevOp& op, int nh, int lo, const LatticeConsts& lc) {
  double kr, ki;
  kick_horner(op.c, nh, lo, L.x, L.y, kr, ki);
  double px0 = L.px, py0 = L.py, rel0 = 1.0 + L.pz;
  double ps2 = rel0 * rel0 - px0 * px0 - py0 * py0;
  double ps = sqrt(ps2) / rel0;
  double f = (1.0 + op.c[13] * L.x) / (ps * L.beta);
  double kx = f * kr, ky = -f * ki;
  double alpha = (kx * (2.0 * px0 + kx) + ky * (2.0 * py0 + ky)) / (rel0 * rel0);
  L.px = px0 + kx;
  L.py = py0 + ky;
  L.pz = L.pz + rel0 * d_sqrt_one(alpha);
  double mc = lc.mc2 / lc.p0c;
  double beta_new = (1.0 + L.pz) / sqrt(ps2 + L.px * L.px + L.py * L.py + mc * mc);
  L.z = L.z * beta_new / L.beta;
  L.beta = beta_new;
  refresh_invariants(L, lc);
}

// track_a_rfcavity body (bmad/low_level/track_a_rfcavity.f90:60-140).
// particle_rf_time (bmad/code/particle_rf_time.f90:32-120): relative time
// -z/(beta c) (:89); absolute time (F_RF_ABS) t - ref_time_start of the
// element (:72-74), formed as c[8] + (k t_lat + r) + dt - z/(beta c) with
// c[8] the lattice-frame reference time of the first track edge less
// ref_time_start (0 forward, delta_ref_time backward) and the whole lattice
// periods k t_lat taken as k frac(f t_lat) cycles (c[9]).  F_RF_BACK: the
// drifts between the kicks are backward drifts (c[10] beta_ref).
// c: 0 L, 1 voltage x charge_dir, 2 phi0 [+ phi0_multipass (relative) or
// - frac(f dt_ref) (absolute, orientation x direction = -1)], 3 rf_frequency,
// 4 1/(2 f), 5, 6 coupler dp_x, dp_y, 7 2 pi coupler_phase; n = n_slice
__device__ void op_rfcavity(Lane& L, const DevOp& op, int n_slice, int flags, const LatticeConsts& lc,
                            const TurnTime& tt) {
  double len = op.c[0], volt = op.c[1], freq = op.c[3], half_t = op.c[4];
  double time = -L.z / (L.beta * D_C_LIGHT);
  double kcyc = 0.0;
#if BMX_ABS_TIME
  if (flags & F_RF_ABS) {
    time = (op.c[8] + tt.r + L.dt) + time;
    kcyc = (double)tt.k * op.c[9];
    kcyc -= rint(kcyc);
  }
#endif
  if (freq != 0.0 && !(time >= -half_t && time < half_t)) time = d_modulo2(time, half_t);
  double phase = d_modulo2(D_TWOPI * (op.c[2] - time * freq - kcyc), D_PI);
  double dl = len / n_slice;
  // rf_coupler_kick (bmad/low_level/rf_coupler_kick.f90:20-106) at one edge
  auto coupler = [&]() {
    double ph = 0.5 * D_PI - phase + op.c[7];
    double dpx = op.c[5], dpy = op.c[6];
    L.px += dpx * cos(ph) / lc.p0c;
    L.py += dpy * cos(ph) / lc.p0c;
    double f2 = sin(ph) * D_TWOPI * freq / D_C_LIGHT;
    energy_kick(L, (dpx * L.x + dpy * L.y) * f2, lc);
  };
  if (flags & F_COUPLER_ENTER) {
    coupler();
    if (L.state != BMX_ALIVE) return;
  }
  // sin(phase) of every slice: one sincos at the entrance, then a rotation by
  // each slice's phase advance as the running phase took it, (phase + d) -
  // phase, so that (sp, cp) follow the rounded phase the reference passes to
  // sin.  The advance is 2 pi f dz / (c beta), a few 1e-9 rad for relativistic
  // particles; its sin and cos are series while it stays below 2^-8 (the
  // next terms below 1e-21), and a wave holding a larger one takes sincos of
  // the phase itself.
  double sp, cp;
  sincos(phase, &sp, &cp);
  for (int i = 0; i <= n_slice; i++) {
    double dE = volt * sp / n_slice;
    if (i == 0 || i == n_slice) dE = dE / 2.0;
    energy_kick(L, dE, lc);
    if (L.state != BMX_ALIVE) return;
    if (L.pz == -1.0) { L.state = BMX_LOST_PZ; return; }
    if (i != n_slice && dl != 0.0) {
      double z0 = L.z;
      if (flags & F_RF_BACK) {
        drift_back(L, dl, op.c[10]);
      } else {
        op_drift(L, dl);
      }
      if (L.state != BMX_ALIVE) return;
      double p1 = phase + D_TWOPI * freq * (L.z - z0) / (D_C_LIGHT * L.beta);
      double d = p1 - phase;
      phase = p1;
      if (__builtin_expect(wave_any(!(fabs(d) < 0x1p-8)), 0)) {
        sincos(phase, &sp, &cp);
      } else {
        double d2 = d * d;
        double sd = d + d * d2 * (-1.0 / 6.0 + d2 * (1.0 / 120.0 - d2 * (1.0 / 5040.0)));
        double cd = 1.0 + d2 * (-0.5 + d2 * (1.0 / 24.0 - d2 * (1.0 / 720.0)));
        double s1 = sp * cd + cp * sd;
        cp = cp * cd - sp * sd;
        sp = s1;
      }
    }
  }
  if (flags & F_COUPLER_EXIT) coupler();
}

// Screen of the momentum conditions of orbit_too_large: |px| + |py| <
// (1 + pz) / 2 implies 1 + pz > 0 and px^2 + py^2 < (1 + pz)^2 with room to
// spare for rounding (halving is exact), in two VALU operations and a compare
// where the exact test takes four and two compares.  A lane that fails it
// (or holds a NaN) takes the exact tests.
__device__ __forceinline__ bool p_screen(const Lane& L) { return fabs(L.px) + fabs(L.py) < 0.5 + 0.5 * L.pz; }

// orbit_too_large (bmad/low_level/orbit_too_large.f90:26-120).  A straight-
// line screen of all four conditions runs first; only lanes that fail it take
// the reference's ordered tests, which pick the loss state.
__device__ __forceinline__ bool too_large(Lane& L, const LatticeConsts& lc) {
  double mal = lc.max_aperture_limit;
  bool ok = (fabs(L.x) <= mal) & (fabs(L.y) <= mal) & p_screen(L);
  if (__builtin_expect(ok, 1)) return false;
  double rel_p = 1.0 + L.pz;
  if (L.x > mal) { L.state = BMX_LOST_POS_X; return true; }
  if (-L.x > mal) { L.state = BMX_LOST_NEG_X; return true; }
  if (L.y > mal) { L.state = BMX_LOST_POS_Y; return true; }
  if (-L.y > mal) { L.state = BMX_LOST_NEG_Y; return true; }
  if (rel_p < 0.0) { L.state = BMX_LOST_PZ; return true; }
  double f = L.px * L.px + L.py * L.py - rel_p * rel_p;
  if (!(f > 0.0)) return false;
  if (fabs(L.px) > fabs(L.py)) L.state = (L.px > 0.0) ? BMX_LOST_POS_X : BMX_LOST_NEG_X;
  else L.state = (L.py > 0.0) ? BMX_LOST_POS_Y : BMX_LOST_NEG_Y;
  return true;
}

// aperture part of check_aperture_limit (bmad/low_level/check_aperture_limit.f90:257-309)
// c, rectangular: 0 x1_lim, 1 x2_lim, 2 y1_lim, 3 y2_lim;
// elliptical: 0 cx, 1 cy, 2 1/wx, 3 1/wy, 4 wx, 5 wy
// The phase-space limits of a collimator (check_aperture_limit.f90, the
// ecollimator / rcollimator branch before the x, y limits; check_this_limit):
// ps = px, py, z, pz (width2, center) pairs; an ecollimator takes each of
// (px, py) and (z, pz) as an ellipse when both widths are set.  uf: the
// unstable factor the x, y limits are compared with.
__device__ void ps_limits(Lane& L, const bmx_f64* ps, bool ell, double& uf) {
  auto one = [&](double v, double w2, double c0) {
    if (w2 == 0.0) return;
    if (v > c0 - w2 && v < c0 + w2) return;
    L.state = BMX_LOST;
    double d1 = (c0 - w2) - v, d2 = v - (c0 + w2);
    uf = d1 > d2 ? d1 : d2;
  };
  auto pair = [&](double u, double v, const bmx_f64* q) {
    if (!ell || q[0] == 0.0 || q[2] == 0.0) {
      one(u, q[0], q[1]);
      one(v, q[2], q[3]);
      return;
    }
    double a = (u - q[1]) / q[0], b = (v - q[3]) / q[2];
    double r = a * a + b * b;
    if (r > 1.0) {
      L.state = BMX_LOST;
      uf = sqrt(r) - 1.0;
    }
  };
  pair(L.px, L.py, ps);
  pair(L.z, L.pz, ps + 4);
}

// true when ps_limits would leave the particle alone (the same tests)
__device__ __forceinline__ bool ps_inside(const Lane& L, const bmx_f64* ps, bool ell) {
  auto one = [&](double v, double w2, double c0) { return w2 == 0.0 || (v > c0 - w2 && v < c0 + w2); };
  auto pair = [&](double u, double v, const bmx_f64* q) {
    if (!ell || q[0] == 0.0 || q[2] == 0.0) return one(u, q[0], q[1]) && one(v, q[2], q[3]);
    double a = (u - q[1]) / q[0], b = (v - q[3]) / q[2];
    return !(a * a + b * b > 1.0);
  };
  return pair(L.px, L.py, ps) && pair(L.z, L.pz, ps + 4);
}

__device__ __forceinline__ void aperture(Lane& L, const bmx_f64* c, bool elliptical, double uf = 0.0) {
  if (elliptical) {
    double xp = L.x - c[0], yp = L.y - c[1];
    // screen with the reciprocal widths: it is within a few ulp of the
    // reference's r, so r_fast < 1 - 1e-13 decides r < 1 exactly
    double fx = xp * c[2], fy = yp * c[3];
    if (__builtin_expect(fx * fx + fy * fy < 1.0 - 1e-13, 1)) return;
    double ax = xp / c[4], ay = yp / c[5];
    double r = ax * ax + ay * ay;
    if (r < 1.0) return;
    if (fabs(ax) > fabs(ay)) L.state = (xp > 0.0) ? BMX_LOST_POS_X : BMX_LOST_NEG_X;
    else L.state = (yp > 0.0) ? BMX_LOST_POS_Y : BMX_LOST_NEG_Y;
    return;
  }
  double x1 = c[0], x2 = c[1], y1 = c[2], y2 = c[3], f;
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
// c: the aperture constants (as aperture()); ps: collimator phase-space
// limits (F_END_PS).
__device__ __forceinline__ void ele_end(Lane& L, const bmx_f64* c, int flags, const LatticeConsts& lc,
                                        const bmx_f64* ps) {
  double mal = lc.max_aperture_limit;
  bool inside = p_screen(L);
  // F_END_INBOX: the aperture lies inside the max_aperture_limit box, so its
  // screen implies |x|, |y| <= max_aperture_limit
  if (!(flags & F_END_INBOX)) inside &= (fabs(L.x) <= mal) & (fabs(L.y) <= mal);
  if (flags & F_END_APERTURE) {
    if (flags & F_END_ELLIPTICAL) {
      double fx, fy;
      if (flags & F_END_CENTERED) {  // cx = cy = 0: x - 0 = x exactly
        fx = L.x * c[2];
        fy = L.y * c[3];
      } else {
        fx = (L.x - c[0]) * c[2];
        fy = (L.y - c[1]) * c[3];
      }
      inside &= fx * fx + fy * fy < 1.0 - 1e-13;
    } else {
      inside &= (c[0] == 0.0 || L.x >= c[0]) & (c[1] == 0.0 || L.x <= c[1]) & (c[2] == 0.0 || L.y >= c[2]) &
                (c[3] == 0.0 || L.y <= c[3]);
    }
  }
  if (flags & F_END_PS) inside &= ps_inside(L, ps, flags & F_PS_ELLIPTICAL);
  bool check = (L.state == BMX_ALIVE) & !inside;
  if (__builtin_expect(wave_any(check), 0)) {
    if (check && !too_large(L, lc) && (flags & F_END_APERTURE)) {
      double uf = 0.0;
      if (flags & F_END_PS) ps_limits(L, ps, flags & F_PS_ELLIPTICAL, uf);
      aperture(L, c, flags & F_END_ELLIPTICAL, uf);
    }
  }
}

#if BMX_RAD
// Philox4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): ten rounds of two
// 32x32 -> 64-bit products on the counter, the key bumped by the Weyl
// constants between rounds.  Counter-based, so a particle's deviates depend
// only on (key, counter), not on which lane or launch tracks it.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Two standard normal deviates (Box-Muller) from one Philox block: u1 in
// (0, 1] and u2 in [0, 1) from 53 bits each.
__device__ __forceinline__ void gauss_pair(const uint32_t r[4], bmx_f64* n0, bmx_f64* n1) {
  bmx_f64 u1 = ((bmx_f64)(((uint64_t)r[0] << 21) | (r[1] >> 11)) + 1.0) * 0x1p-53;
  bmx_f64 u2 = (bmx_f64)(((uint64_t)r[2] << 21) | (r[3] >> 11)) * 0x1p-53;
  bmx_f64 rad = ::sqrt(-2.0 * ::log(u1));
  bmx_f64 sn, cs;
  ::sincos(D_TWOPI * u2, &sn, &cs);
  *n0 = rad * cs;
  *n1 = rad * sn;
}

// track1_radiation (bmad/modules/radiation_mod.f90:62-141) at one edge, with
// the map radiation_map_setup made for that half of the element
// (ele%rad_map%rm0 at the entrance, %rm1 at the exit), pre-scaled by
// synch_rad_scale: vec += D (vec - ref_orb) + xfer_damp_vec, then
// vec += S N(0, 1)^6 with the Cholesky factor S, then pz < -1 is lost_pz.
// At the exit edge track1 first checks orbit_too_large (track1.f90:281-284).
// beta keeps its value, as in the reference; 1/(1+pz) and the drift's
// beta/beta0 - 1, which this code forms from pz, follow pz.  The particle
// time does not change, so its offset from the reference relation takes
// the change of z.  In the transfer-matrix build (BMX_TANGENT) the kick
// moves the orbit's values and leaves the derivatives as they are: track1
// applies track1_radiation to the orbit only (track1.f90:198-201, :316-319),
// so make_mat6's ele%mat6 is the body's matrix along the radiating orbit.
__device__ void op_rad(Lane& L, const DevOp& op, int flags, const LatticeConsts& lc, const TurnTime& tt) {
  if ((flags & F_RAD_EXIT) && too_large(L, lc)) return;
  // the map's address, wave-uniform: its entries come in through the scalar cache
  uint64_t a = (uint64_t)(uintptr_t)lc.rad + (uint64_t)(int64_t)(int)op.c[0] * (RAD_STRIDE * sizeof(bmx_f64));
  uint32_t a_lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
  uint32_t a_hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32));
  const cdouble* m = (const cdouble*)(uintptr_t)(((uint64_t)a_hi << 32) | a_lo);
  double v[6] = {L.x, L.px, L.y, L.py, L.z, L.pz};
#if BMX_TANGENT
  if (flags & F_RAD_DAMP) {
    bmx_f64 d[6];
#pragma unroll
    for (int c = 0; c < 6; c++) d[c] = v[c].v - sg(m[RAD_REF + c]);
#pragma unroll
    for (int r = 0; r < 6; r++) {
      bmx_f64 a = sg(m[RAD_DAMP + r * 6]) * d[0];
#pragma unroll
      for (int c = 1; c < 6; c++) a = a + sg(m[RAD_DAMP + r * 6 + c]) * d[c];
      v[r].v = v[r].v + a;
    }
#pragma unroll
    for (int r = 0; r < 6; r++) v[r].v = v[r].v + sg(m[RAD_VEC + r]);
  }
#else
  if (flags & F_RAD_DAMP) {
    double d[6];
#pragma unroll
    for (int c = 0; c < 6; c++) d[c] = v[c] - sg(m[RAD_REF + c]);
#pragma unroll
    for (int r = 0; r < 6; r++) {
      double a = sg(m[RAD_DAMP + r * 6]) * d[0];
#pragma unroll
      for (int c = 1; c < 6; c++) a = a + sg(m[RAD_DAMP + r * 6 + c]) * d[c];
      v[r] = v[r] + a;
    }
#pragma unroll
    for (int r = 0; r < 6; r++) v[r] = v[r] + sg(m[RAD_VEC + r]);
  }
  if (flags & F_RAD_FLUCT) {
    bmx_f64 g[6];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      uint32_t ctr[4] = {L.pid, (uint32_t)(op.ie * 2 + ((flags & F_RAD_EXIT) ? 1 : 0)), (uint32_t)tt.turn,
                         lc.rad_call * 4u + (uint32_t)q};
      philox4x32_10(ctr, lc.rad_key0, lc.rad_key1);
      gauss_pair(ctr, &g[2 * q], &g[2 * q + 1]);
    }
    int k = RAD_STOC;
#pragma unroll
    for (int r = 0; r < 6; r++) {
      double a = 0.0;
#pragma unroll
      for (int c = 0; c <= r; c++) a = a + sg(m[k++]) * g[c];
      v[r] = v[r] + a;
    }
  }
#endif
#if BMX_ABS_TIME
  L.dt = L.dt + (v[4] - L.z) / (L.beta * D_C_LIGHT);
#endif
  L.x = v[0]; L.px = v[1]; L.y = v[2]; L.py = v[3]; L.z = v[4]; L.pz = v[5];
  if (L.pz < -1.0) L.state = BMX_LOST_PZ;
  refresh_invariants(L, lc);
}
#endif

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
__device__ __forceinline__ void run_op(Lane& L, const DevOp& op, int type, int flags, int n, const LatticeConsts& lc,
                                       const TurnTime& tt) {
  unsigned tb = (unsigned)__builtin_amdgcn_readfirstlane((int)(1u << type));
  if (tb & (TB(OP_DRIFT) | TB(OP_KICK) | TB(OP_KD_SLICES))) {
    if (tb & TB(OP_DRIFT)) op_drift(L, op.c[0]);
    if (tb & TB(OP_KICK)) op_kick(L, op, flags & 0xff, n);
    if (tb & TB(OP_KD_SLICES)) op_kd_slices(L, op, n, flags);
  }
  if (tb & (TB(OP_QUAD) | TB(OP_BEND) | TB(OP_SBEND_FUSED))) {
    if (tb & TB(OP_QUAD)) op_quad(L, op, lc);
    if (tb & (TB(OP_BEND) | TB(OP_SBEND_FUSED))) op_sbend(L, op, flags, (tb & TB(OP_SBEND_FUSED)) != 0, lc);
  }
  if (tb & (TB(OP_HWANG) | TB(OP_RFCAVITY))) {
    if (tb & TB(OP_HWANG)) op_hwang(L, op, flags);
    if (tb & TB(OP_RFCAVITY)) op_rfcavity(L, op, n, flags, lc, tt);
  }
  constexpr unsigned RARE = TB(OP_KICK_REF0) | TB(OP_BEND_K1) | TB(OP_BEND_KICK) | TB(OP_LINEAR_EDGE) | TB(OP_TILT) |
                            TB(OP_OFFSET_SET) | TB(OP_OFFSET_UNSET) | TB(OP_TOO_LARGE) | TB(OP_APERTURE_IN) |
                            TB(OP_HARD_MULT_EDGE) | TB(OP_EKICK) | TB(OP_EFRINGE) | TB(OP_BEND_EKICK) |
                            TB(OP_DRIFT_BACK) | TB(OP_ZBETA) | TB(OP_RAD);
  if (tb & RARE) {
    if (tb & TB(OP_KICK_REF0)) op_kick_ref0(L, op);
    if (tb & TB(OP_BEND_K1)) op_bend_k1(L, op, lc);
    if (tb & TB(OP_BEND_KICK)) op_bend_kick(L, op, flags & 0xff, n);
    if (tb & TB(OP_LINEAR_EDGE)) op_linear_edge(L, op);
    if (tb & TB(OP_TILT)) op_tilt(L, op);
    if (tb & TB(OP_OFFSET_SET)) op_offset(L, op, true, flags);
    if (tb & TB(OP_OFFSET_UNSET)) op_offset(L, op, false, flags);
    if (tb & TB(OP_TOO_LARGE)) too_large(L, lc);
    if (tb & TB(OP_APERTURE_IN)) {
      double uf = 0.0;
      if (flags & F_APERTURE_PS) ps_limits(L, op.c + 6, flags & F_PS_ELLIPTICAL, uf);
      aperture(L, op.c, flags & F_APERTURE_ELLIPTICAL, uf);
    }
    if (tb & TB(OP_HARD_MULT_EDGE)) op_hard_mult_edge(L, op, n);
    if (tb & TB(OP_EKICK)) op_ekick(L, op, flags & 0xff, n, lc);
    if (tb & TB(OP_EFRINGE)) op_efringe(L, op, flags & 0xff, n, lc);
    if (tb & TB(OP_BEND_EKICK)) op_bend_ekick(L, op, flags & 0xff, n, lc);
    if (tb & TB(OP_DRIFT_BACK)) op_drift_back(L, op);
    if (tb & TB(OP_ZBETA)) L.z -= op.c[0] * L.beta;
#if BMX_RAD
    if (tb & TB(OP_RAD)) op_rad(L, op, flags, lc, tt);
#endif
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
  if (flags & F_END) ele_end(L, op.c + ((flags & F_END_PS) ? 8 : END_C), flags, lc, op.c);
}
#undef TB
