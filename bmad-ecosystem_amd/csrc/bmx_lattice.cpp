// bmx_lattice.cpp -- compiles bmx_ele_rec element lists into device op
// programs (bmx_internal.h).  This is the hoisting pass that replaces Bmad's
// per-particle element bookkeeping: multipole_ele_to_ab
// (bmad/code/multipole_ele_to_ab.f90:31-235), the tilt/pitch set-up of
// offset_particle (bmad/code/offset_particle.f90:66-530), the constants of the
// Hwang edge map (bmad/modules/fringe_mod.f90:240-418) and the RF set-up of
// track_a_rfcavity (bmad/low_level/track_a_rfcavity.f90:17-80).
//
// Scope of the device path: forward and backward tracking (direction = +1 or
// -1, time_dir = +1), element orientation +1 and -1 for every key, the
// reference species or its antiparticle, relative-time and absolute-time RF
// (bmad_com%absolute_time_tracking), a constant reference energy apart from
// rf cavities.  Anything else is rejected with an error message naming the
// element; the CPU side never silently substitutes another path.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bmx.h"
#include "bmx_internal.h"

namespace bmx {

namespace {

const double kPi = 3.141592653589793238462643383279;
const double kCLight = 2.99792458e8;

double species_mass(int s) {
  switch (s) {
    case BMX_ELECTRON: case BMX_POSITRON: return 0.51099895069e6;
    case BMX_PROTON: case BMX_ANTIPROTON: return 0.93827208943e9;
    case BMX_MUON: case BMX_ANTIMUON: return 105.6583755e6;
  }
  return 0;
}

// charge_of (particle_species_mod.f90), in units of e
double species_charge(int s) {
  switch (s) {
    case BMX_ELECTRON: case BMX_ANTIPROTON: case BMX_MUON: return -1;
    case BMX_POSITRON: case BMX_PROTON: case BMX_ANTIMUON: return 1;
  }
  return 0;
}

int nint(double x) { return (int)(x < 0 ? -std::floor(-x + 0.5) : std::floor(x + 0.5)); }

bool has_hkick_attr(int key) {
  switch (key) {
    case BMX_KEY_QUADRUPOLE: case BMX_KEY_SBEND: case BMX_KEY_SEXTUPOLE: case BMX_KEY_OCTUPOLE:
    case BMX_KEY_THICK_MULTIPOLE: case BMX_KEY_RFCAVITY: case BMX_KEY_KICKER: case BMX_KEY_INSTRUMENT:
    case BMX_KEY_MONITOR: case BMX_KEY_ECOLLIMATOR: case BMX_KEY_RCOLLIMATOR:
      return true;
  }
  return false;
}

// Magnetic An/Bn of an element as multipole_ele_to_ab(ele, use_ele_tilt =
// false, magnetic, include_kicks) returns them; *b1 receives the pulled
// quadrupole term when b1 != nullptr.
struct AB {
  double a[BMX_N_POLE + 1], b[BMX_N_POLE + 1];  // + 1: the electric potential runs one order higher
  int ix_max;
};

int max_nonzero(const double* a, const double* b, int n_hi) {
  int ix = -1;
  for (int n = 0; n <= n_hi; n++)
    if (a[n] != 0 || b[n] != 0) ix = n;
  return ix;
}

bool magnetic_ab(const bmx_ele_rec& e, bool include_kicks, AB* out, double* b1, std::string* err) {
  AB& r = *out;
  std::memset(&r, 0, sizeof(r));
  r.ix_max = -1;
  if (b1) *b1 = 0;
  if (!e.is_on || e.key == BMX_KEY_PIPE) return true;
  double nfact[BMX_N_POLE];  // n! (a local table: variants compile on several threads)
  nfact[0] = 1;
  for (int n = 1; n < BMX_N_POLE; n++) nfact[n] = n * nfact[n - 1];

  if (e.key == BMX_KEY_MULTIPOLE) {
    if (!e.has_mag_poles) return true;
    for (int n = 0; n < BMX_N_POLE; n++) {  // multipole1_kt_to_ab
      double knl = e.a[n], tn = e.b[n];
      if (knl == 0) continue;
      double ang = -tn * (n + 1);
      r.b[n] = knl * std::cos(ang) / nfact[n];
      r.a[n] = knl * std::sin(ang) / nfact[n];
    }
  } else {
    if (e.has_mag_poles) {
      if (e.field_master) { *err = "field_master multipoles are not supported"; return false; }
      for (int n = 0; n < BMX_N_POLE; n++) { r.a[n] = e.a[n]; r.b[n] = e.b[n]; }
      if (max_nonzero(r.a, r.b, BMX_N_POLE - 1) >= 0 && e.scale_multipoles) {
        double cnst = 0;
        int ref_exp = 0;
        bool scale = true;
        switch (e.key) {
          case BMX_KEY_SBEND: cnst = e.l * (e.g + e.dg); ref_exp = 0; break;
          case BMX_KEY_KICKER:
            cnst = (e.hkick == 0) ? e.vkick : (e.vkick == 0 ? e.hkick : std::sqrt(e.hkick * e.hkick + e.vkick * e.vkick));
            break;
          case BMX_KEY_HKICKER: case BMX_KEY_VKICKER: cnst = e.kick; break;
          case BMX_KEY_QUADRUPOLE: cnst = e.k1 * e.l; ref_exp = 1; break;
          case BMX_KEY_SEXTUPOLE: cnst = e.k2 * e.l; ref_exp = 2; break;
          case BMX_KEY_OCTUPOLE: cnst = e.k3 * e.l; ref_exp = 3; break;
          case BMX_KEY_AB_MULTIPOLE: case BMX_KEY_THICK_MULTIPOLE: case BMX_KEY_RFCAVITY: scale = false; break;
          default: *err = "multipoles on a non-scalable element"; return false;
        }
        if (scale) {
          if (e.r0_mag == 0) {
            for (int n = 0; n < BMX_N_POLE; n++) { r.a[n] *= cnst; r.b[n] *= cnst; }
          } else {
            double f = cnst * std::pow(e.r0_mag, ref_exp + 1);
            for (int n = 0; n < BMX_N_POLE; n++) { f /= e.r0_mag; r.a[n] *= f; r.b[n] *= f; }
          }
        }
      }
    }
    // this_ele_non_multipoles (multipole_ele_to_ab.f90:326-405)
    double ak[4] = {0, 0, 0, 0}, bk[4] = {0, 0, 0, 0};
    switch (e.key) {
      case BMX_KEY_HKICKER: bk[0] = -e.kick; break;
      case BMX_KEY_VKICKER: ak[0] = e.kick; break;
      case BMX_KEY_KICKER: bk[0] = -e.hkick; ak[0] = e.vkick; break;
      default:
        if (has_hkick_attr(e.key) && (e.hkick != 0 || e.vkick != 0)) {
          double tilt = (e.key == BMX_KEY_SBEND) ? e.ref_tilt : e.tilt;
          if (tilt != 0) {
            double c = std::cos(tilt), s = std::sin(tilt);
            bk[0] = -e.hkick * c - e.vkick * s;
            ak[0] = -e.hkick * s + e.vkick * c;
          } else {
            bk[0] = -e.hkick;
            ak[0] = e.vkick;
          }
        }
    }
    switch (e.key) {
      case BMX_KEY_QUADRUPOLE: bk[1] = e.k1 * e.l; break;
      case BMX_KEY_SBEND: bk[0] += e.dg * e.l; bk[1] = e.k1 * e.l; bk[2] = e.k2 * e.l / 2.0; break;
      case BMX_KEY_SEXTUPOLE: bk[2] = e.k2 * e.l / 2.0; break;
      case BMX_KEY_OCTUPOLE: bk[3] = e.k3 * e.l / 6.0; break;
    }
    if (include_kicks)
      for (int n = 0; n < 4; n++) { r.a[n] += ak[n]; r.b[n] += bk[n]; }
  }
  r.ix_max = max_nonzero(r.a, r.b, BMX_N_POLE - 1);
  if (b1) {  // pull_this_b1
    *b1 = r.b[1];
    r.b[1] = 0;
    if (r.ix_max <= 1) r.ix_max = max_nonzero(r.a, r.b, 1);
  }
  return true;
}

// multipole_ele_to_ab for the electric poles (multipole_ele_to_ab.f90:31-235,
// pole_type electric, no kicks, no tilt): a_elec, b_elec, scaled by
// r0_elec^-n when scale_multipoles and r0_elec is neither 0 nor 1.  Thin
// multipole elements track kt kicks of their magnetic poles alone.
bool electric_ab(const bmx_ele_rec& e, AB* out, std::string* err) {
  AB& r = *out;
  std::memset(&r, 0, sizeof(r));
  r.ix_max = -1;
  if (!e.is_on || e.key == BMX_KEY_PIPE || !e.has_elec_poles) return true;
  if (e.key == BMX_KEY_MULTIPOLE || e.key == BMX_KEY_AB_MULTIPOLE) return true;
  for (int n = 0; n < BMX_N_POLE; n++) { r.a[n] = e.a_elec[n]; r.b[n] = e.b_elec[n]; }
  if (max_nonzero(r.a, r.b, BMX_N_POLE - 1) < 0) return true;
  if (e.field_master) { *err = "field_master multipoles are not supported"; return false; }
  if (e.scale_multipoles && e.r0_elec != 0 && e.r0_elec != 1) {
    double f = e.r0_elec;
    for (int n = 0; n < BMX_N_POLE; n++) { f /= e.r0_elec; r.a[n] *= f; r.b[n] *= f; }
  }
  r.ix_max = max_nonzero(r.a, r.b, BMX_N_POLE - 1);
  return true;
}

DevOp make_op(int type, int ie) {
  DevOp op;
  std::memset(&op, 0, sizeof(op));
  op.type = type;
  op.ie = ie;
  return op;
}

// Thin kick ops for orders 0..ix_max scaled by `scale` (and by (1 + g x) in a
// bend).  Each op carries up to 7 complex coefficients c_n = scale (b_n + i a_n)
// for orders n_lo..n_lo+k; the kick is dpx = -Re K, dpy = Im K with
// K = sum c_n (x + i y)^n  (ab_multipole_kick, bmad/modules/multipole_mod.f90:314-401).
void emit_kicks(std::vector<DevOp>& prog, int ie, const AB& ab, double scale, int type, double g = 0) {
  const bool bend = (type == OP_BEND_KICK || type == OP_BEND_EKICK);
  const int cap = bend ? 6 : 7;  // bend kicks keep g in c[13]
  if (ab.ix_max < 0) return;
  int lo = 0;
  while (lo <= ab.ix_max && ab.a[lo] == 0 && ab.b[lo] == 0) lo++;
  while (lo <= ab.ix_max) {
    int hi = std::min(ab.ix_max, lo + cap - 1);
    DevOp op = make_op(type, ie);
    op.n = lo;             // lowest order
    op.flags = hi - lo;    // number of Horner steps
    for (int n = lo; n <= hi; n++) {
      op.c[2 * (n - lo)] = scale * ab.b[n];
      op.c[2 * (n - lo) + 1] = scale * ab.a[n];
    }
    if (bend) op.c[13] = g;
    prog.push_back(op);
    lo = hi + 1;
    while (lo <= ab.ix_max && ab.a[lo] == 0 && ab.b[lo] == 0) lo++;
  }
}

// A Hwang edge with e = 0 and no fringe integral: tan(e) = 0 and sec(e) = 1,
// so gt, gt2, k1 tan(e) and the fringe factor vanish and gs2 = g_tot.
bool simple_edge(const DevOp& op) {
  return op.type == OP_HWANG && op.c[0] == 0 && op.c[1] == 0 && op.c[3] == 0 && op.c[4] == 0 && op.c[6] == 0 &&
         op.c[2] == op.c[5];
}

// Replace [HWANG] TOO_LARGE BEND TOO_LARGE [HWANG] at prog[b ..] by one
// OP_SBEND_FUSED when both edges are simple: the body constants stay in
// c[0 .. 10], the entrance edge's g_tot goes to c[11], minus the exit
// edge's to c[12] (the exit kick is the entrance kick with the sign of g_tot
// flipped).
void fuse_sbend(std::vector<DevOp>& prog, size_t b) {
  size_t n = prog.size() - b, i = b;
  if (n < 3 || n > 5) return;
  bool in_edge = prog[i].type == OP_HWANG;
  if (in_edge && !simple_edge(prog[i])) return;
  i += in_edge;
  if (i + 3 > prog.size() || prog[i].type != OP_TOO_LARGE || prog[i + 1].type != OP_BEND ||
      prog[i + 2].type != OP_TOO_LARGE)
    return;
  size_t tail = i + 3;
  bool out_edge = tail < prog.size();
  if (out_edge && (tail + 1 != prog.size() || !simple_edge(prog[tail]))) return;
  DevOp f = prog[i + 1];
  f.type = OP_SBEND_FUSED;
  if (in_edge) { f.flags |= F_HWANG_ENTER; f.c[11] = prog[b].c[5]; }
  if (out_edge) { f.flags |= F_EDGE_EXIT; f.c[12] = -prog[tail].c[5]; }
  prog.resize(b);
  prog.push_back(f);
}

bool is_misaligned(const bmx_ele_rec& e) {
  return e.x_offset != 0 || e.y_offset != 0 || e.z_offset != 0 || e.x_pitch != 0 || e.y_pitch != 0;
}

// offset_particle set/unset for a straight element (bmad/code/offset_particle.f90:150-330):
// OP_OFFSET_* c[0..2] = x_off, y_off, z_off; c[3] = L/2; c[4..12] = W row-major
// (W^-1 when setting); the element tilt follows (set) or precedes (unset) as OP_TILT.
// half_hv: offset_particle's half hkick / vkick (set_hvkicks, :300-304 and
// :393-397), a dipole kick between the frame change and the tilt, scaled by
// charge_dir = orientation x direction x rel_tracking_charge_to_mass.  rev:
// sign_z_vel = orientation x direction is -1, the affine form of
// emit_reversed_offset.
void emit_reversed_offset(std::vector<DevOp>& prog, int ie, const bmx_ele_rec& e, bool set);
void emit_offset(std::vector<DevOp>& prog, int ie, const bmx_ele_rec& e, bool rev, bool set, bool with_tilt,
                 bool half_hv = false, double charge_dir = 1.0) {
  bool mis = is_misaligned(e);
  bool tilt = with_tilt && e.tilt != 0;
  DevOp t = make_op(OP_TILT, ie);
  if (tilt) {
    t.c[0] = std::cos(set ? e.tilt : -e.tilt);
    t.c[1] = std::sin(set ? e.tilt : -e.tilt);
  }
  DevOp hv = make_op(OP_KICK, ie);  // px += hkick / 2, py += vkick / 2: K = -hkick / 2 + i vkick / 2
  hv.c[0] = -0.5 * charge_dir * e.hkick;
  hv.c[1] = 0.5 * charge_dir * e.vkick;
  bool kick = half_hv && (e.hkick != 0 || e.vkick != 0);
  if (!set && tilt) prog.push_back(t);
  if (!set && kick) prog.push_back(hv);
  if (mis && rev) {
    emit_reversed_offset(prog, ie, e, set);
  } else if (mis) {
    DevOp op = make_op(set ? OP_OFFSET_SET : OP_OFFSET_UNSET, ie);
    op.flags |= F_OFFSET_MOVE;
    if (e.x_pitch != 0 || e.y_pitch != 0) op.flags |= F_OFFSET_PITCH;
    op.c[0] = e.x_offset;
    op.c[1] = e.y_offset;
    op.c[2] = e.z_offset;
    op.c[3] = 0.5 * e.l;
    // floor_angles_to_w_mat(theta = x_pitch, phi = y_pitch, psi = 0)
    double st = std::sin(e.x_pitch), ct = std::cos(e.x_pitch), sp = std::sin(e.y_pitch), cp = std::cos(e.y_pitch);
    double w[9] = {ct, -st * sp, st * cp, 0.0, cp, sp, -st, -ct * sp, ct * cp};
    if (set) {  // inverse = transpose
      double wi[9] = {w[0], w[3], w[6], w[1], w[4], w[7], w[2], w[5], w[8]};
      std::memcpy(w, wi, sizeof(w));
    }
    for (int i = 0; i < 9; i++) op.c[4 + i] = w[i];
    prog.push_back(op);
  }
  if (set && kick) prog.push_back(hv);
  if (set && tilt) prog.push_back(t);
}

// Frame changes of offset_particle for a bend (bmad/code/offset_particle.f90:
// 208-235 set, :409-440 unset).  Every step there is a rotation S and a shift
// applied to the position r and (the rotation alone) to the orientation w, so
// the whole chain from (x, y, 0) is r = W (x, y, 0) + t with W the final w.
struct Affine {
  double w[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double t[3] = {0, 0, 0};
  void rotate(const double S[3][3]) {  // r -> S r, w -> S w
    double w2[3][3], t2[3];
    for (int i = 0; i < 3; i++) {
      t2[i] = S[i][0] * t[0] + S[i][1] * t[1] + S[i][2] * t[2];
      for (int j = 0; j < 3; j++) w2[i][j] = S[i][0] * w[0][j] + S[i][1] * w[1][j] + S[i][2] * w[2][j];
    }
    std::memcpy(w, w2, sizeof(w));
    std::memcpy(t, t2, sizeof(t));
  }
};

void mat3_unit(double m[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = (i == j) ? 1.0 : 0.0;
}

// rotate_mat (sim_utils/math/rotation_3d_mod.f90), left multiplication:
// m = R m with R the rotation by `angle` about the y (axis 1) or z (axis 2) axis
void rotate_left(double m[3][3], int axis, double angle) {
  if (angle == 0) return;
  double ca = std::cos(angle), sa = std::sin(angle);
  int a = 0, b = (axis == 1) ? 2 : 1;  // rows mixed: (x, z) for y, (x, y) for z
  double sgn = (axis == 1) ? -1.0 : 1.0;  // y: x' = ca x + sa z; z: x' = ca x - sa y
  for (int j = 0; j < 3; j++) {
    double u = m[a][j], v = m[b][j];
    m[a][j] = ca * u - sgn * sa * v;
    m[b][j] = sgn * sa * u + ca * v;
  }
}

// bend_shift (bmad/geometry/bend_shift.f90): the frame delta_s further along
// a reference arc of curvature g, rolled by ref_tilt when has_tilt
void bend_shift(Affine& p, double g, double delta_s, bool has_tilt, double ref_tilt) {
  double angle = delta_s * g;
  if (angle == 0) {
    p.t[2] -= delta_s;
    return;
  }
  double tlt = has_tilt ? ref_tilt : 0.0;
  double S[3][3];
  mat3_unit(S);
  if (tlt != 0) {
    rotate_left(S, 2, -ref_tilt);
    rotate_left(S, 1, angle);
    rotate_left(S, 2, ref_tilt);
  } else {
    rotate_left(S, 1, angle);
  }
  double sh = std::sin(0.5 * angle);
  double L[3] = {-2.0 * sh * sh / g, 0.0, -std::sin(angle) / g};  // cos_one(angle) = -2 sin^2(angle/2)
  if (has_tilt && ref_tilt != 0) {
    double c = std::cos(ref_tilt), s = std::sin(ref_tilt), l0 = L[0], l1 = L[1];
    L[0] = c * l0 - s * l1;
    L[1] = s * l0 + c * l1;
  }
  p.rotate(S);
  for (int i = 0; i < 3; i++) p.t[i] += L[i];
}

// floor_angles_to_w_mat (bmad/geometry/floor_angles_to_w_mat.f90), psi = 0
void pitch_w(double theta, double phi, bool inverse, double w[3][3]) {
  double st = std::sin(theta), ct = std::cos(theta), sp = std::sin(phi), cp = std::cos(phi);
  double m[3][3] = {{ct, -st * sp, st * cp}, {0.0, cp, sp}, {-st, -ct * sp, ct * cp}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) w[i][j] = inverse ? m[j][i] : m[i][j];
}

// offset_particle of a straight element with sign_z_vel = orientation x
// direction = -1 (offset_particle.f90; both orientation -1 forward and
// orientation +1 backward give the same frame change): the particle enters the body
// frame at s = L with the z component of its momentum pointing backwards, and
// the drift to the edge runs against s when setting.  Set: r = W^-1 (x - x_off,
// y - y_off, L/2 - z_off) + (0, 0, L/2), drift against s over L - r(3).
// Unset: r = W (x, y, -L/2) + (x_off, y_off, z_off + L/2), s -> L - r(3),
// drift over L - s.  As an OP_OFFSET_* with F_OFFSET_AFFINE: rows 0 and 1 of
// W with the momentum's z column negated, row 2 holding L - r(3), so the op's
// noref drift over (0 or L) - row 2 is either reference drift (a drift of
// length d against s is the noref drift of length -d).
void emit_reversed_offset(std::vector<DevOp>& prog, int ie, const bmx_ele_rec& e, bool set) {
  double len = e.l, lh = 0.5 * e.l, w[3][3], v[3], t[3];
  pitch_w(e.x_pitch, e.y_pitch, set, w);
  if (set) { v[0] = -e.x_offset; v[1] = -e.y_offset; v[2] = lh - e.z_offset; }
  else { v[0] = 0.0; v[1] = 0.0; v[2] = -lh; }
  for (int i = 0; i < 3; i++) t[i] = w[i][0] * v[0] + w[i][1] * v[1] + w[i][2] * v[2];
  if (set) {
    t[2] += lh;
  } else {
    t[0] += e.x_offset;
    t[1] += e.y_offset;
    t[2] += e.z_offset + lh;
  }
  DevOp op = make_op(set ? OP_OFFSET_SET : OP_OFFSET_UNSET, ie);
  op.flags |= F_OFFSET_MOVE | F_OFFSET_PITCH | F_OFFSET_AFFINE;
  op.c[0] = t[0];
  op.c[1] = t[1];
  op.c[2] = len - t[2];
  op.c[3] = lh;
  double* c = op.c + 4;
  c[0] = w[0][0]; c[1] = w[0][1]; c[2] = -w[0][2];
  c[3] = w[1][0]; c[4] = w[1][1]; c[5] = -w[1][2];
  c[6] = -w[2][0]; c[7] = -w[2][1]; c[8] = -w[2][2];
  prog.push_back(op);
}

// offset_particle of a misaligned bend (g != 0 or ref_tilt != 0), orientation
// +1, drift to the edge: one OP_OFFSET_* with F_OFFSET_AFFINE, c[0..2] = t,
// c[3] = L/2, c[4..12] = W row-major.  rev: direction -1 (sign_z_vel = -1):
// the arc shifts to the centre and to the edge change sign, the momentum's z
// column is negated and row 2 is rewritten so that the op's noref drift over
// (0 or L) - row 2 is the reference's drift against s (set: over the chain's
// r(3); unset: over r(3) as well, the edge being s = 0).
void emit_bend_offset(std::vector<DevOp>& prog, int ie, const bmx_ele_rec& e, bool set, bool rev) {
  double g = e.g, L_half = 0.5 * e.l, rt = e.ref_tilt, sz = rev ? -1.0 : 1.0;
  double ws[3][3], wt[3][3];
  Affine p;
  if (set) {
    bend_shift(p, g, sz * L_half, true, rt);  // orientation x ds_center, ds_center = L/2 - s_pos0
    p.t[0] -= e.x_offset;
    p.t[1] -= e.y_offset;
    p.t[2] -= e.z_offset;
    pitch_w(e.x_pitch, e.y_pitch, true, ws);  // ele_misalignment_L_S_calc, transposed
    p.rotate(ws);
    if (rt != 0) {
      bend_shift(p, g, -L_half, true, rt);
      mat3_unit(wt);
      rotate_left(wt, 2, -rt);  // w_mat_for_tilt(-ref_tilt)
      p.rotate(wt);
      bend_shift(p, g, L_half, false, 0.0);
    }
    bend_shift(p, g, -sz * L_half, false, 0.0);  // drift_to upstream: r(3) += (1 - sign_z_vel) L/2, folded below
  } else {
    bend_shift(p, g, -sz * L_half, false, 0.0);  // ds_center = L/2 - s_pos0
    if (rt != 0) {
      bend_shift(p, g, -L_half, false, 0.0);
      mat3_unit(wt);
      rotate_left(wt, 2, rt);  // w_mat_for_tilt(ref_tilt)
      p.rotate(wt);
      bend_shift(p, g, L_half, true, rt);
    }
    pitch_w(e.x_pitch, e.y_pitch, false, ws);
    p.rotate(ws);
    p.t[0] += e.x_offset;
    p.t[1] += e.y_offset;
    p.t[2] += e.z_offset;
    bend_shift(p, g, sz * L_half, true, rt);  // drift_to downstream
    if (!rev) p.t[2] += e.l;                  // r(3) += (1 + sign_z_vel) L/2
  }
  DevOp op = make_op(set ? OP_OFFSET_SET : OP_OFFSET_UNSET, ie);
  op.flags |= F_OFFSET_MOVE | F_OFFSET_PITCH | F_OFFSET_AFFINE;
  for (int i = 0; i < 3; i++) op.c[i] = p.t[i];
  op.c[3] = L_half;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) op.c[4 + 3 * i + j] = p.w[i][j];
  if (rev) {
    op.c[4 + 2] = -op.c[4 + 2];  // p_z = -sqrt(...)
    op.c[4 + 5] = -op.c[4 + 5];
    for (int j = 0; j < 3; j++) op.c[4 + 6 + j] = -op.c[4 + 6 + j];
    op.c[2] = (set ? 0.0 : e.l) - p.t[2];
  }
  prog.push_back(op);
}

}  // namespace

namespace {

// Number of leading constants c[0 ..] an op's body uses (14 when not
// listed); F_SHIFT_OUT offsets sit apart, at c[SHIFT_C ..].
int used_consts(const DevOp& op) {
  int nh = op.flags & 0xff;
  switch (op.type) {
    case OP_ELE_END: return 0;
    case OP_DRIFT: case OP_ZBETA: case OP_RAD: return 1;
    case OP_DRIFT_BACK: return 2;
    case OP_QUAD: return 3;
    case OP_TILT: case OP_HARD_MULT_EDGE: return 2;
    case OP_KICK: case OP_EKICK: case OP_EFRINGE: return 2 * (nh + 1);
    case OP_BEND_KICK: case OP_BEND_EKICK: return 14;
    case OP_KD_SLICES: return 1 + 2 * (nh + 1);
    case OP_RFCAVITY:
      if (op.flags & (F_RF_ABS | F_RF_BACK)) return 11;
      return (op.flags & (F_COUPLER_ENTER | F_COUPLER_EXIT)) ? 8 : 5;
    default: return 14;
  }
}

}  // namespace

// Compile one variant.  Returns false and fills *err for unsupported input.
bool compile_variant(const bmx_ele_rec* eles, int n_ele, const bmx_ref& ref, const bmx_config& cfg,
                     std::vector<DevOp>& prog, std::vector<int32_t>& ele_op_begin, LatticeConsts* lc,
                     std::string* err, const RadInput* rad, std::vector<double>* rad_table) {
  char buf[256];
  if (n_ele <= 0) { *err = "empty lattice"; return false; }
  double p0c = eles[0].p0c, e_tot = eles[0].e_tot;
  if (!(p0c > 0) || !(e_tot > 0)) { *err = "element 0 has no reference energy (p0c, e_tot)"; return false; }
  double mc2 = species_mass(ref.species);
  if (mc2 == 0) { *err = "unknown species"; return false; }
  lc->p0c = p0c;
  lc->e_tot = e_tot;
  lc->mc2 = mc2;
  lc->beta0 = p0c / e_tot;
  lc->species = ref.species;
  // the tracked species: the reference species or its antiparticle, whose
  // rel_tracking_charge_to_mass (rel_tracking_charge_to_mass.f90:15-35) is the
  // charge ratio; the device's mass, beta and energy constants stay valid
  int tracked = ref.tracked_species ? ref.tracked_species : ref.species;
  if (species_mass(tracked) != mc2 || species_charge(tracked) == 0) {
    *err = "the tracked species must be the reference species or its antiparticle";
    return false;
  }
  const double rel_charge = species_charge(tracked) / species_charge(ref.species);
  // tracking direction (bmx_config.direction): backward programs run the
  // elements last to first (track_bunch, beam_mod.f90:131-147), each element
  // with direction = -1 in its maps
  if (cfg.direction != 0 && cfg.direction != 1 && cfg.direction != -1) {
    *err = "direction must be +1 or -1";
    return false;
  }
  const int dir = (cfg.direction == -1) ? -1 : 1;
  const double beta_ref_drift = p0c / std::sqrt(p0c * p0c + mc2 * mc2);  // track_a_drift's backward beta_ref
  lc->mc2sq = mc2 * mc2;
  lc->p0c_sq = p0c * p0c;
  // reference time of one pass (the sum of delta_ref_time, in extended
  // precision like the reference's quad-precision particle time) and the
  // per-call time fields at their whole-turn forward values
  long double t_lat_ld = 0.0L;
  for (int k = 0; k < n_ele; k++) t_lat_ld += (long double)eles[k].delta_ref_time;
  lc->t_lat = (double)t_lat_ld;
  lc->abs_time = cfg.absolute_time_tracking ? 1 : 0;
  lc->n_ele = n_ele;
  lc->tt_sign = 1;
  lc->tt_d = 0.0;
  lc->ix_first = 0;
  lc->t_start = lc->t_end = 0.0;
  lc->rt = nullptr;
  {
    double b0 = lc->beta0, me = mc2 / e_tot;
    lc->lez_k = mc2 * b0 * b0;
    lc->lez_thr = 3e-7 * e_tot;
    lc->lez_a = 1.5 * b0 * b0;
    lc->lez_f = b0 * b0 * (2.0 * b0 * b0 - me * me / 2.0);
    lc->lez_me2 = me * me;
  }

  for (int k = 0; k < n_ele; k++) {
    const int ie = (dir == 1) ? k : n_ele - 1 - k;
    const bmx_ele_rec& e = eles[ie];
    auto fail = [&](const char* what) {
      std::snprintf(buf, sizeof(buf), "element %d (key %d): %s", ie, e.key, what);
      *err = buf;
      return false;
    };
    ele_op_begin.push_back((int32_t)prog.size());
    if (e.p0c != p0c || e.p0c_start != p0c || e.e_tot != e_tot)
      return fail("reference energy changes along the lattice (only rfcavity-style lattices are supported)");
    if (e.orientation != 1 && e.orientation != -1) return fail("orientation must be +1 or -1");
    // charge_dir of the magnetic fields: orientation x direction x
    // rel_tracking_charge_to_mass; with orientation x direction = -1 the
    // particle meets the element's exit end at its first track edge
    // (physical_ele_end.f90), which fringe_here and the aperture and coupler
    // switches test
    const bool rev = e.orientation * dir == -1;
    const double cdir = e.orientation * dir * rel_charge;
    // a drift of the body: track_a_drift with include_ref_motion, forward or backward
    auto emit_drift = [&](double len) {
      DevOp op = make_op(dir == 1 ? OP_DRIFT : OP_DRIFT_BACK, ie);
      op.c[0] = len;
      if (dir == -1) op.c[1] = beta_ref_drift;
      prog.push_back(op);
    };
    auto at_end = [&](int where, bool first_edge) {
      int phys = (first_edge != rev) ? BMX_ENTRANCE_END : BMX_EXIT_END;
      return where == phys || where == BMX_BOTH_ENDS;
    };
    auto scale_ab = [&](AB& t, double f) {
      for (int n = 0; n < BMX_N_POLE + 1; n++) { t.a[n] *= f; t.b[n] *= f; }
    };
    AB ae;  // electric multipoles
    {
      std::string ee;
      if (!electric_ab(e, &ae, &ee)) return fail(ee.c_str());
      int lo_e = 0;
      while (lo_e <= ae.ix_max && ae.a[lo_e] == 0 && ae.b[lo_e] == 0) lo_e++;
      if (ae.ix_max >= 0 && ae.ix_max - lo_e >= ((e.key == BMX_KEY_SBEND) ? 6 : 7))
        return fail("electric multipoles spanning more than 7 orders (6 in a bend) are not supported on the device path");
    }
    // collimator phase-space limits (check_aperture_limit.f90, ecollimator /
    // rcollimator only; other elements ignore these fields)
    bool is_coll = e.key == BMX_KEY_ECOLLIMATOR || e.key == BMX_KEY_RCOLLIMATOR;
    bool ps_set = is_coll && (e.px_aperture_width2 != 0 || e.py_aperture_width2 != 0 || e.z_aperture_width2 != 0 ||
                              e.pz_aperture_width2 != 0);
    const double ps_c[8] = {e.px_aperture_width2, e.px_aperture_center, e.py_aperture_width2, e.py_aperture_center,
                            e.z_aperture_width2,  e.z_aperture_center,  e.pz_aperture_width2, e.pz_aperture_center};
    int ps_flags = ps_set ? (e.key == BMX_KEY_ECOLLIMATOR ? F_PS_ELLIPTICAL : 0) : 0;
    if (e.exact_multipoles != 0 && e.exact_multipoles != 1) return fail("exact_multipoles must be off");
    if (e.num_steps < 1 && e.l != 0) return fail("num_steps must be >= 1");

    // entrance aperture (check_aperture_limit at the first track edge)
    bool ap_set = (e.x1_limit != 0 || e.x2_limit != 0 || e.y1_limit != 0 || e.y2_limit != 0);
    int ap_flags = (e.aperture_type == BMX_APERTURE_ELLIPTICAL) ? F_APERTURE_ELLIPTICAL : 0;
    if (e.aperture_type != BMX_APERTURE_ELLIPTICAL && e.aperture_type != BMX_APERTURE_RECTANGULAR)
      return fail("aperture_type must be rectangular or elliptical");
    if (e.aperture_type == BMX_APERTURE_ELLIPTICAL && ap_set &&
        ((e.x1_limit == 0) != (e.x2_limit == 0) || (e.y1_limit == 0) != (e.y2_limit == 0)))
      return fail("elliptical aperture needs all four limits");
    // aperture constants at c[k0 ..]: rectangular x1, x2, y1, y2; elliptical
    // (elliptical_params_setup) cx, cy, 1/wx, 1/wy, wx, wy -- the reciprocal
    // widths for the device's division-free screen
    auto ap_consts = [&](DevOp& op, int k0) {
      double x1 = -e.x1_limit, x2 = e.x2_limit, y1 = -e.y1_limit, y2 = e.y2_limit;
      double* c = op.c + k0;
      if (e.aperture_type == BMX_APERTURE_ELLIPTICAL) {
        double wx = (x2 - x1) / 2, cx = (x2 + x1) / 2, wy = (y2 - y1) / 2, cy = (y2 + y1) / 2;
        if (wx == 0) wx = cfg.max_aperture_limit;
        if (wy == 0) wy = cfg.max_aperture_limit;
        wx = std::max(wx, 0.0);
        wy = std::max(wy, 0.0);
        c[0] = cx; c[1] = cy; c[2] = 1.0 / wx; c[3] = 1.0 / wy; c[4] = wx; c[5] = wy;
      } else {
        c[0] = x1; c[1] = x2; c[2] = y1; c[3] = y2;
      }
    };
    if (cfg.aperture_limit_on && (ap_set || ps_set) && at_end(e.aperture_at, true)) {
      DevOp op = make_op(OP_APERTURE_IN, ie);
      op.flags = ap_flags;
      ap_consts(op, 0);
      if (ps_set) {
        op.flags |= F_APERTURE_PS | ps_flags;
        for (int i = 0; i < 8; i++) op.c[6 + i] = ps_c[i];
      }
      prog.push_back(op);
    }

    // track1_radiation at the element's two edges (track1.f90:198-201,
    // :316-319): the first half after the entrance aperture, the second
    // after orbit_too_large (F_RAD_EXIT) and before the exit aperture.  No
    // step for the elements track1_radiation skips (radiation_mod.f90:87-91)
    // or for a map that changes nothing.
    const bool rad_here = rad && (rad->cfg.damping_on || rad->cfg.fluctuations_on) && e.is_on && e.l != 0 &&
                          e.key != BMX_KEY_DRIFT && e.key != BMX_KEY_PIPE && e.key != BMX_KEY_MARKER &&
                          e.key != BMX_KEY_MULTIPOLE && e.key != BMX_KEY_AB_MULTIPOLE;
    auto emit_rad = [&](bool exit_edge) -> bool {
      if (!rad_here) return true;
      const bmx_rad_map& rm = rad->maps[2 * ie + (exit_edge ? 1 : 0)];
      const double sc = rad->cfg.synch_rad_scale;
      double m[RAD_STRIDE] = {0.0};
      bool damp = false, fluct = false;
      for (int i = 0; i < 6; i++) m[RAD_REF + i] = rm.ref_orb[i];
      if (rad->cfg.damping_on) {
        for (int i = 0; i < 36; i++) { m[RAD_DAMP + i] = sc * rm.damp_dmat[i]; damp |= m[RAD_DAMP + i] != 0.0; }
        if (!rad->cfg.zero_average)
          for (int i = 0; i < 6; i++) { m[RAD_VEC + i] = sc * rm.xfer_damp_vec[i]; damp |= m[RAD_VEC + i] != 0.0; }
      }
      if (rad->cfg.fluctuations_on) {
        int k = RAD_STOC;
        for (int r = 0; r < 6; r++)
          for (int c = 0; c <= r; c++) { m[k] = sc * rm.stoc_mat[r * 6 + c]; fluct |= m[k] != 0.0; k++; }
      }
      if (!damp && !fluct) return true;
      if (dir == -1) return fail("radiation is tracked forward only");
      DevOp op = make_op(OP_RAD, ie);
      op.flags = (damp ? F_RAD_DAMP : 0) | (fluct ? F_RAD_FLUCT : 0) | (exit_edge ? F_RAD_EXIT : 0);
      op.c[0] = (double)(rad_table->size() / RAD_STRIDE);
      rad_table->insert(rad_table->end(), m, m + RAD_STRIDE);
      if (fluct) lc->rad_fluct = 1;
      prog.push_back(op);
      return true;
    };
    if (!emit_rad(false)) return false;

    int key = e.key;
    if (!e.is_on) {
      if (key == BMX_KEY_MULTIPOLE || key == BMX_KEY_AB_MULTIPOLE) key = BMX_KEY_MARKER;
      else if (key != BMX_KEY_SBEND) key = BMX_KEY_DRIFT;
    }
    // hard_multipole_edge_kick of a quadrupole (n = 1, b = k1) or sextupole
    // (n = 2, b = k2) at one track edge, where fringe_type and fringe_at put
    // one (fringe_mod.f90:714-721, fringe_here.f90:15-35); the soft
    // quadrupole edge kick has no strength here (fq1 = fq2 = 0)
    auto hard_edge = [&](int n, double b, bool enter) {
      if (e.fringe_type != BMX_FRINGE_HARD_EDGE_ONLY && e.fringe_type != BMX_FRINGE_FULL) return;
      if (!at_end(e.fringe_at, enter)) return;
      DevOp op = make_op(OP_HARD_MULT_EDGE, ie);
      op.n = n;
      double c0 = cdir * b / (4.0 * (n + 2));
      op.c[0] = enter ? -c0 : c0;
      op.c[1] = (double)(n + 3) / (double)(n + 1);
      prog.push_back(op);
    };
    // ab_multipole_kicks of the electric poles (multipole_mod.f90:82-174,
    // ab_multipole_kick's electric branch :338-345): the ab kick of a f and
    // -b f with f = charge / (beta p0c) x scale, as kick coefficients
    // scale charge / p0c (-b, a); the device divides by beta and updates pz,
    // beta and z (OP_EKICK)
    double charge = species_charge(tracked);
    auto emit_ekicks = [&](double scale) {
      if (ae.ix_max < 0) return;
      AB t = ae;
      double f = scale * charge / p0c;
      for (int n = 0; n <= ae.ix_max; n++) { t.b[n] = -f * ae.b[n]; t.a[n] = f * ae.a[n]; }
      emit_kicks(prog, ie, t, 1.0, OP_EKICK);
    };
    // electric_longitudinal_fringe (apply_element_edge_kick.f90:283-329) at
    // one edge, where fringe_at puts an edge kick: the energy kick
    // -ff phi = Re sum_i C_i z^(i+1), C_i = ff (b_i - i a_i) / (i+1),
    // ff = at_sign charge (at_sign +1 entering, -1 leaving) (OP_EFRINGE)
    auto emit_efringe = [&](bool enter) {
      if (e.l == 0 || ae.ix_max < 0) return;
      if (!at_end(e.fringe_at, enter)) return;
      double ff = (enter ? 1.0 : -1.0) * charge;
      AB t;
      std::memset(&t, 0, sizeof(t));
      for (int i = 0; i <= ae.ix_max; i++) {
        t.b[i + 1] = ff * ae.b[i] / (i + 1);
        t.a[i + 1] = -ff * ae.a[i] / (i + 1);
      }
      t.ix_max = ae.ix_max + 1;
      emit_kicks(prog, ie, t, 1.0, OP_EFRINGE);
    };
    std::string e2;
    switch (key) {
      case BMX_KEY_MARKER:
        break;
      case BMX_KEY_DRIFT: {
        if (e.l != 0) emit_drift(e.l);
        break;
      }
      case BMX_KEY_MULTIPOLE: case BMX_KEY_AB_MULTIPOLE: {
        // track1_bmad.f90:142-166: offsets without tilt, kt kicks (tilt folded into ab)
        AB ab;
        if (!magnetic_ab(e, false, &ab, nullptr, &e2)) return fail(e2.c_str());
        if (key == BMX_KEY_AB_MULTIPOLE && e.tilt != 0) {  // use_ele_tilt = true
          for (int n = 0; n <= ab.ix_max; n++) {
            if (ab.a[n] == 0 && ab.b[n] == 0) continue;
            double an = ab.a[n], bn = ab.b[n], c = std::cos((n + 1) * e.tilt), s = std::sin((n + 1) * e.tilt);
            ab.b[n] = bn * c + an * s;
            ab.a[n] = -bn * s + an * c;
          }
        }
        if (key == BMX_KEY_MULTIPOLE && e.tilt != 0) {  // tilt_tot adds to every Tn
          for (int n = 0; n < BMX_N_POLE; n++) {
            double an = ab.a[n], bn = ab.b[n];
            if (an == 0 && bn == 0) continue;
            double c = std::cos((n + 1) * e.tilt), s = std::sin((n + 1) * e.tilt);
            ab.b[n] = bn * c + an * s;
            ab.a[n] = -bn * s + an * c;
          }
        }
        bool mis = is_misaligned(e);
        if (mis) emit_offset(prog, ie, e, rev, true, false);
        if (key == BMX_KEY_MULTIPOLE && e.has_mag_poles && e.a[0] != 0) {
          DevOp op = make_op(OP_KICK_REF0, ie);
          double t = e.b[0] + e.tilt;
          op.c[0] = cdir * e.a[0] * std::cos(t);
          op.c[1] = cdir * e.a[0] * std::sin(t);
          prog.push_back(op);
          ab.a[0] = ab.b[0] = 0;
          ab.ix_max = max_nonzero(ab.a, ab.b, BMX_N_POLE - 1);
        }
        scale_ab(ab, cdir);
        emit_kicks(prog, ie, ab, 1.0, OP_KICK);
        if (mis) emit_offset(prog, ie, e, rev, false, false);
        break;
      }
      case BMX_KEY_QUADRUPOLE: {
        // track_a_quadrupole.f90:17-148
        AB ab;
        double b1;
        if (!magnetic_ab(e, true, &ab, &b1, &e2)) return fail(e2.c_str());
        scale_ab(ab, cdir);
        b1 *= cdir;
        int n_step = 1;
        if (ab.ix_max > -1 || ae.ix_max > -1) n_step = std::max(nint(std::fabs(e.l) / (e.l / e.num_steps)), 1);
        if (e.l == 0) n_step = 0;
        double step_len = e.l / (n_step > 0 ? n_step : 1);
        double r_step = 1.0 / (n_step > 0 ? n_step : 1);
        emit_offset(prog, ie, e, rev, true, true);
        hard_edge(1, e.k1, true);
        emit_efringe(true);
        emit_kicks(prog, ie, ab, r_step / 2, OP_KICK);
        emit_ekicks(step_len / 2);
        for (int i = 1; i <= n_step; i++) {
          DevOp op = make_op(OP_QUAD, ie);
          op.c[0] = step_len;
          op.c[1] = b1 / e.l;  // k1 * rel_p
          op.c[2] = op.c[1] * step_len * step_len;
          prog.push_back(op);
          if (dir == -1) {  // track_a_quadrupole.f90: z -= 2 c beta delta_ref_time, every step
            DevOp zb = make_op(OP_ZBETA, ie);
            zb.c[0] = 2.0 * kCLight * e.delta_ref_time;
            prog.push_back(zb);
          }
          emit_kicks(prog, ie, ab, (i == n_step) ? r_step / 2 : r_step, OP_KICK);
          emit_ekicks((i == n_step) ? step_len / 2 : step_len);
        }
        emit_efringe(false);
        hard_edge(1, e.k1, false);
        emit_offset(prog, ie, e, rev, false, true);
        break;
      }
      case BMX_KEY_SEXTUPOLE: case BMX_KEY_OCTUPOLE: case BMX_KEY_THICK_MULTIPOLE: case BMX_KEY_HKICKER:
      case BMX_KEY_VKICKER: case BMX_KEY_KICKER: case BMX_KEY_INSTRUMENT: case BMX_KEY_MONITOR:
      case BMX_KEY_ECOLLIMATOR: case BMX_KEY_RCOLLIMATOR: case BMX_KEY_PIPE: {
        // track_a_thick_multipole.f90:18-133
        AB ab;
        if (!magnetic_ab(e, true, &ab, nullptr, &e2)) return fail(e2.c_str());
        scale_ab(ab, cdir);
        int n_step = (ab.ix_max == -1 && ae.ix_max == -1) ? 1 : std::max(nint(std::fabs(e.l / (e.l / e.num_steps))), 1);
        double step_len = e.l / n_step, r_step = 1.0 / n_step;
        emit_offset(prog, ie, e, rev, true, true);
        if (key == BMX_KEY_SEXTUPOLE) hard_edge(2, e.k2, true);
        emit_efringe(true);
        int lo = 0;
        while (lo <= ab.ix_max && ab.a[lo] == 0 && ab.b[lo] == 0) lo++;
        if (ab.ix_max >= 0 && step_len != 0 && ab.ix_max - lo < 6 && ae.ix_max < 0 && dir == 1) {
          // one op for the body; the half kicks are the full-step kick x 0.5,
          // which is exact (the reference scales by r_step / 2)
          DevOp op = make_op(OP_KD_SLICES, ie);
          op.n = n_step;
          op.flags = (ab.ix_max - lo) | (lo << 8);
          op.c[0] = step_len;
          for (int n = lo; n <= ab.ix_max; n++) {
            op.c[1 + 2 * (n - lo)] = r_step * ab.b[n];
            op.c[2 + 2 * (n - lo)] = r_step * ab.a[n];
          }
          prog.push_back(op);
        } else {
          emit_kicks(prog, ie, ab, r_step / 2, OP_KICK);
          emit_ekicks(step_len / 2);
          for (int i = 1; i <= n_step; i++) {
            if (step_len != 0) emit_drift(step_len);
            emit_kicks(prog, ie, ab, (i == n_step) ? r_step / 2 : r_step, OP_KICK);
            emit_ekicks((i == n_step) ? step_len / 2 : step_len);
          }
        }
        emit_efringe(false);
        if (key == BMX_KEY_SEXTUPOLE) hard_edge(2, e.k2, false);
        emit_offset(prog, ie, e, rev, false, true);
        break;
      }
      case BMX_KEY_SBEND: {
        // track_a_bend.f90:18-333
        // offset_particle (has_misalign: offsets, pitches or ref_tilt); a bend
        // ignores the element tilt (it has ref_tilt instead)
        bool mis = is_misaligned(e), has_mis = mis || e.ref_tilt != 0;
        auto bend_offset = [&](bool set) {
          if (!has_mis) return;
          if (!mis) {  // tilt_coords(+-ref_tilt) alone (offset_particle.f90:249-251, :472-474)
            DevOp t = make_op(OP_TILT, ie);
            t.c[0] = std::cos(set ? e.ref_tilt : -e.ref_tilt);
            t.c[1] = std::sin(set ? e.ref_tilt : -e.ref_tilt);
            prog.push_back(t);
          } else if (e.g == 0 && e.ref_tilt == 0) {
            emit_offset(prog, ie, e, rev, set, false);  // straight-element frame change
          } else {
            emit_bend_offset(prog, ie, e, set, rev);
          }
        };
        bend_offset(true);
        if (e.fringe_type == BMX_FRINGE_FULL || e.fringe_type == BMX_FRINGE_SAD_FULL || e.fringe_type == BMX_FRINGE_SOFT_EDGE_ONLY)
          return fail("sbend fringe_type full / sad_full / soft_edge_only is not supported on the device path");
        auto emit_edge = [&](bool enter) {
          if (e.fringe_type == BMX_FRINGE_NONE || !at_end(e.fringe_at, enter)) return;
          bool phys_in = enter != rev;  // e1, fint, hgap at the physical entrance end
          double g_tot = e.is_on ? (e.g + e.dg) * cdir : 0.0, k1 = e.is_on ? e.k1 * cdir : 0.0;
          double ang = phys_in ? e.e1 : e.e2;
          double fint_gap = phys_in ? e.fint * e.hgap : e.fintx * e.hgapx;
          if (e.fringe_type == BMX_FRINGE_HARD_EDGE_ONLY) fint_gap = 0;
          double ce = std::cos(ang), se = std::sin(ang), te = se / ce, sec = 1 / ce;
          if (e.fringe_type == BMX_FRINGE_LINEAR_EDGE) {
            double fint = phys_in ? e.fint : e.fintx, hgap = phys_in ? e.hgap : e.hgapx;
            DevOp op = make_op(OP_LINEAR_EDGE, ie);
            op.c[0] = g_tot * te;
            op.c[1] = (fint == 0) ? -op.c[0] : -g_tot * std::tan(ang - 2 * fint * g_tot * hgap * (1 + se * se) / ce);
            prog.push_back(op);
            return;
          }
          DevOp op = make_op(OP_HWANG, ie);
          if (enter) op.flags |= F_HWANG_ENTER;
          op.c[0] = g_tot * te;                                   // gt
          op.c[1] = g_tot * te * te;                              // gt2
          op.c[2] = g_tot * sec * sec;                            // gs2
          op.c[3] = k1 * te;                                      // k1_tane
          op.c[4] = 2 * fint_gap * op.c[2] * g_tot * sec * (1 + se * se);  // fg_factor
          op.c[5] = g_tot;
          op.c[6] = te;
          prog.push_back(op);
        };
        size_t body_begin = prog.size();
        emit_edge(true);
        emit_efringe(true);
        prog.push_back(make_op(OP_TOO_LARGE, ie));
        AB ab;
        double b1;
        if (!magnetic_ab(e, true, &ab, &b1, &e2)) return fail(e2.c_str());
        b1 *= cdir;
        if (std::fabs(b1) < 1e-10) { ab.b[1] = b1; b1 = 0; }
        double g = e.g, dg;
        if (e.l == 0) dg = 0;
        else { dg = ab.b[0] / e.l; ab.b[0] = 0; }
        double g_tot = (g + dg) * cdir;
        if (!e.is_on) { dg = -g; g_tot = 0; b1 = 0; }
        int ix_mag_max = ab.ix_max;
        int n_step = 1;
        if (ix_mag_max > -1 || ae.ix_max > -1) n_step = std::max(nint(std::fabs(e.l) / (e.l / e.num_steps)), 1);
        double r_step = 1.0 / n_step, step_len = e.l / n_step, angle = g * step_len;
        AB abk = ab;  // kicks run over orders 0..ix_mag_max as computed before dg/b1 were pulled out
        abk.ix_max = ix_mag_max;
        scale_ab(abk, cdir);  // c_dir / charge_of(ref) of apply_multipole_kicks
        // bend_apply_multipole_kicks: the magnetic kick, then the electric one
        // from the fields at the same x, y (coefficients step_len coef charge / p0c (b, -a))
        auto bend_kicks = [&](double coef) {
          if (ix_mag_max > -1) emit_kicks(prog, ie, abk, coef * r_step, OP_BEND_KICK, g);
          if (ae.ix_max > -1) {
            AB t = ae;
            double f = step_len * coef * charge / p0c;
            for (int n = 0; n <= ae.ix_max; n++) { t.b[n] = f * ae.b[n]; t.a[n] = -f * ae.a[n]; }
            emit_kicks(prog, ie, t, 1.0, OP_BEND_EKICK, g);
          }
        };
        bend_kicks(0.5);
        for (int n = 1; n <= n_step; n++) {
          if (b1 != 0) {
            DevOp op = make_op(OP_BEND_K1, ie);
            double k1 = b1 / e.l;
            op.c[0] = step_len; op.c[1] = g; op.c[2] = g_tot; op.c[3] = k1;
            op.c[4] = dir * e.orientation;  // the sign of z0 (direction x orientation)
            prog.push_back(op);
          } else if ((g == 0 && dg == 0) || step_len == 0) {
            if (step_len != 0) emit_drift(step_len);
          } else {
            DevOp op = make_op(OP_BEND, ie);
            double sinc_a = (std::fabs(angle) < 1e-8) ? 1.0 : std::sin(angle) / angle;
            double s2 = (std::fabs(angle / 2) < 1e-8) ? 1.0 : std::sin(angle / 2) / (angle / 2);
            double cosc_a = (std::fabs(angle) < 1e-8) ? 0.5 : 0.5 * s2 * s2;
            double x2 = angle * angle;
            double sincc_a = (std::fabs(angle) < 0.1)
                                 ? 1.0 / 6.0 + x2 * (-1.0 / 120.0 + x2 * (1.0 / 5040.0 + x2 * (-1.0 / 362880.0)))
                                 : (angle - std::sin(angle)) / (angle * x2);
            op.c[0] = step_len; op.c[1] = g; op.c[2] = g_tot; op.c[3] = angle;
            op.c[4] = std::cos(angle); op.c[5] = std::sin(angle); op.c[6] = sinc_a; op.c[7] = cosc_a;
            op.c[8] = sincc_a; op.c[9] = lc->beta0; op.c[10] = dg; op.c[13] = step_len / lc->beta0;
            prog.push_back(op);
          }
          bend_kicks(n == n_step ? 0.5 : 1.0);
        }
        prog.push_back(make_op(OP_TOO_LARGE, ie));
        emit_efringe(false);
        emit_edge(false);
        fuse_sbend(prog, body_begin);
        bend_offset(false);
        break;
      }
      case BMX_KEY_RFCAVITY: {
        // track_a_rfcavity.f90:17-146 (standing wave).  Its offset_particle
        // calls keep the default set_hvkicks, so hkick and vkick act twice:
        // as half kicks there and in the ab kicks (multipole_ele_to_ab with
        // include_kicks, :60-66).  The rf phase runs on relative time, or on
        // absolute time with bmx_config.absolute_time_tracking
        // (particle_rf_time.f90:72-89); backward tracking drifts backward
        // between the kicks.
        if (e.rf_frequency == 0 && e.voltage != 0) return fail("rfcavity with zero rf_frequency");
        AB ab;
        if (!magnetic_ab(e, true, &ab, nullptr, &e2)) return fail(e2.c_str());
        scale_ab(ab, cdir);
        emit_offset(prog, ie, e, rev, true, true, true, cdir);
        emit_kicks(prog, ie, ab, 0.5, OP_KICK);
        emit_ekicks(0.5 * e.l);
        int n_slice = std::max(1, nint(std::fabs(e.l) / (e.l / (e.num_steps > 0 ? e.num_steps : 1))));
        if (e.l == 0) n_slice = 1;
        DevOp op = make_op(OP_RFCAVITY, ie);
        op.n = n_slice;
        op.c[0] = e.l;
        // e_accel_field * charge_dir, charge_dir = rel_tracking_charge x orientation (no direction)
        op.c[1] = (e.is_on ? e.voltage : 0.0) * cdir * dir;
        op.c[3] = e.rf_frequency;
        op.c[4] = (e.rf_frequency != 0) ? 0.5 / e.rf_frequency : 0.0;
        if (cfg.absolute_time_tracking) {
          // time = t - ref_time_start: the lattice-frame reference time of the
          // first track edge less ref_time_start is 0 forward and
          // delta_ref_time backward (the particle enters at the exit end);
          // whole lattice periods enter as frac(f t_lat) cycles each
          op.flags |= F_RF_ABS;
          op.c[8] = (dir == 1) ? 0.0 : e.delta_ref_time;
          long double ft = (long double)e.rf_frequency * t_lat_ld;
          op.c[9] = (double)(ft - floorl(ft));
          op.c[2] = e.phi0;
          if (e.orientation * dir == -1) {  // the backward wave: phase0 - 2 pi f dt_ref (:86-88)
            long double fd = (long double)e.rf_frequency * ((long double)e.l / (kCLight * (long double)lc->beta0));
            op.c[2] = (double)((long double)e.phi0 - (fd - floorl(fd)));
          }
        } else {
          op.c[2] = e.phi0 + e.phi0_multipass;
        }
        if (dir == -1) {
          op.flags |= F_RF_BACK;
          op.c[10] = beta_ref_drift;
        }
        if (e.coupler_strength != 0) {  // rf_coupler_kick.f90:41-68
          bool both = e.coupler_at == BMX_BOTH_ENDS;
          if (at_end(e.coupler_at, true)) op.flags |= F_COUPLER_ENTER;
          if (at_end(e.coupler_at, false)) op.flags |= F_COUPLER_EXIT;
          double grad = (e.l == 0) ? 0.0 : e.voltage / e.l;
          double dp = grad * e.coupler_strength;
          op.c[5] = dp * std::cos(2 * kPi * e.coupler_angle);
          op.c[6] = dp * std::sin(2 * kPi * e.coupler_angle);
          if (both) { op.c[5] /= 2; op.c[6] /= 2; }
          op.c[7] = 2 * kPi * e.coupler_phase;
        }
        prog.push_back(op);
        emit_ekicks(0.5 * e.l);
        emit_kicks(prog, ie, ab, 0.5, OP_KICK);
        emit_offset(prog, ie, e, rev, false, true, true, cdir);
        break;
      }
      default:
        return fail("element key not supported");
    }

    // Element end (orbit_too_large, exit aperture): folded into the element's
    // last op when that op leaves c[END_C ..] free, else an OP_ELE_END.  A
    // pure-translation offset_particle unset before it folds into the op
    // before that (F_SHIFT_OUT).
    bool exit_ap = cfg.aperture_limit_on && (ap_set || ps_set) && at_end(e.aperture_at, false);
    int end_flags = F_END | (exit_ap ? (F_END_APERTURE | ((ap_flags & F_APERTURE_ELLIPTICAL) ? F_END_ELLIPTICAL : 0)) : 0);
    size_t first = (size_t)ele_op_begin.back();
    if (prog.size() >= first + 2) {
      DevOp& u = prog.back();
      DevOp& body = prog[prog.size() - 2];
      if (u.type == OP_OFFSET_UNSET && !(u.flags & F_OFFSET_PITCH) && u.c[2] == 0.0 && used_consts(body) <= SHIFT_C) {
        body.flags |= F_SHIFT_OUT;
        body.c[SHIFT_C] = u.c[0];
        body.c[SHIFT_C + 1] = u.c[1];
        prog.pop_back();
      }
    }
    if (!emit_rad(true)) return false;
    if (exit_ap && ps_set) {  // an element end of its own: phase-space limits at c[0..7], aperture at c[8..13]
      DevOp op = make_op(OP_ELE_END, ie);
      op.flags = end_flags | F_END_PS | ps_flags;
      for (int i = 0; i < 8; i++) op.c[i] = ps_c[i];
      ap_consts(op, 8);
      prog.push_back(op);
      continue;
    }
    DevOp* last = (prog.size() > first) ? &prog.back() : nullptr;
    if (!last || used_consts(*last) > END_C) {
      prog.push_back(make_op(OP_ELE_END, ie));
      last = &prog.back();
    }
    last->flags |= end_flags;
    if (exit_ap) {
      ap_consts(*last, END_C);
      const double* c = last->c + END_C;
      double mal = cfg.max_aperture_limit * (1.0 - 1e-12);
      bool inbox, centered = false;
      if (ap_flags & F_APERTURE_ELLIPTICAL) {  // |x - cx| < wx on the screen
        inbox = std::fabs(c[0]) + c[4] < mal && std::fabs(c[1]) + c[5] < mal;
        centered = c[0] == 0.0 && c[1] == 0.0;
      } else {  // x1 <= x <= x2, y1 <= y <= y2 on the screen when all four limits are set
        inbox = c[0] != 0.0 && c[1] != 0.0 && c[2] != 0.0 && c[3] != 0.0 && std::fabs(c[0]) < mal &&
                std::fabs(c[1]) < mal && std::fabs(c[2]) < mal && std::fabs(c[3]) < mal;
      }
      last->flags |= (inbox ? F_END_INBOX : 0) | (centered ? F_END_CENTERED : 0);
    }
  }
  ele_op_begin.push_back((int32_t)prog.size());
  (void)kCLight;
  (void)kPi;
  return true;
}

}  // namespace bmx
