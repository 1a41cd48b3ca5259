// bmx_internal.h -- device program format shared by the lattice compiler
// (bmx_lattice.cpp) and the tracking kernels (bmx_track.hip).
//
// A lattice is compiled once into a flat program of "ops".  Each Bmad element
// becomes a short op sequence that reproduces track1 -> track1_bmad ->
// track_a_* for that element (bmad/code/track1.f90:43-342,
// bmad/low_level/track1_bmad.f90:22-283), with every particle-independent
// quantity (multipole_ele_to_ab results, tilt/pitch rotations, edge-kick
// coefficients, RF constants) folded into the op's constants.  The kernels
// walk the op list with a wave-uniform index, so the ops are read through
// the scalar cache and the dispatch is a scalar branch.
#pragma once
#include <stdint.h>

// Particles per lane of the tracking kernels (bmx_kernels.h).
#ifndef BMX_NP
#define BMX_NP 1
#endif

// Lanes per workgroup of the survival-scan kernel (bmx_grid_block_kernel),
// also its LDS chunk of ops.  The host launches it with this many lanes and
// passes the value (and BMX_NP) to the runtime-compiled kernels.  On scans
// whose variants keep a few hundred live particles each, 64- and 128-lane
// workgroups ran 2-7 % slower than 256 (more op staging per particle; the
// partly filled last workgroup of a variant was not what limited them).
#ifndef BMX_GRID_BLOCK
#define BMX_GRID_BLOCK 256
#endif

namespace bmx {

// FP64 for values that are never a particle coordinate (op constants, series
// coefficients).  The particle arithmetic of bmx_ops.h is written with
// `double`, which the transfer-matrix build redefines as a dual number; this
// name stays plain FP64 there.
typedef double bmx_f64;

enum OpType : int32_t {
  OP_DRIFT = 1,          // exact drift, c[0] = L
  OP_KICK = 2,           // thin ab multipole kick, complex Horner over n = 0..nmax
  OP_KICK_REF0 = 3,      // n = 0 kt kick of a multipole element with ref_orb_offset
  OP_QUAD = 4,           // linear quad body step with quadratic z and low-energy z
  OP_BEND = 5,           // exact sector-bend body step (g != 0, k1 = 0)
  OP_BEND_K1 = 6,        // sbend body step with k1 (sbend_body_with_k1_map)
  OP_BEND_KICK = 7,      // multipole kick inside a bend, scaled by (1 + g x)
  OP_HWANG = 8,          // Hwang bend edge kick (basic_bend / hard_edge_only)
  OP_LINEAR_EDGE = 9,    // linear bend edge kick
  OP_HARD_MULT_EDGE = 10,// hard multipole edge kick of quadrupole / sextupole, n = 1 / 2
  OP_OFFSET_SET = 11,    // lab -> body frame (offsets, pitches, tilt)
  OP_OFFSET_UNSET = 12,  // body -> lab frame
  OP_TOO_LARGE = 13,     // orbit_too_large check inside an element (early return)
  OP_RFCAVITY = 14,      // standing-wave cavity body (n_slice energy kicks and drifts)
  OP_APERTURE_IN = 15,   // aperture check at the entrance edge
  OP_ELE_END = 16,       // no body: only the element-end work of F_END
  OP_TILT = 17,          // tilt_coords rotation, c[0] = cos, c[1] = sin
  OP_KD_SLICES = 18,     // whole kick-drift body of a thick multipole: half kick, n x (drift, kick), last half
  OP_SBEND_FUSED = 19,   // hard edge (e = 0, no fringe integral), orbit_too_large, exact body, orbit_too_large, hard edge
  OP_EKICK = 20,         // thin electric ab multipole kick (1/beta) with its pz, beta, z update
  OP_EFRINGE = 21,       // electric longitudinal fringe: energy kick of the multipole potential
  OP_BEND_EKICK = 22,    // electric multipole kick inside a bend, scaled by (1 + g x) / (ps beta)
  OP_DRIFT_BACK = 23,    // exact drift tracked backward (direction = -1), c[0] = L, c[1] = beta_ref
  OP_ZBETA = 24,         // z -= c[0] beta (the quadrupole's backward 2 c beta delta_ref_time per step)
  OP_RAD = 25,           // track1_radiation at one edge: c[0] = the map's index in LatticeConsts::rad
};

// op flags
enum : int32_t {
  F_APERTURE_ELLIPTICAL = 1,
  F_HWANG_ENTER = 2,      // OP_HWANG at the first track edge
  F_OFFSET_PITCH = 4,     // OP_OFFSET_*: general rotation, else pure translation
  F_OFFSET_AFFINE = 8,    // OP_OFFSET_*: bend frame change, r = W (x, y, 0) + t, t at c[0..2]
  F_OFFSET_MOVE = 32,     // OP_OFFSET_*: has a translation/pitch part
  F_EDGE_EXIT = 64,       // OP_SBEND_FUSED: exit edge present (F_HWANG_ENTER: entrance edge present)
  F_COUPLER_ENTER = 1 << 8,  // OP_RFCAVITY: coupler kick at the entrance (c[5..7])
  F_COUPLER_EXIT = 1 << 9,   // OP_RFCAVITY: coupler kick at the exit
  F_RF_ABS = 1 << 12,        // OP_RFCAVITY: absolute-time phase (c[8] time offset, c[9] frac(f T_lattice))
  F_RF_BACK = 1 << 13,       // OP_RFCAVITY: backward drifts between the kicks (c[10] beta_ref)
  // collimator phase-space limits (px, py, z, pz width2 / center pairs):
  // OP_APERTURE_IN with F_APERTURE_PS holds them at c[6..13]; an OP_ELE_END
  // with F_END_PS at c[0..7], its aperture constants then at c[8..13];
  // F_PS_ELLIPTICAL: (px, py) and (z, pz) are ellipses (ecollimator)
  F_PS_ELLIPTICAL = 1 << 10,
  F_APERTURE_PS = 1 << 11,
  // the element ends after this op's body: orbit_too_large, then (with
  // F_END_APERTURE) the exit aperture with its constants at c[END_C ..]
  F_END = 1 << 16,
  F_END_APERTURE = 1 << 17,
  F_END_ELLIPTICAL = 1 << 18,
  // after the body (and before the element end), for a particle still alive:
  // offset_particle unset of a pure transverse offset, x += c[SHIFT_C],
  // y += c[SHIFT_C + 1] (the element's OP_OFFSET_UNSET, folded)
  F_SHIFT_OUT = 1 << 19,
  // first op of a run of the lattice's repeating cell (runtime-compiled cell kernels)
  F_CELL = 1 << 20,
  // with F_END_APERTURE: the aperture lies inside the max_aperture_limit box
  // (its screen implies orbit_too_large's position limits), and (elliptical)
  // it is centred on the axis
  F_END_INBOX = 1 << 21,
  F_END_CENTERED = 1 << 22,
  F_END_PS = 1 << 23,
  // OP_RAD: damping (damp_dmat, xfer_damp_vec), fluctuations (stoc_mat), and
  // the exit edge (orbit_too_large first, as track1 checks it before the
  // exit half of the radiation)
  F_RAD_DAMP = 1,
  F_RAD_FLUCT = 2,
  F_RAD_EXIT = 4,
};

// One radiation map in LatticeConsts::rad, pre-scaled by synch_rad_scale:
// ref_orb[6], damp_dmat[36] (row-major), xfer_damp_vec[6] (zero with
// radiation_zero_average), the lower triangle of stoc_mat row by row (21),
// padded to 72 doubles.
constexpr int RAD_REF = 0, RAD_DAMP = 6, RAD_VEC = 42, RAD_STOC = 48, RAD_STRIDE = 72;

// aperture constants of an op carrying F_END (6 slots, see aperture())
constexpr int END_C = 6;
constexpr int SHIFT_C = 12;

struct alignas(64) DevOp {
  int32_t type;   // OpType
  int32_t ie;     // element index in its variant (0-based)
  int32_t n;      // aux integer (kick order, slice count)
  int32_t flags;  // F_*
  double c[14];   // constants, meaning per op type (see bmx_lattice.cpp)
};
static_assert(sizeof(DevOp) == 128, "DevOp must be 128 bytes");

// Per-lattice constants shared by every op (one reference particle, constant p0c).
// Everything the kernels would otherwise derive from these with FP64
// arithmetic is precomputed here: such values would be hoisted out of the
// element loop into long-lived vector registers (there is no scalar FP64 ALU),
// while kernel-argument constants stay in scalar registers.
struct LatticeConsts {
  double p0c;        // reference momentum [eV]
  double e_tot;      // reference energy [eV]
  double mc2;        // particle mass [eV]
  double beta0;      // p0c / e_tot
  double max_aperture_limit;
  double mc2sq;      // mc2^2
  double p0c_sq;     // p0c^2
  double lez_k;      // mc2 * beta0^2       (low_energy_z_correction branch test)
  double lez_thr;    // 3e-7 * e_tot
  double lez_a;      // 3/2 beta0^2
  double lez_f;      // beta0^2 (2 beta0^2 - (mc2/e_tot)^2 / 2)
  double lez_me2;    // (mc2/e_tot)^2
  double two_max_ap; // 2 * max_aperture_limit
  int32_t aperture_limit_on;
  int32_t species;
  const int32_t* cell_prv;  // device: last F_CELL op at or before each op (-1: none); cell kernels only
  // Reference time (absolute-time rf and the particle time t).  Lattice:
  // t_lat = the reference time of one pass through the lattice; rt = device
  // array of the reference times of the element boundaries (n_ele + 1:
  // rt[j] = exit of array element j - 1, rt[0] = entrance of element 0).
  // Per call (set by the entry point on its copy): tt_sign = +1 forward, -1
  // backward (the reference time runs against s backward); tt_d = the
  // reference time of one pass of the tracked range minus t_lat; ix_first =
  // the array index of the first element tracked; t_start, t_end = the
  // reference times at the start edge of the first pass and the end edge of
  // the last.  The wave-uniform time of a pass (TurnTime) is k t_lat + r with
  // k = tt_sign (turn + wrapped) and r = tt_sign turn tt_d.
  double t_lat;
  const double* rt;
  double tt_d, t_start, t_end;
  int32_t tt_sign, ix_first;
  int32_t abs_time;  // the program has absolute-time rf cavities
  int32_t n_ele;
  // radiation maps of the program's OP_RAD ops (RAD_STRIDE doubles each),
  // and the fluctuation stream: Philox key (the seed) and this call's number
  const double* rad;
  uint32_t rad_key0, rad_key1, rad_call;
  int32_t rad_fluct;  // some OP_RAD of the program carries F_RAD_FLUCT
};

// Radiation input of one variant for the lattice compiler: two maps per
// element (rm0, rm1) and the switches.  The compiler appends each radiation
// step's pre-scaled map to its table (RAD_STRIDE doubles) and points the
// OP_RAD at it.
struct RadInput {
  const bmx_rad_map* maps;
  bmx_radiation cfg;
};

// Reference time of the current pass relative to the lattice frame, as
// k t_lat + r (k whole lattice periods, kept apart so that absolute-time rf
// phases stay exact over many turns: each cavity holds frac(f t_lat)).
struct TurnTime {
  int32_t k;
  double r;
  int32_t turn;  // pass number within the call (the radiation stream's counter)
};

}  // namespace bmx
