/*
 * bmx.h -- C ABI of libbmx, an MI355X (gfx950) multi-particle tracker for
 * Bmad's bmad_standard element-by-element tracking.
 *
 * Everything that crosses this boundary is plain C: fixed-layout structs,
 * pointers and sizes.  No C++ or torch types appear here.  The Fortran side
 * binds these entry points with ISO_C_BINDING (bmad-ecosystem_amd/fortran).
 *
 * Coordinates are Bmad's canonical 6-D phase space (x, px, y, py, z, pz),
 * pz = dp/p0, stored structure-of-arrays.  Particle loss is reported only
 * through state[] with Bmad's codes (bmad/modules/bmad_struct.f90:522-527);
 * a negative return code is an error (bad input, device failure), never a loss.
 *
 * Entry point  <->  reference interface it replaces
 *   bmx_track_bunch  track_bunch / track1_bunch / track1_bunch_hom
 *                    (bmad/multiparticle/beam_mod.f90:103-170, :238-386;
 *                     bmad/multiparticle/beam_utils.f90:32-78) reached through
 *                    track1_bunch_hook_ptr (bmad/modules/bmad_routine_interface.f90:3927-3936)
 *   bmx_track_grid   track_many through track_many_hook_ptr
 *                    (bmad/code/track_many.f90:50-157,
 *                     bmad/modules/bmad_routine_interface.f90:3949-3959) as used by
 *                    dynamic_aperture_scan (bmad/modules/dynamic_aperture_mod.f90:27-207)
 *   bmx_lattice_upload  replaces the per-particle multipole_ele_to_ab calls
 *                    (bmad/code/multipole_ele_to_ab.f90:31-235) by one hoisting pass
 */
#ifndef BMX_H
#define BMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMX_ABI_VERSION 4
#define BMX_N_POLE 22 /* n_pole_maxx + 1, bmad/modules/bmad_struct.f90 */

/* ---- particle state codes (bmad/modules/bmad_struct.f90:520-527) ---- */
enum {
  BMX_NOT_SET = -1,
  BMX_PRE_BORN = 0,
  BMX_ALIVE = 1,
  BMX_LOST = 2,
  BMX_LOST_NEG_X = 3,
  BMX_LOST_POS_X = 4,
  BMX_LOST_NEG_Y = 5,
  BMX_LOST_POS_Y = 6,
  BMX_LOST_Z = 7,
  BMX_LOST_PZ = 8
};

/* ---- element keys: the bmad_standard element mix this library tracks ---- */
enum {
  BMX_KEY_MARKER = 0,
  BMX_KEY_DRIFT = 1,
  BMX_KEY_QUADRUPOLE = 2,
  BMX_KEY_SBEND = 3,
  BMX_KEY_SEXTUPOLE = 4,
  BMX_KEY_OCTUPOLE = 5,
  BMX_KEY_THICK_MULTIPOLE = 6,
  BMX_KEY_MULTIPOLE = 7,    /* thin KnL/Tn multipole: a[] = KnL, b[] = Tn */
  BMX_KEY_AB_MULTIPOLE = 8, /* thin An/Bn multipole */
  BMX_KEY_RFCAVITY = 9,
  BMX_KEY_HKICKER = 10,
  BMX_KEY_VKICKER = 11,
  BMX_KEY_KICKER = 12,
  BMX_KEY_INSTRUMENT = 13,
  BMX_KEY_MONITOR = 14,
  BMX_KEY_ECOLLIMATOR = 15,
  BMX_KEY_RCOLLIMATOR = 16,
  BMX_KEY_PIPE = 17,
  BMX_N_KEY = 18
};

/* fringe_type codes (bmad/modules/bmad_struct.f90:27,167-168) */
enum {
  BMX_FRINGE_NONE = 1,
  BMX_FRINGE_SOFT_EDGE_ONLY = 2,
  BMX_FRINGE_HARD_EDGE_ONLY = 3,
  BMX_FRINGE_FULL = 4,
  BMX_FRINGE_SAD_FULL = 5,
  BMX_FRINGE_LINEAR_EDGE = 6,
  BMX_FRINGE_BASIC_BEND = 7
};

/* element-end codes for fringe_at / aperture_at / coupler_at
 * (bmad/modules/bmad_struct.f90:372) */
enum { BMX_ENTRANCE_END = 1, BMX_EXIT_END = 2, BMX_BOTH_ENDS = 3, BMX_NO_END = 4 };

/* aperture_type codes (bmad/modules/bmad_struct.f90:157) */
enum { BMX_APERTURE_RECTANGULAR = 2, BMX_APERTURE_ELLIPTICAL = 3 };

/* particle species */
enum {
  BMX_ELECTRON = 1,
  BMX_POSITRON = 2,
  BMX_PROTON = 3,
  BMX_ANTIPROTON = 4,
  BMX_MUON = 5,
  BMX_ANTIMUON = 6
};

/*
 * One lattice element after Bmad's bookkeeping (attribute_bookkeeper), in the
 * element frame: the "_tot" attributes of the reference are the plain fields
 * here.  Layout is fixed (all doubles 8-byte aligned); the Fortran shim fills
 * it from ele%value(...).  ds_step = l / num_steps.
 */
typedef struct bmx_ele_rec {
  int32_t key;              /* BMX_KEY_* */
  int32_t is_on;            /* ele%is_on */
  int32_t orientation;      /* ele%orientation, +1 or -1 */
  int32_t num_steps;        /* >= 1 */
  int32_t fringe_type;      /* BMX_FRINGE_* */
  int32_t fringe_at;        /* BMX_*_END */
  int32_t aperture_type;    /* BMX_APERTURE_* */
  int32_t aperture_at;      /* BMX_*_END */
  int32_t field_master;     /* multipoles given in field units (T m^-n) */
  int32_t scale_multipoles; /* ele%scale_multipoles (default 1) */
  int32_t has_mag_poles;    /* associated(ele%a_pole) */
  int32_t has_elec_poles;   /* associated(ele%a_pole_elec) */
  int32_t coupler_at;       /* BMX_*_END, rfcavity only */
  int32_t exact_multipoles; /* 1 = off (only supported value) */
  int32_t reserved[2];
  double l, k1, k2, k3, g, dg, e1, e2, fint, fintx, hgap, hgapx;
  double tilt, ref_tilt;
  double x_offset, y_offset, z_offset, x_pitch, y_pitch;
  double hkick, vkick, kick;
  double voltage, rf_frequency, phi0, phi0_multipass;
  double coupler_strength, coupler_angle, coupler_phase;
  double x1_limit, x2_limit, y1_limit, y2_limit;
  double p0c, e_tot, p0c_start, e_tot_start, delta_ref_time, ref_time;
  double r0_mag, r0_elec;
  /* collimator phase-space limits (ecollimator / rcollimator) */
  double px_aperture_width2, px_aperture_center, py_aperture_width2, py_aperture_center;
  double z_aperture_width2, z_aperture_center, pz_aperture_width2, pz_aperture_center;
  double a[BMX_N_POLE], b[BMX_N_POLE];           /* magnetic An/Bn (MULTIPOLE: KnL/Tn) */
  double a_elec[BMX_N_POLE], b_elec[BMX_N_POLE]; /* electric An/Bn */
} bmx_ele_rec;

/* Reference particle of the branch (lat%param%particle) and the species the
 * compiled lattice tracks (bunch%species): 0 = the reference species, else
 * the reference species or its antiparticle (rel_tracking_charge_to_mass,
 * bmad/code/rel_tracking_charge_to_mass.f90:15-35, is then -1). */
typedef struct bmx_ref {
  int32_t species;          /* BMX_ELECTRON, ... */
  int32_t tracked_species;  /* 0 or a BMX_* species of the same mass */
} bmx_ref;

/* Global switches (subset of bmad_com, bmad/modules/bmad_struct.f90:2296-2343). */
typedef struct bmx_config {
  int32_t aperture_limit_on;     /* default 1 */
  int32_t direction;             /* bmx_lattice_check only: 0 or +1 forward, -1 backward tracking
                                    (bmx_lattice_upload ignores it and compiles both directions) */
  double max_aperture_limit;     /* default 1e3 m */
  int32_t absolute_time_tracking; /* bmad_com%absolute_time_tracking: rf phases from the particle
                                     time t (particle_rf_time.f90:72-74) instead of -z/(beta c);
                                     backward tracking of rf cavities needs it, as in the reference
                                     harness (tracking_method_test.f90:170-172) */
  int32_t reserved;
} bmx_config;

/*
 * Particle store, structure-of-arrays.  x..pz, beta and state are required;
 * ix_lost, turn_lost and t may be NULL.  beta is the particle velocity / c and
 * is carried exactly as coord_struct%beta (energy kicks update it).  t is the
 * particle time coord_struct%t [s] in the lattice's reference-time frame
 * (ref_time of the elements); bmx_track_bunch reads it (absolute-time rf
 * phases) and writes it back, for a lost particle at the far edge of the
 * element where it was lost.  It is carried as t - t_ref(s) + z/(beta c),
 * which tracking conserves, so its precision is that of the double given.
 * on_device = 1: pointers are device (HBM) addresses; 0: host memory.
 */
typedef struct bmx_particles {
  double *x, *px, *y, *py, *z, *pz;
  double *beta;
  int32_t *state;
  int32_t *ix_lost;    /* element index (0-based) where the particle was lost, -1 if alive */
  int32_t *turn_lost;  /* turn index where the particle was lost, -1 if alive */
  double *t;           /* particle time [s] (coord_struct%t), may be NULL */
  int64_t n;
  int32_t on_device;
  int32_t reserved;
} bmx_particles;

/* Radiation map of one half of an element (rad_map_struct,
 * bmad/modules/bmad_struct.f90:974-980), as radiation_map_setup leaves it in
 * ele%rad_map%rm0 (entrance half) and %rm1 (exit half)
 * (bmad/modules/radiation_mod.f90:160-232): the damping correction damp_dmat
 * acting on (vec - ref_orb), the 0th-order damping vector xfer_damp_vec and
 * the Cholesky factor stoc_mat of the stochastic kick variance (lower
 * triangular).  Matrices are row-major [row][col] over (x, px, y, py, z, pz):
 * damp_dmat[r * 6 + c] = damp_dmat(r+1, c+1) of the reference.  xfer_damp_mat
 * is not read by tracking and is not carried.  An element with no radiation
 * has all-zero maps. */
typedef struct bmx_rad_map {
  double ref_orb[6];
  double damp_dmat[36];
  double xfer_damp_vec[6];
  double stoc_mat[36];
} bmx_rad_map;

/* Radiation switches (bmad_com%radiation_damping_on, %radiation_fluctuations_on,
 * %radiation_zero_average, %synch_rad_scale; bmad/modules/bmad_struct.f90:2312-2330)
 * and the seed of the fluctuation stream.  The Gaussian deviates come from
 * Philox4x32-10 keyed by the seed, with the counter (particle index, element
 * index and edge, turn of the call, call number and pair): the same particle
 * sees the same kicks whatever the launch shape, and every tracking call on a
 * lattice handle draws fresh ones (the call number advances per call).  The
 * reference draws them from its own generator (ran_gauss), so fluctuating
 * tracks agree with it in distribution, not particle by particle. */
typedef struct bmx_radiation {
  int32_t damping_on;
  int32_t fluctuations_on;
  int32_t zero_average;   /* 1: leave out xfer_damp_vec (no closed-orbit sawtooth) */
  int32_t reserved;
  double synch_rad_scale; /* default 1 */
  uint64_t seed;
} bmx_radiation;

/* Per-call statistics returned by the tracking entry points. */
typedef struct bmx_stats {
  int64_t n_live;           /* particles still alive after the call */
  int64_t n_lost;           /* particles lost during the call */
  double kernel_ms;         /* device time of the tracking kernel(s), HIP events */
  double pet;               /* live particle-element-turns processed */
} bmx_stats;

/* Batched dynamic-aperture grid: n_var lattice variants, each with the same
 * polar grid of initial conditions n_angle x n_amp x n_pz.  Initial condition
 * ic = (i_pz * n_angle + i_angle) * n_amp + i_amp is the reference orbit of
 * (variant, pz slice) plus the offset x = a cos(th), y = a amp_ratio_y sin(th)
 * with a = amp_max_x (i_amp + 1) / n_amp (dynamic_aperture_point,
 * bmad/modules/dynamic_aperture_mod.f90:333-336).  Without a reference orbit
 * it is (0, 0, 0, 0, 0, pz_i), pz_i evenly spaced over [pz_min, pz_max].
 * Tracking runs in blocks of turns_per_block turns; survivors are compacted
 * between blocks (turns_per_block <= 0: 32). */
typedef struct bmx_grid {
  int32_t n_angle, n_amp, n_pz, turns_per_block;
  double amp_max_x;         /* largest x amplitude [m]; y amplitude scales by amp_ratio_y */
  double amp_ratio_y;
  double pz_min, pz_max;
  double angle_min, angle_max; /* polar angle range [rad], default (0, pi) */
} bmx_grid;

/* Turn-block schedule of the survival scans (bmx_track_grid, bmx_track_ics):
 * turns_per_block turns per kernel launch, survivors compacted between
 * blocks; graph_depth > 1 captures that many blocks (rounded up to even) into
 * a hipGraph replayed between host synchronisations, 1 launches block by
 * block.  probe_pet_per_s: the rate bmx_tune_grid measured for it. */
typedef struct bmx_schedule {
  int32_t turns_per_block;
  int32_t graph_depth;
  double probe_pet_per_s;
} bmx_schedule;

typedef struct bmx_ctx bmx_ctx;
typedef struct bmx_lattice bmx_lattice;

/* Library version (BMX_ABI_VERSION). */
int32_t bmx_abi_version(void);

/* sizeof of the ABI structs, for binding-side layout checks:
 * which = 0 bmx_ele_rec, 1 bmx_particles, 2 bmx_stats, 3 bmx_grid, 4 bmx_config, 5 bmx_schedule,
 * 6 bmx_rad_map, 7 bmx_radiation. */
int32_t bmx_sizeof(int32_t which);

/* Create a context on HIP device `device`.  flags: reserved, pass 0. */
int32_t bmx_ctx_create(bmx_ctx **out, int32_t device, uint32_t flags);
void bmx_ctx_destroy(bmx_ctx *ctx);
const char *bmx_last_error(const bmx_ctx *ctx);

/* Compile and upload a lattice.  eles holds n_variants consecutive element
 * lists; variant v is eles[variant_offsets[v] .. variant_offsets[v+1]).
 * variant_offsets has n_variants+1 entries (NULL with n_variants = 1 means the
 * whole array).  Replaces the per-particle multipole_ele_to_ab bookkeeping. */
int32_t bmx_lattice_upload(bmx_ctx *ctx, const bmx_ele_rec *eles, int32_t n_ele_total,
                           const int32_t *variant_offsets, int32_t n_variants,
                           const bmx_ref *ref, const bmx_config *cfg, bmx_lattice **out);
void bmx_lattice_destroy(bmx_lattice *lat);

/* bmx_lattice_upload with radiation damping and/or fluctuations: maps holds
 * two bmx_rad_map per element of eles (maps[2 i] = rm0, applied at the
 * element's entrance; maps[2 i + 1] = rm1, at its exit), the maps
 * radiation_map_setup computed for them (a Bmad caller passes
 * ele%rad_map%rm0 / %rm1).  Each element gets track1's radiation steps
 * (bmad/code/track1.f90:198-201, :316-319; track1_radiation,
 * bmad/modules/radiation_mod.f90:62-141): after the entrance aperture,
 * vec += synch_rad_scale (damp_dmat (vec - ref_orb) + xfer_damp_vec) and,
 * with fluctuations, vec += synch_rad_scale stoc_mat N(0, 1)^6, then a
 * particle with pz < -1 is lost (lost_pz); the same at the exit after
 * orbit_too_large and before the exit aperture.  beta is left as it was
 * (the reference does not update it there).  Elements that track1_radiation
 * skips (zero length, drift, pipe, marker, multipole, ab_multipole, off)
 * and all-zero maps add no step.  Radiation is tracked forward (direction
 * -1 refuses the lattice) in relative or absolute time.  Fluctuations are
 * accepted by bmx_track_bunch(_noalloc) and bmx_track_multi only; damping by
 * every tracking entry point.  rad NULL or both switches off: the same as
 * bmx_lattice_upload. */
int32_t bmx_lattice_upload_rad(bmx_ctx *ctx, const bmx_ele_rec *eles, int32_t n_ele_total,
                               const int32_t *variant_offsets, int32_t n_variants, const bmx_rad_map *maps,
                               const bmx_radiation *rad, const bmx_ref *ref, const bmx_config *cfg,
                               bmx_lattice **out);

/* The fluctuation stream of a lattice handle: seed, and the call number the
 * next tracking call draws with (it advances by one per call; 0 after
 * upload).  No reference counterpart (ran_seed_put plays this role for the
 * reference's generator). */
int32_t bmx_lattice_set_rad_stream(bmx_lattice *lat, uint64_t seed, uint64_t next_call);

/* Strength attributes bmx_lattice_upload_scaled can scale. */
enum { BMX_ATTR_K1 = 1, BMX_ATTR_K2 = 2, BMX_ATTR_K3 = 3 };

/* n_variants variants of one element list without building them all on the
 * caller's side: variant v is eles[0 .. n_ele) with the attribute `attr`
 * (BMX_ATTR_K1, K2 or K3) of element sel[j] multiplied by
 * factors[v * n_sel + j] -- knob or error studies of magnet strengths, the
 * same lattice as bmx_lattice_upload of the scaled copies, record for record.
 * The variants compile on several host threads.  No reference counterpart
 * (the reference sets each value with set_flags_for_changed_attribute and
 * lattice_bookkeeper, lattice by lattice). */
int32_t bmx_lattice_upload_scaled(bmx_ctx *ctx, const bmx_ele_rec *eles, int32_t n_ele, int32_t n_variants,
                                  int32_t attr, const int32_t *sel, int32_t n_sel, const double *factors,
                                  const bmx_ref *ref, const bmx_config *cfg, bmx_lattice **out);

/* Host-only check of one element list against the device path, with no
 * device, context or allocation on it: 0 when bmx_lattice_upload would
 * accept it (*n_ops: the ops it compiles to), else -1 with the reason (the
 * element and the unsupported feature) in msg.  n_ops and msg may be NULL.
 * No reference counterpart: it answers ahead of time what the upload would
 * reject. */
int32_t bmx_lattice_check(const bmx_ele_rec *eles, int32_t n_ele, const bmx_ref *ref, const bmx_config *cfg,
                          int64_t *n_ops, char *msg, int32_t msg_len);

/* The lattice's runtime-compiled cell kernels: *cell_len = ops in its
 * repeating cell and *n_cells = runs of that cell the kernels track straight
 * through (both 0 when the lattice runs on the built-in op interpreter alone:
 * no cell covers half of its ops, BMX_JIT=0 in the environment, or the
 * runtime compiler failed).  Results are the same either way.  No reference
 * counterpart. */
int32_t bmx_lattice_cells(const bmx_lattice *lat, int32_t *cell_len, int64_t *n_cells);

/* Track a bunch through elements ix_start+1 .. ix_end of variant 0 with
 * track_bunch wrap-around semantics (ix_end <= ix_start wraps through the
 * end of the lattice), repeated n_turns times.  direction = -1 tracks
 * backward through elements ix_start down to ix_end+1 (wrapping when
 * ix_start <= ix_end), beam_mod.f90:131-147, with the lattice's backward
 * program, compiled at upload (a lattice the backward path refuses returns
 * the reason here).  Equivalent to n_turns calls of track_bunch(lat, bunch,
 * ele_start, ele_end, direction) (bmad/multiparticle/beam_mod.f90:103-170). */
int32_t bmx_track_bunch(bmx_ctx *ctx, const bmx_lattice *lat, int32_t ix_start, int32_t ix_end,
                        int32_t n_turns, int32_t direction, bmx_particles *p, bmx_stats *out);

/* bmx_track_bunch that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight turn or serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as bmx_track_bunch.  The loss counter
 * and, for host-resident arrays, their staging copies live in a static
 * device array of 256 B + 65536 x 76 B (about 5.0 MB) that each GPU holds
 * until the process ends: host-resident bunches of more than
 * BMX_NOALLOC_MAX_PARTICLES particles are refused with an error, device-
 * resident arrays are tracked in place at any size.  Calls that overlap on
 * one GPU (threads or contexts) share that array and race: serialise them. */
#define BMX_NOALLOC_MAX_PARTICLES 65536
int32_t bmx_track_bunch_noalloc(bmx_ctx *ctx, const bmx_lattice *lat, int32_t ix_start, int32_t ix_end,
                                int32_t n_turns, int32_t direction, bmx_particles *p, bmx_stats *out);

/* Track particles of several lattice variants in one launch: particles
 * [part_offsets[v], part_offsets[v+1]) go through variant v for n_turns whole
 * turns; part_offsets has n_variants + 1 entries from 0 to p->n.  Replaces one
 * track_many / track_bunch call per lattice (bmad/code/track_many.f90,
 * bmad/multiparticle/beam_mod.f90:40-170) when many lattices are tracked. */
int32_t bmx_track_multi(bmx_ctx *ctx, const bmx_lattice *lat, const int64_t *part_offsets, int32_t n_turns,
                        bmx_particles *p, bmx_stats *out);

/* Track every variant's grid of initial conditions for n_turns full turns.
 * ref_orbit: NULL, or n_variants x n_pz x 6 reference orbits (closed orbit
 * plus the pz offset, as dynamic_aperture_scan forms ap_scan%ref_orb,
 * bmad/modules/dynamic_aperture_mod.f90:130-140).  survival_turn[v*n_ic + i]
 * = turns survived (n_turns if alive at the end); da_area[v] = area of the
 * polygon through the origin and the per-angle boundary offsets of the middle
 * pz slice [m^2] (may be NULL).  Host output arrays. */
int32_t bmx_track_grid(bmx_ctx *ctx, const bmx_lattice *lat, const bmx_grid *grid, const double *ref_orbit,
                       int32_t n_turns, int32_t *survival_turn, double *da_area, bmx_stats *out);

/* Survival turns of explicit initial conditions (turns_per_block <= 0: the
 * context's schedule): ics is 6 x n host SoA
 * (x[n], px[n], y[n], py[n], z[n], pz[n]) with n = ic_offsets[n_variants];
 * ICs [ic_offsets[v], ic_offsets[v+1]) are tracked through variant v.  The
 * per-IC survival loop of dynamic_aperture_point
 * (bmad/modules/dynamic_aperture_mod.f90:339-349), batched. */
int32_t bmx_track_ics(bmx_ctx *ctx, const bmx_lattice *lat, const double *ics, const int64_t *ic_offsets,
                      int32_t n_turns, int32_t turns_per_block, int32_t *survival_turn, bmx_stats *out);

/* The context's turn-block schedule (default 32 turns per block, no graph);
 * a scan call passing turns_per_block <= 0 uses it. */
int32_t bmx_ctx_set_schedule(bmx_ctx *ctx, const bmx_schedule *s);
int32_t bmx_ctx_get_schedule(const bmx_ctx *ctx, bmx_schedule *s);

/* Autotuner of the turn-block schedule (no counterpart in the reference):
 * times every candidate (turns_per_block in {4, 8, 16, 32, 64, 128} x
 * graph_depth in {1, 8}) on a sparse sample of this grid (every 4th angle and
 * amplitude, all variants and pz slices) for min(n_turns, 64) turns, keeps
 * the highest live particle-element-turns per second of wall time, stores it
 * as the context's schedule and returns it in *best. */
int32_t bmx_tune_grid(bmx_ctx *ctx, const bmx_lattice *lat, const bmx_grid *grid, const double *ref_orbit,
                      int32_t n_turns, bmx_schedule *best);

/* Sums over the live particles of a bunch, the moments of ltt_calc_bunch_data
 * (bsim/modules/lt_tracking_mod.f90:2391-2402): moments[0] = n_live,
 * moments[1..6] = sum vec_i, moments[7..27] = sum vec_i vec_j for i <= j
 * (upper triangle, row by row).  Deterministic reduction order. */
int32_t bmx_bunch_moments(bmx_ctx *ctx, const bmx_particles *p, double *moments);

/* Closed orbit of every variant (closed_orbit_calc, bmad/code/closed_orbit_calc.f90:81):
 * orbit is n_variants x 6, the guess on entry and the closed orbit on return.
 * n_dim = 4 closes x, px, y, py at the guess's pz; n_dim = 6 closes all six
 * (rf on), with z kept within half an rf wavelength (of the lowest frequency
 * among the cavities that are on) of zero and, as closed_orbit_calc does
 * (:491-520), moved off a longitudinally unstable fixed point to the stable
 * one.  status[v]: 0 converged, 1 not converged, 2 a probe was lost,
 * 3 singular one-turn matrix. */
int32_t bmx_closed_orbit(bmx_ctx *ctx, const bmx_lattice *lat, int32_t n_dim, double *orbit, int32_t *status,
                         bmx_stats *out);

/* One-turn transfer matrix of every variant by tracking, the matrix step of
 * twiss_from_tracking (bmad/code/twiss_from_tracking.f90:51-70): central
 * differences of one turn about orbit (n_variants x 6) with steps d_orb[6]
 * (NULL: 1e-5 each, bmad_com%d_orb): mat6 is n_variants x 36, row-major,
 * mat6[v][i][j] = d out_i / d in_j.  status[v]: 0, or 2 when a probe was
 * lost (its matrix is zeros).  Cavities act as uploaded; the reference turns
 * them off for its Twiss (upload the lattice without rf for that). */
int32_t bmx_one_turn_matrix(bmx_ctx *ctx, const bmx_lattice *lat, const double *orbit, const double *d_orb,
                            double *mat6, int32_t *status, bmx_stats *out);

/* Exact transfer matrices: for each of n orbits (vec_in, n x 6, host), the
 * 6x6 Jacobian mat6[i][r][q] = d out_r / d in_q (row-major, n x 36) of
 * tracking it forward through elements ix_start+1 .. ix_end of variant 0 for
 * n_turns passes (the element range of bmx_track_bunch; ix_start == ix_end:
 * whole turns), and the tracked orbit (vec_out, n x 6).  Replaces make_mat6
 * (bmad/code/make_mat6.f90:23) with the make_matrix branches of the
 * track_a_* routines (e.g. bmad/low_level/track_a_drift.f90:94-114), chained
 * over the range as transfer_matrix_calc does; the derivatives are those of
 * libbmx's own tracking map, carried by forward-mode differentiation, so one
 * element or a thousand turns are exact to rounding.  beta at entry follows
 * init_coord (the reference beta at pz = 0).  state[i]: BMX_ALIVE, or the
 * loss state (its matrix is zeros, vec_out the coordinates at the loss).
 * With radiation damping (bmx_lattice_upload_rad) the kicks move the orbit
 * and leave the derivatives alone, as track1 applies track1_radiation to
 * the orbit and not to ele%mat6 (bmad/code/track1.f90:198-201, :316-319).
 * Relative-time rf only. */
int32_t bmx_transfer_matrix(bmx_ctx *ctx, const bmx_lattice *lat, int32_t ix_start, int32_t ix_end,
                            int32_t n_turns, int64_t n, const double *vec_in, double *vec_out, double *mat6,
                            int32_t *state);

/* ------------------------------------------------------------------------
 * Dynamic-aperture surrogate kernels (no counterpart in the reference; they
 * serve the learned screening of lattice variants).  Device pointers; work
 * is queued on the context's stream (bmx_ctx_sync waits for it).
 * ------------------------------------------------------------------------ */

/* C[M x N] = op(A) op(B), BF16 inputs (op = transpose when trans_*), FP32
 * matrix-core accumulation.  Row-major with leading dimensions lda, ldb, ldc.
 * epilogue: 0 plain; 1 + bias[n]; 2 pre = acc + bias[n] -> C (FP32, may be
 * NULL), C_bf16 = SiLU(pre); 3 acc * SiLU'(aux[m, n]).  C (FP32) and/or
 * C_bf16 receive the result. */
int32_t bmx_gemm_bf16(bmx_ctx *ctx, int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K,
                      const void *A, int32_t lda, const void *B, int32_t ldb, float *C, void *C_bf16, int32_t ldc,
                      const float *bias, int32_t epilogue, const float *aux);

/* out[n] = sum_m X[m * ld + n] (bias gradients). */
int32_t bmx_colsum_f32(bmx_ctx *ctx, int32_t M, int32_t N, const float *X, int32_t ld, float *out);

/* One Adam step on FP32 master weights w (and their BF16 copy, may be NULL). */
int32_t bmx_adam_step(bmx_ctx *ctx, int64_t n, float *w, void *w_bf16, const float *g, float *m, float *v, float lr,
                      float beta1, float beta2, float eps, int32_t step);

/* y = (bf16) x, element-wise. */
int32_t bmx_f32_to_bf16(bmx_ctx *ctx, int64_t n, const float *x, void *y);

/* Mean-squared-error loss over n predictions: *loss = sum (pred - target)^2 / n,
 * d = 2 (pred - target) / n in FP32 and (d_bf16) BF16. */
int32_t bmx_mse_grad(bmx_ctx *ctx, int32_t n, const float *pred, const float *target, float *d, void *d_bf16,
                     float *loss);

/* Surrogate inputs of a DA scan in BF16, row l * n_ic + i = (knobs[l, :] /
 * knob_range, geo[i, :]) for n_lat lattices x n_ic grid points (row width
 * n_knob + n_geo).  Device pointers. */
int32_t bmx_da_features(bmx_ctx *ctx, int32_t n_lat, int32_t n_ic, int32_t n_knob, int32_t n_geo,
                        const float *knobs, float knob_range, const float *geo, void *X_bf16);

/* Boundary index of each of n_rays predicted survival profiles pred[ray, n_amp]
 * (device pointers): the k in 0..n_amp minimising
 * sum_{j<k} (p_j - alive_level)^2 + sum_{j>=k} (p_j - mean_{j>=k} p)^2, ties to
 * the larger k (n_amp: the whole ray survives).  n_amp <= 2048.  No reference
 * counterpart (the DA surrogate's screening step). */
int32_t bmx_step_fit(bmx_ctx *ctx, int32_t n_rays, int32_t n_amp, const float *pred, float alive_level,
                     int32_t *k_out);

/* The whole forward pass of a surrogate network dims[0] -> 256 -> ... -> 256
 * -> 1 (dims[0] <= 64, SiLU after each hidden layer; n_dims entries, at
 * least one hidden layer) in one kernel, activations kept on chip: y[i] for
 * the n BF16 input rows X (n x dims[0]).  w: FP32 parameters, w_bf16 their
 * BF16 copy (16-byte aligned), both laid out per layer as W [out][in] then
 * b [out].  Equal bit for bit to the layer-by-layer bmx_gemm_bf16 pass. */
int32_t bmx_mlp_forward(bmx_ctx *ctx, int32_t n_dims, const int32_t *dims, const float *w, const void *w_bf16,
                        const void *X_bf16, int64_t n, float *y);

/* Surrogate screening of a DA scan: for each of n_lat knob vectors (knobs,
 * n_lat x n_knob) and each of the n_ic grid ICs (geo, n_ic x n_geo, ordered
 * ray by ray with n_amp amplitudes per ray), the network's prediction
 * (inputs as bmx_da_features forms them, forward pass as bmx_mlp_forward),
 * then each ray's step fit (bmx_step_fit): kb, n_lat x (n_ic / n_amp)
 * predicted boundary indices.  n_knob + n_geo = dims[0]. */
int32_t bmx_da_screen(bmx_ctx *ctx, int32_t n_dims, const int32_t *dims, const float *w, const void *w_bf16,
                      int32_t n_lat, int32_t n_knob, const float *knobs, float knob_range, int32_t n_ic,
                      int32_t n_geo, const float *geo, int32_t n_amp, float alive_level, int32_t *kb);

/* A surrogate network owned by the library, for callers without the Python
 * wrapper (a Fortran DA driver, a C optimiser).  Layer widths dims[0 ..
 * n_dims-1] (dims[n_dims-1] = 1), SiLU after every hidden layer; params
 * (host, FP32, *n_params of them) per layer W [out][in] row-major then
 * b [out].  Training is minibatch Adam (beta 0.9, 0.999, eps 1e-8) on the
 * mean-squared error with a cosine learning-rate decay from lr to lr_final
 * over all steps, each epoch one pass over a shuffle (seeded) of the n rows
 * of X (device, FP32, n x dims[0]) and Y (device, FP32, n); the remainder of
 * n / batch is left out.  Each step is captured once into a hipGraph and
 * replayed (flags & 1: launch every step eagerly instead; same results).
 * epoch_loss (host, may be NULL): per epoch the mean of the losses sampled
 * every 16 steps.  bmx_surrogate_infer: y (device, n) for BF16 input rows
 * (one kernel when the network is dims[0] <= 64 -> 256 ... -> 1, see
 * bmx_mlp_forward); bmx_surrogate_screen: bmx_da_screen with this network.
 * A network belongs to its context: calls on it must not overlap, and it
 * is used only while the context lives (destroying it needs no context). */
typedef struct bmx_net bmx_net;
int32_t bmx_surrogate_create(bmx_ctx *ctx, int32_t n_dims, const int32_t *dims, const float *params,
                             int64_t *n_params, bmx_net **out);
void bmx_surrogate_destroy(bmx_net *net);
int32_t bmx_surrogate_set_params(bmx_net *net, const float *params);
int32_t bmx_surrogate_get_params(const bmx_net *net, float *params);
int32_t bmx_surrogate_train(bmx_net *net, const float *X, const float *Y, int64_t n, int32_t epochs, int32_t batch,
                            float lr, float lr_final, uint64_t seed, int32_t flags, float *epoch_loss);
int32_t bmx_surrogate_infer(bmx_net *net, const void *X_bf16, int64_t n, float *y);
int32_t bmx_surrogate_screen(bmx_net *net, int32_t n_lat, int32_t n_knob, const float *knobs, float knob_range,
                             int32_t n_ic, int32_t n_geo, const float *geo, int32_t n_amp, float alive_level,
                             int32_t *kb);

/* Wait for the work queued on the context's stream. */
int32_t bmx_ctx_sync(bmx_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* BMX_H */
