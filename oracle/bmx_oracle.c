/*
 * bmx_oracle.c -- CPU restatement of Bmad's bmad_standard single-particle
 * tracking for the element mix of libbmx.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP tracker is
 * checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; libbmx never links or calls it.
 *
 * Every routine restates the reference routine cited in its comment,
 * statement by statement, in the reference's operation order, written anew in
 * C.  Paths are relative to the bmad-ecosystem root.  Supported: forward and
 * backward direction (time_dir = +1), both element orientations, the
 * elements enumerated in include/bmx.h.  Time t is carried in long double
 * (the reference uses quad precision, bmad/modules/bmad_struct.f90:549).
 *
 * Pinning (tests/test_oracle_golden.py, tests/test_oracle_invariants.py):
 * per-element maps by the tracking_method_test Bmad_Standard rows (absolute
 * time and backward rf included), multipole_ele_to_ab by multipole_test, the
 * Hwang edge by fringe_test, apertures by aperture_test; the linear edge,
 * the coupler angle, reversed bends and every branch switch by closed forms,
 * symmetries and high-precision evaluation of the other branch.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_real.h"
#include "../include/bmx.h"

#ifdef ORACLE_COUNT_FLOPS
extern "C" {
long long oracle_flop_counter = 0;
}
#define ORACLE_API extern "C"
#elif defined(ORACLE_DUAL)
#define ORACLE_API extern "C"
#else
#define ORACLE_API
#endif

/* Time t is carried for the golden-vector checks only; libbmx does not track
 * it, so its arithmetic is left out of the FLOP count. */
#ifdef ORACLE_COUNT_FLOPS
#define NOCOUNT_BEGIN long long nocount_saved_ = oracle_flop_counter
#define NOCOUNT_END oracle_flop_counter = nocount_saved_
#else
#define NOCOUNT_BEGIN ((void)0)
#define NOCOUNT_END ((void)0)
#endif

/* ------------------------------------------------------------------------ */
/* constants: sim_utils/interfaces/physical_constants.f90:12-37             */
/* ------------------------------------------------------------------------ */
static const double O_PI = 3.141592653589793238462643383279;
static const double O_TWOPI = 2 * 3.141592653589793238462643383279;
static const double O_C_LIGHT = 2.99792458e8;
static const double O_M_ELECTRON = 0.51099895069e6;
static const double O_M_PROTON = 0.93827208943e9;
static const double O_M_MUON = 105.6583755e6;

enum { O_MAGNETIC = 1, O_ELECTRIC = 2 };
enum { O_UPSTREAM = 1, O_DOWNSTREAM = 2, O_INSIDE = 3 };
enum { O_FIRST_EDGE = 11, O_SECOND_EDGE = 12 };

/* Particle record: the coord_struct fields this path reads or writes
 * (bmad/modules/bmad_struct.f90:546-575). */
/* The particle time is real(qp) in the reference (bmad_struct.f90:549); the
 * plain builds carry it in long double, the counting and dual-number builds
 * in their scalar type. */
#if defined(ORACLE_COUNT_FLOPS) || defined(ORACLE_DUAL)
typedef real treal;
#else
typedef long double treal;
#endif

typedef struct o_coord {
  real vec[6];
  treal t;
  real beta;
  real p0c;
  real s;
  real phase1;
  int state, direction, time_dir, species, location;
} o_coord;

/* Global switches, subset of bmad_com (bmad/modules/bmad_struct.f90:2296-2343). */
typedef struct o_com_t {
  int aperture_limit_on;
  int absolute_time_tracking;
  double max_aperture_limit;
  double significant_length;
} o_com_t;

static o_com_t o_com = {1, 0, 1e3, 1e-10};

ORACLE_API void oracle_set_com(int aperture_limit_on, int absolute_time_tracking, double max_aperture_limit) {
  o_com.aperture_limit_on = aperture_limit_on;
  o_com.absolute_time_tracking = absolute_time_tracking;
  o_com.max_aperture_limit = max_aperture_limit;
}

/* Species of the particles the multi-particle entry points track (bunch%species);
 * 0 = the reference species passed to them. */
static int o_tracked_species = 0;
ORACLE_API void oracle_set_tracked_species(int species) { o_tracked_species = species; }
#define O_TRACKED(ref) (o_tracked_species ? o_tracked_species : (ref))
/* Direction of oracle_track_bunch: -1 tracks every element last to first with
 * direction = -1 (track_bunch, bmad/multiparticle/beam_mod.f90:131-147). */
static int o_direction = 1;
ORACLE_API void oracle_set_direction(int direction) { o_direction = (direction == -1) ? -1 : 1; }
/* Particle times of oracle_track_bunch: NULL = each element starts from t = 0
 * (relative-time tracking never reads t); else t[i] is carried from element
 * to element and turn to turn, as coord_struct%t is (absolute-time rf). */
static double *o_times = NULL;
ORACLE_API void oracle_set_times(double *t) { o_times = t; }

/* ------------------------------------------------------------------------ */
/* species: sim_utils/interfaces/particle_species_mod.f90:1505-1680         */
/* ------------------------------------------------------------------------ */
static double mass_of(int species) {
  switch (species) {
    case BMX_ELECTRON: case BMX_POSITRON: return O_M_ELECTRON;
    case BMX_PROTON: case BMX_ANTIPROTON: return O_M_PROTON;
    case BMX_MUON: case BMX_ANTIMUON: return O_M_MUON;
  }
  return 0;
}

static double charge_of(int species) {
  switch (species) {
    case BMX_ELECTRON: case BMX_ANTIPROTON: case BMX_MUON: return -1;
    case BMX_POSITRON: case BMX_PROTON: case BMX_ANTIMUON: return 1;
  }
  return 0;
}


/* charge_to_mass_of: particle_species_mod.f90:1669-1680 */
static real charge_to_mass_of(int species) { return (real)(charge_of(species)) / (real)(mass_of(species)); }

/* rel_tracking_charge_to_mass: bmad/code/rel_tracking_charge_to_mass.f90:15-35 */
static real rel_tracking_charge_to_mass(const o_coord *orb, int ref_species) {
  if (orb->species == ref_species) return (real)(1.0);
  return ((real)(charge_of(orb->species)) / (real)(mass_of(orb->species))) /
         ((real)(charge_of(ref_species)) / (real)(mass_of(ref_species)));
}

/* ------------------------------------------------------------------------ */
/* special functions: sim_utils/special_functions/ (sqrt_one, sinc, cosc, sincc, cos_one)                    */
/* ------------------------------------------------------------------------ */
/* sqrt_one.f90:18-41 */
static real sqrt_one(real x) {
  real sq = sqrt(1 + x);
  real rad = sq + 1;
  return x / rad;
}

/* sinc.f90:15-48 (nd = 0) */
static real sinc(real x) {
  if (fabs(x) < 1e-8) return (real)(1.0);
  return sin(x) / x;
}

/* cosc.f90:15-40 (nd = 0) */
static real cosc(real x) {
  if (fabs(x) < 1e-8) return (real)(0.5);
  real s = sinc(x / 2);
  return 0.5 * s * s;
}

/* sincc.f90:15-48 (nd = 0) */
static real sincc(real x) {
  const double c0 = 1.0 / 6.0, c1 = -1.0 / 120.0, c2 = 1.0 / 5040.0, c3 = -1.0 / 362880.0;
  real x2 = x * x;
  if (fabs(x) < 0.1) return c0 + x2 * (c1 + x2 * (c2 + x2 * c3));
  return (x - sin(x)) / (x * x2);
}

/* cos_one.f90 */
static real cos_one(real angle) {
  real s = sin(angle / 2);
  return -2 * s * s;
}

/* modulo2: sim_utils/special_functions/modulo2_mod.f90:85-100.
 * mod2 = modulo(x, 2 amp) (Fortran MODULO: fmod shifted to the sign of the
 * divisor); if (mod2 >= amp) mod2 -= 2 amp. */
static real modulo2(real x, real amp) {
  real p = 2 * amp;
  real m = fmod(x, p);
  if (m != 0 && m < 0) m = m + p;
  if (m >= amp) m = m - p;
  return m;
}

/* modulo2 evaluated in quad precision, as particle_rf_time does on its real(qp)
 * time (bmad/code/particle_rf_time.f90:116-118): exact when |x| < amp. */
#if defined(ORACLE_COUNT_FLOPS) || defined(ORACLE_DUAL)
static real floor_t(real x) { return floor(x); }
#else
static long double floor_t(long double x) { return floorl(x); }
#endif

static real modulo2_qp(real x, real amp) {
  long double xl, pl, m;
  if (x >= -amp && x < amp) return x;
  xl = (long double)RV(x);
  pl = 2.0L * (long double)RV(amp);
  m = fmodl(xl, pl);
  if (m != 0 && m < 0) m = m + pl;
  if (m >= (long double)RV(amp)) m = m - pl;
  return (real)((double)m);
}

static int nint_r(real x) {
  double v = RV(x);
  return (int)(v < 0 ? -floor(-v + 0.5) : floor(v + 0.5));
}

/* x**m with integer m (gfortran expands to repeated multiplication). */
static real powi(real x, int m) {
  real r = 1.0;
  int i;
  for (i = 0; i < m; i++) r = (i == 0) ? x : r * x;
  return r;
}

/* mexp: bmad/low_level/mexp.f90:14-32 */
static real mexp(real x, int m) {
  if (m < 0) return (real)(0.0);
  if (m == 0) return (real)(1.0);
  return powi(x, m);
}

/* c_multi: bmad/low_level/c_multi.f90:19-75.  Table of binomial coefficients
 * / n! with the sign pattern on m mod 4 = 0, 3. */
static double cm_tab[BMX_N_POLE][BMX_N_POLE];
static double nfact[BMX_N_POLE];
static int cm_init = 0;

static void c_multi_init(void) {
  int in, im;
  if (cm_init) return;
  memset(cm_tab, 0, sizeof(cm_tab));
  cm_tab[0][0] = 1;
  for (in = 1; in < BMX_N_POLE; in++) {
    cm_tab[in][0] = 1;
    cm_tab[in][in] = 1;
    for (im = 1; im < in; im++) cm_tab[in][im] = cm_tab[in - 1][im - 1] + cm_tab[in - 1][im];
  }
  nfact[0] = 1;
  for (in = 0; in < BMX_N_POLE; in++) {
    if (in > 0) nfact[in] = in * nfact[in - 1];
    for (im = 0; im <= in; im++) {
      cm_tab[in][im] = cm_tab[in][im] / nfact[in];
      if (im % 4 == 0) cm_tab[in][im] = -cm_tab[in][im];
      if (im % 4 == 3) cm_tab[in][im] = -cm_tab[in][im];
    }
  }
  cm_init = 1;
}

static double c_multi(int n, int m, int no_n_fact) {
  c_multi_init();
  if (no_n_fact) return cm_tab[n][m] * nfact[n];
  return cm_tab[n][m];
}

/* ------------------------------------------------------------------------ */
/* element helpers                                                           */
/* ------------------------------------------------------------------------ */
static int has_orientation_attributes(int key) {
  switch (key) {
    case BMX_KEY_MARKER: case BMX_KEY_DRIFT: return key == BMX_KEY_DRIFT; /* drift has x_offset */
    default: return 1;
  }
}

/* keys carrying HKICK/VKICK (bmad/modules/attribute_mod.f90:698-767) */
static int has_hkick_attributes(int key) {
  switch (key) {
    case BMX_KEY_QUADRUPOLE: case BMX_KEY_SBEND: case BMX_KEY_SEXTUPOLE: case BMX_KEY_OCTUPOLE:
    case BMX_KEY_THICK_MULTIPOLE: case BMX_KEY_RFCAVITY: case BMX_KEY_KICKER: case BMX_KEY_INSTRUMENT:
    case BMX_KEY_MONITOR: case BMX_KEY_ECOLLIMATOR: case BMX_KEY_RCOLLIMATOR:
      return 1;
  }
  return 0;
}

static int has_kick_attributes(int key) { return key == BMX_KEY_HKICKER || key == BMX_KEY_VKICKER; }

static int has_misalign(const bmx_ele_rec *ele) {
  /* bmad/code/attribute_bookkeeper.f90:155-180 */
  int m;
  if (ele->key == BMX_KEY_SBEND) m = (ele->ref_tilt != 0);
  else m = (ele->tilt != 0);
  m = m || ele->x_offset != 0 || ele->y_offset != 0 || ele->z_offset != 0 || ele->x_pitch != 0 ||
      ele->y_pitch != 0;
  return m;
}

static double ds_step_of(const bmx_ele_rec *ele) {
  int n = ele->num_steps > 0 ? ele->num_steps : 1;
  return ele->l / n;
}

/* physical_ele_end: bmad/low_level/physical_ele_end.f90 */
static int physical_ele_end(int track_end, const o_coord *orb, int ele_orientation) {
  int s = ele_orientation * orb->direction * orb->time_dir;
  if (s == 1) return track_end == O_FIRST_EDGE ? BMX_ENTRANCE_END : BMX_EXIT_END;
  return track_end == O_FIRST_EDGE ? BMX_EXIT_END : BMX_ENTRANCE_END;
}

/* at_this_ele_end: bmad/code/at_this_ele_end.f90 */
static int at_this_ele_end(int now_at, int where_at) {
  if (where_at == BMX_NO_END) return 0;
  if (now_at == BMX_ENTRANCE_END) return where_at == BMX_ENTRANCE_END || where_at == BMX_BOTH_ENDS;
  if (now_at == BMX_EXIT_END) return where_at == BMX_EXIT_END || where_at == BMX_BOTH_ENDS;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* multipole bookkeeping: bmad/code/multipole_ele_to_ab.f90                  */
/* ------------------------------------------------------------------------ */
static int max_nonzero(const real *a, const real *b, int n_max) {
  int n, ix = -1;
  for (n = 0; n <= n_max; n++)
    if (a[n] != 0 || b[n] != 0) ix = n;
  return ix;
}

/* tilt_this_multipole: multipole_ele_to_ab.f90:383-405 */
static void tilt_this_multipole(const bmx_ele_rec *ele, int sgn, real *a, real *b, int ix_pole_max) {
  real tilt, an, bn, cos_t, sin_t;
  int n;
  if (ele->key == BMX_KEY_SBEND) tilt = sgn * ele->ref_tilt;
  else tilt = sgn * ele->tilt;
  if (tilt == 0) return;
  for (n = 0; n <= ix_pole_max; n++) {
    if (a[n] == 0 && b[n] == 0) continue;
    an = a[n];
    bn = b[n];
    cos_t = cos((n + 1) * tilt);
    sin_t = sin((n + 1) * tilt);
    b[n] = bn * cos_t + an * sin_t;
    a[n] = -bn * sin_t + an * cos_t;
  }
}

/* pull_this_b1: multipole_ele_to_ab.f90:548-560 */
static real pull_this_b1(real *a, real *b, int *ix_pole_max) {
  real b1 = b[1];
  b[1] = 0;
  if (*ix_pole_max > 1) return b1;
  *ix_pole_max = max_nonzero(a, b, 1);
  return b1;
}

/* convert_this_ab: multipole_ele_to_ab.f90:408-500 (this_ele == ele) */
static int convert_this_ab(const bmx_ele_rec *ele, int p_type, const double *a_pole, const double *b_pole,
                           real *a, real *b) {
  int n, ix_max, ref_exp = 0;
  real factor, cnst = 0, radius;
  for (n = 0; n < BMX_N_POLE; n++) {
    a[n] = a_pole[n];
    b[n] = b_pole[n];
  }
  ix_max = max_nonzero(a, b, BMX_N_POLE - 1);
  if (ix_max < 0) return 0;
  /* tilt difference with itself is zero */
  if (ele->field_master) return -1; /* field_master multipoles: not supported */
  if (!ele->scale_multipoles) return 0;
  if (p_type == O_ELECTRIC) {
    radius = ele->r0_elec;
    if (radius != 0 && radius != 1) {
      factor = radius;
      for (n = 0; n < BMX_N_POLE; n++) {
        factor = factor / radius;
        a[n] = factor * a[n];
        b[n] = factor * b[n];
      }
    }
    return 0;
  }
  switch (ele->key) {
    case BMX_KEY_SBEND:
      cnst = (real)(ele->l) * ((real)(ele->g) + (real)(ele->dg));
      ref_exp = 0;
      break;
    case BMX_KEY_KICKER:
      if (ele->hkick == 0) cnst = ele->vkick;
      else if (ele->vkick == 0) cnst = ele->hkick;
      else cnst = sqrt((real)(ele->hkick) * ele->hkick + (real)(ele->vkick) * ele->vkick);
      ref_exp = 0;
      break;
    case BMX_KEY_HKICKER: case BMX_KEY_VKICKER:
      cnst = ele->kick;
      ref_exp = 0;
      break;
    case BMX_KEY_QUADRUPOLE:
      cnst = (real)(ele->k1) * ele->l;
      ref_exp = 1;
      break;
    case BMX_KEY_SEXTUPOLE:
      cnst = (real)(ele->k2) * ele->l;
      ref_exp = 2;
      break;
    case BMX_KEY_OCTUPOLE:
      cnst = (real)(ele->k3) * ele->l;
      ref_exp = 3;
      break;
    case BMX_KEY_AB_MULTIPOLE: case BMX_KEY_THICK_MULTIPOLE: case BMX_KEY_RFCAVITY:
      return 0;
    default:
      return -1; /* non-scalable element with multipoles: reference err_exit */
  }
  radius = ele->r0_mag;
  if (radius == 0) {
    for (n = 0; n < BMX_N_POLE; n++) {
      a[n] = cnst * a[n];
      b[n] = cnst * b[n];
    }
  } else {
    factor = cnst * powi(radius, ref_exp + 1);
    for (n = 0; n < BMX_N_POLE; n++) {
      factor = factor / radius;
      a[n] = factor * a[n];
      b[n] = factor * b[n];
    }
  }
  return 0;
}

/* this_ele_non_multipoles: multipole_ele_to_ab.f90:326-405 */
static void this_ele_non_multipoles(const bmx_ele_rec *ele, int p_type, real *a_kick, real *b_kick) {
  int i;
  real hk, vk, tilt, cos_t, sin_t;
  for (i = 0; i < 4; i++) a_kick[i] = b_kick[i] = 0;
  if (p_type == O_MAGNETIC) {
    switch (ele->key) {
      case BMX_KEY_HKICKER: b_kick[0] = -ele->kick; break;
      case BMX_KEY_VKICKER: a_kick[0] = ele->kick; break;
      case BMX_KEY_KICKER:
        b_kick[0] = -ele->hkick;
        a_kick[0] = ele->vkick;
        break;
      default:
        hk = ele->hkick;
        vk = ele->vkick;
        if (has_hkick_attributes(ele->key) && (hk != 0 || vk != 0)) {
          if (ele->key == BMX_KEY_SBEND) tilt = ele->ref_tilt;
          else tilt = ele->tilt;
          if (tilt != 0) {
            cos_t = cos(tilt);
            sin_t = sin(tilt);
            b_kick[0] = -hk * cos_t - vk * sin_t;
            a_kick[0] = -hk * sin_t + vk * cos_t;
          } else {
            b_kick[0] = -hk;
            a_kick[0] = vk;
          }
        }
    }
    switch (ele->key) {
      case BMX_KEY_QUADRUPOLE: b_kick[1] = (real)(ele->k1) * ele->l; break;
      case BMX_KEY_SBEND:
        b_kick[0] = b_kick[0] + (real)(ele->dg) * ele->l;
        b_kick[1] = (real)(ele->k1) * ele->l;
        b_kick[2] = (real)(ele->k2) * ele->l / 2.0;
        break;
      case BMX_KEY_SEXTUPOLE: b_kick[2] = (real)(ele->k2) * ele->l / 2.0; break;
      case BMX_KEY_OCTUPOLE: b_kick[3] = (real)(ele->k3) * ele->l / 6.0; break;
    }
  }
}

/* multipole1_kt_to_ab: bmad/low_level/multipole1_kt_to_ab.f90 */
static void multipole1_kt_to_ab(real knl, real knsl, real tn, int n, real *an, real *bn) {
  real angle;
  if (knl == 0 && knsl == 0) {
    *bn = 0;
    *an = 0;
  } else {
    angle = -tn * (n + 1);
    *bn = (knl * cos(angle) - knsl * sin(angle)) / nfact[n];
    *an = (knl * sin(angle) + knsl * cos(angle)) / nfact[n];
  }
}

/* multipole1_ab_to_kt: bmad/low_level/multipole1_ab_to_kt.f90 */
static void multipole1_ab_to_kt(real an, real bn, int n, real *knl, real *tn) {
  if (an == 0 && bn == 0) {
    *knl = 0;
    *tn = 0;
  } else {
    *knl = nfact[n] * sqrt(an * an + bn * bn);
    *tn = -atan2(an, bn) / (n + 1);
    if (2 * (n + 1) * fabs(*tn) > O_PI) {
      *knl = -*knl;
      *tn = (*tn > 0 ? 1.0 : -1.0) * (fabs(*tn) - O_PI / (n + 1));
    }
  }
}

/* multipole_ele_to_ab: bmad/code/multipole_ele_to_ab.f90:31-235 (no lords, no cache) */
static int multipole_ele_to_ab(const bmx_ele_rec *ele, int use_ele_tilt, int p_type, int include_kicks,
                               int *ix_pole_max, real *a, real *b, real *b1) {
  int n;
  real a_kick[4], b_kick[4];
  c_multi_init();
  *ix_pole_max = -1;
  for (n = 0; n < BMX_N_POLE; n++) a[n] = b[n] = 0;
  if (b1) *b1 = 0;
  if (!ele->is_on || ele->key == BMX_KEY_PIPE) return 0;

  if (ele->key == BMX_KEY_MULTIPOLE) {
    if (p_type == O_ELECTRIC || !ele->has_mag_poles) return 0;
    for (n = 0; n < BMX_N_POLE; n++) multipole1_kt_to_ab(ele->a[n], 0.0, ele->b[n], n, &a[n], &b[n]);
    *ix_pole_max = max_nonzero(a, b, BMX_N_POLE - 1);
    if (b1) *b1 = pull_this_b1(a, b, ix_pole_max);
    return 0;
  }

  if (p_type == O_MAGNETIC && ele->has_mag_poles) {
    if (convert_this_ab(ele, p_type, ele->a, ele->b, a, b) < 0) return -1;
  } else if (p_type == O_ELECTRIC && ele->has_elec_poles) {
    if (convert_this_ab(ele, p_type, ele->a_elec, ele->b_elec, a, b) < 0) return -1;
  }
  this_ele_non_multipoles(ele, p_type, a_kick, b_kick);
  if (include_kicks) {
    for (n = 0; n < 4; n++) {
      a[n] = a[n] + a_kick[n];
      b[n] = b[n] + b_kick[n];
    }
  }
  *ix_pole_max = max_nonzero(a, b, BMX_N_POLE - 1);
  if (use_ele_tilt) tilt_this_multipole(ele, +1, a, b, *ix_pole_max);
  if (b1) *b1 = pull_this_b1(a, b, ix_pole_max);
  return 0;
}

/* multipole_ele_to_kt: bmad/code/multipole_ele_to_kt.f90 (magnetic) */
static void multipole_ele_to_kt(const bmx_ele_rec *ele, int use_ele_tilt, int *ix_pole_max, real *knl, real *tilt) {
  int n;
  real a[BMX_N_POLE], b[BMX_N_POLE];
  c_multi_init();
  if (ele->key == BMX_KEY_MULTIPOLE) {
    for (n = 0; n < BMX_N_POLE; n++) {
      knl[n] = ele->has_mag_poles ? ele->a[n] : 0.0;
      tilt[n] = (ele->has_mag_poles ? ele->b[n] : 0.0) + ele->tilt;
    }
    *ix_pole_max = -1;
    for (n = 0; n < BMX_N_POLE; n++)
      if (knl[n] != 0) *ix_pole_max = n;
    return;
  }
  multipole_ele_to_ab(ele, use_ele_tilt, O_MAGNETIC, 0, ix_pole_max, a, b, NULL);
  if (*ix_pole_max > -1) {
    for (n = 0; n < BMX_N_POLE; n++) multipole1_ab_to_kt(a[n], b[n], n, &knl[n], &tilt[n]);
  } else {
    for (n = 0; n < BMX_N_POLE; n++) knl[n] = tilt[n] = 0;
  }
}

/* ------------------------------------------------------------------------ */
/* track_a_drift: bmad/low_level/track_a_drift.f90:25-116                    */
/* ------------------------------------------------------------------------ */
static void track_a_drift(o_coord *orb, real length, int ele_orientation, int include_ref_motion, real *time) {
  real delta, rel_pc, px_rel, py_rel, pxy2, ps_rel, mc2, dz, beta_ref, dt;
  int rel_z_vel;
  if (length == 0) return;
  delta = orb->vec[5];
  rel_pc = 1 + delta;
  rel_z_vel = (ele_orientation == 0 ? orb->direction : ele_orientation) * orb->direction;
  px_rel = rel_z_vel * orb->vec[1] / rel_pc;
  py_rel = rel_z_vel * orb->vec[3] / rel_pc;
  pxy2 = px_rel * px_rel + py_rel * py_rel;
  if (pxy2 >= 1) {
    orb->state = BMX_LOST_PZ;
    return;
  }
  ps_rel = sqrt(1 - pxy2);
  orb->vec[0] = orb->vec[0] + length * px_rel / ps_rel;
  orb->vec[2] = orb->vec[2] + length * py_rel / ps_rel;
  if (orb->beta > 0) {
    if (include_ref_motion) {
      mc2 = mass_of(orb->species);
      if (orb->direction == 1) {
        dz = length * (sqrt_one((mc2 * mc2 * (2 * delta + delta * delta)) /
                                ((orb->p0c * rel_pc) * (orb->p0c * rel_pc) + mc2 * mc2)) +
                       sqrt_one(-pxy2) / ps_rel);
      } else {
        beta_ref = orb->p0c / sqrt(orb->p0c * orb->p0c + mc2 * mc2);
        dz = length * ((-orb->beta / beta_ref - 1.0) - (1.0 / ps_rel - 1.0));
      }
      orb->s = orb->s + orb->direction * length;
    } else {
      dz = -length / ps_rel;
    }
    NOCOUNT_BEGIN;
    dt = rel_z_vel * length / (orb->beta * ps_rel * O_C_LIGHT);
    orb->t = orb->t + dt;
    if (time) *time = *time + dt;
    NOCOUNT_END;
  } else {
    if (include_ref_motion) {
      dz = length * (1 - 1 / ps_rel);
      orb->s = orb->s + orb->direction * length;
    } else {
      dz = -length / ps_rel;
    }
  }
  orb->vec[4] = orb->vec[4] + rel_z_vel * dz;
}

/* tilt_coords: bmad/code/tilt_coords.f90 */
static void tilt_coords(real tilt_val, real *vec) {
  real s, c, x, y, px, py;
  if (tilt_val == 0) return;
  s = sin(tilt_val);
  c = cos(tilt_val);
  x = vec[0]; y = vec[2];
  vec[0] = c * x + s * y;
  vec[2] = -s * x + c * y;
  px = vec[1]; py = vec[3];
  vec[1] = c * px + s * py;
  vec[3] = -s * px + c * py;
}

/* ------------------------------------------------------------------------ */
/* 3-D geometry helpers: bmad/geometry, sim_utils/math/rotation_3d_mod.f90 */
/* ------------------------------------------------------------------------ */
typedef struct o_pos { real r[3]; real w[3][3]; } o_pos;

static void mat3_unit(real m[3][3]) {
  int i, j;
  for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) m[i][j] = (i == j) ? 1.0 : 0.0;
}

static void mat3_vec(real m[3][3], const real *v, real *out) {
  int i;
  real t[3];
  for (i = 0; i < 3; i++) t[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
  for (i = 0; i < 3; i++) out[i] = t[i];
}

static void mat3_mul(real a[3][3], real b[3][3], real out[3][3]) {
  int i, j;
  real t[3][3];
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
  for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) out[i][j] = t[i][j];
}

static void mat3_transpose(real a[3][3]) {
  int i, j;
  real t;
  for (i = 0; i < 3; i++)
    for (j = i + 1; j < 3; j++) { t = a[i][j]; a[i][j] = a[j][i]; a[j][i] = t; }
}

enum { O_X_AXIS = 1, O_Y_AXIS = 2, O_Z_AXIS = 3 };

/* rotate_mat (left multiply): sim_utils/math/rotation_3d_mod.f90 */
static void rotate_mat(real m[3][3], int axis, real angle) {
  real ca, sa, p, q;
  int j;
  if (angle == 0) return;
  ca = cos(angle);
  sa = sin(angle);
  for (j = 0; j < 3; j++) {
    if (axis == O_X_AXIS) {
      p = m[1][j]; q = m[2][j];
      m[1][j] = ca * p - sa * q;
      m[2][j] = sa * p + ca * q;
    } else if (axis == O_Y_AXIS) {
      p = m[0][j]; q = m[2][j];
      m[0][j] = sa * q + ca * p;
      m[2][j] = ca * q - sa * p;
    } else {
      p = m[0][j]; q = m[1][j];
      m[0][j] = ca * p - sa * q;
      m[1][j] = sa * p + ca * q;
    }
  }
}

/* rotate_vec: sim_utils/math/rotation_3d_mod.f90 */
static void rotate_vec(real *v, int axis, real angle) {
  real ca, sa, p, q;
  if (angle == 0) return;
  ca = cos(angle);
  sa = sin(angle);
  if (axis == O_X_AXIS) {
    p = v[1]; q = v[2];
    v[1] = ca * p - sa * q;
    v[2] = sa * p + ca * q;
  } else if (axis == O_Y_AXIS) {
    p = v[0]; q = v[2];
    v[0] = sa * q + ca * p;
    v[2] = ca * q - sa * p;
  } else {
    p = v[0]; q = v[1];
    v[0] = ca * p - sa * q;
    v[1] = sa * p + ca * q;
  }
}

/* floor_angles_to_w_mat: bmad/geometry/floor_angles_to_w_mat.f90 */
static void floor_angles_to_w_mat(real theta, real phi, real psi, real w[3][3], int inverse) {
  real s_the = sin(theta), c_the = cos(theta);
  real s_phi = sin(phi), c_phi = cos(phi);
  real s_psi = sin(psi), c_psi = cos(psi);
  if (!inverse) {
    w[0][0] = c_the * c_psi - s_the * s_phi * s_psi;
    w[0][1] = -c_the * s_psi - s_the * s_phi * c_psi;
    w[0][2] = s_the * c_phi;
    w[1][0] = c_phi * s_psi;
    w[1][1] = c_phi * c_psi;
    w[1][2] = s_phi;
    w[2][0] = -s_the * c_psi - c_the * s_phi * s_psi;
    w[2][1] = s_the * s_psi - c_the * s_phi * c_psi;
    w[2][2] = c_the * c_phi;
  } else {
    w[0][0] = c_the * c_psi - s_the * s_phi * s_psi;
    w[0][1] = c_phi * s_psi;
    w[0][2] = -s_the * c_psi - c_the * s_phi * s_psi;
    w[1][0] = -c_the * s_psi - s_the * s_phi * c_psi;
    w[1][1] = c_phi * c_psi;
    w[1][2] = s_the * s_psi - c_the * s_phi * c_psi;
    w[2][0] = s_the * c_phi;
    w[2][1] = s_phi;
    w[2][2] = c_the * c_phi;
  }
}

/* w_mat_for_tilt: bmad/geometry/w_mat_for_tilt.f90 */
static void w_mat_for_tilt(real tilt, real w[3][3]) {
  real c = cos(tilt), s = sin(tilt);
  w[0][0] = c; w[0][1] = -s; w[0][2] = 0;
  w[1][0] = s; w[1][1] = c; w[1][2] = 0;
  w[2][0] = 0; w[2][1] = 0; w[2][2] = 1;
}

/* bend_shift: bmad/geometry/bend_shift.f90 */
static o_pos bend_shift(o_pos p1, real g, real delta_s, int has_ref_tilt, real ref_tilt) {
  o_pos p2;
  real S[3][3], L[3], angle, tlt;
  angle = delta_s * g;
  if (angle == 0) {
    p2 = p1;
    p2.r[2] = p2.r[2] - delta_s;
    return p2;
  }
  tlt = has_ref_tilt ? ref_tilt : (real)(0.0);
  mat3_unit(S);
  if (tlt != 0) {
    rotate_mat(S, O_Z_AXIS, -ref_tilt);
    rotate_mat(S, O_Y_AXIS, angle);
    rotate_mat(S, O_Z_AXIS, ref_tilt);
  } else {
    rotate_mat(S, O_Y_AXIS, angle);
  }
  L[0] = cos_one(angle) / g;
  L[1] = 0.0;
  L[2] = -sin(angle) / g;
  if (has_ref_tilt) rotate_vec(L, O_Z_AXIS, ref_tilt);
  mat3_vec(S, p1.r, p2.r);
  p2.r[0] = p2.r[0] + L[0];
  p2.r[1] = p2.r[1] + L[1];
  p2.r[2] = p2.r[2] + L[2];
  mat3_mul(S, p1.w, p2.w);
  return p2;
}

/* ele_misalignment_L_S_calc: bmad/geometry/ele_misalignment_L_S_calc.f90 (roll = 0) */
static void ele_misalignment_L_S_calc(const bmx_ele_rec *ele, real *L_mis, real S_mis[3][3]) {
  L_mis[0] = ele->x_offset;
  L_mis[1] = ele->y_offset;
  L_mis[2] = ele->z_offset;
  if (ele->key == BMX_KEY_SBEND) floor_angles_to_w_mat(ele->x_pitch, ele->y_pitch, 0.0, S_mis, 0);
  else floor_angles_to_w_mat(ele->x_pitch, ele->y_pitch, ele->tilt, S_mis, 0);
}

/* ------------------------------------------------------------------------ */
/* offset_particle: bmad/code/offset_particle.f90:66-530                     */
/*   drift_to: 0 = no$, 1 = upstream_end$, 2 = downstream_end$, -1 default   */
/* ------------------------------------------------------------------------ */
static void offset_particle(const bmx_ele_rec *ele, int ref_species, int set, o_coord *orb, int set_tilt,
                            int set_hvkicks, int drift_to_edge) {
  int drift_to, sign_z_vel, key = ele->key, is_misaligned = 0, set_hv, set_t, set_hv1 = 0, set_hv2 = 0;
  real length, L_half, s_pos0, ds_center, s_target, rel_p, charge_dir = 0, x_off = 0, y_off = 0, z_off = 0;
  real xp = 0, yp = 0, ref_tilt = 0, s_body, s_lab, pz, ds, ws[3][3], L_mis[3], p_vec0[3], p_vec[3];
  o_pos position;
  int i;

  set_hv = set_hvkicks && ele->is_on && (has_kick_attributes(key) || has_hkick_attributes(key));
  set_t = set_tilt && has_orientation_attributes(key);

  if (!has_orientation_attributes(key)) drift_to = 0;
  else if (set) drift_to = (drift_to_edge < 0) ? O_UPSTREAM : drift_to_edge;
  else drift_to = (drift_to_edge < 0) ? O_DOWNSTREAM : drift_to_edge;

  length = ele->l;
  L_half = 0.5 * length;
  sign_z_vel = ele->orientation * orb->direction;
  if ((set && orb->direction * orb->time_dir == 1) || (!set && sign_z_vel * orb->time_dir == -1)) s_pos0 = 0;
  else s_pos0 = length;
  ds_center = L_half - s_pos0;
  if (drift_to == O_UPSTREAM) s_target = (1 - sign_z_vel * orb->time_dir) * L_half;
  else s_target = (1 + orb->direction * orb->time_dir) * L_half;

  rel_p = 1 + orb->vec[5];
  if (set_hv) {
    charge_dir = rel_tracking_charge_to_mass(orb, ref_species) * sign_z_vel * orb->time_dir;
    if (key == BMX_KEY_KICKER || key == BMX_KEY_HKICKER || key == BMX_KEY_VKICKER) set_hv2 = 1;
    else set_hv1 = 1;
  }

  if (has_orientation_attributes(key)) {
    x_off = ele->x_offset;
    y_off = ele->y_offset;
    z_off = ele->z_offset;
    xp = ele->x_pitch;
    yp = ele->y_pitch;
    ref_tilt = ele->ref_tilt;
    is_misaligned = (x_off != 0 || y_off != 0 || z_off != 0 || xp != 0 || yp != 0);
  }

  if (set) {
    if (is_misaligned) {
      position.r[0] = orb->vec[0];
      position.r[1] = orb->vec[2];
      position.r[2] = 0;
      mat3_unit(position.w);
      if (key == BMX_KEY_SBEND && (ele->g != 0 || ref_tilt != 0)) {
        position = bend_shift(position, ele->g, ele->orientation * ds_center, 1, ref_tilt);
        ele_misalignment_L_S_calc(ele, L_mis, ws);
        mat3_transpose(ws);
        for (i = 0; i < 3; i++) position.r[i] = position.r[i] - L_mis[i];
        mat3_vec(ws, position.r, position.r);
        mat3_mul(ws, position.w, position.w);
        if (ref_tilt != 0) {
          position = bend_shift(position, ele->g, -L_half, 1, ref_tilt);
          w_mat_for_tilt(-ref_tilt, ws);
          mat3_vec(ws, position.r, position.r);
          mat3_mul(ws, position.w, position.w);
          position = bend_shift(position, ele->g, L_half, 0, 0.0);
        }
        if (drift_to != 0) {
          position = bend_shift(position, ele->g, -sign_z_vel * orb->time_dir * L_half, 0, 0.0);
          position.r[2] = position.r[2] + (1 - sign_z_vel * orb->time_dir) * L_half;
        } else {
          if (ele->g == 0) ds = -ele->orientation * ds_center;
          else ds = (1 / (real)(ele->g)) * atan2(position.r[2] * ele->g, 1.0 + position.r[0] * ele->g);
          position = bend_shift(position, ele->g, ds, 0, 0.0);
          position.r[2] = position.r[2] + L_half + ds;
        }
      } else {
        position.r[0] = position.r[0] - x_off;
        position.r[1] = position.r[1] - y_off;
        position.r[2] = position.r[2] - (z_off + ele->orientation * ds_center);
        floor_angles_to_w_mat(xp, yp, 0.0, ws, 1);
        mat3_vec(ws, position.r, position.r);
        mat3_mul(ws, position.w, position.w);
        position.r[2] = position.r[2] + L_half;
      }
      s_body = position.r[2];
      pz = rel_p * rel_p - orb->vec[1] * orb->vec[1] - orb->vec[3] * orb->vec[3];
      if (pz <= 0) {
        orb->state = BMX_LOST_PZ;
      } else {
        p_vec0[0] = orb->vec[1];
        p_vec0[1] = orb->vec[3];
        p_vec0[2] = sign_z_vel * sqrt(pz);
        mat3_vec(position.w, p_vec0, p_vec);
        orb->vec[1] = p_vec[0];
        orb->vec[3] = p_vec[1];
        orb->vec[0] = position.r[0];
        orb->vec[2] = position.r[1];
      }
    } else {
      if (key == BMX_KEY_SBEND && ele->ref_tilt != 0) tilt_coords(ele->ref_tilt, orb->vec);
      if (ele->orientation == 1) s_body = s_pos0;
      else s_body = length - s_pos0;
    }
    ds = s_target - s_body;
    if (fabs(ds) > 1e-14 && drift_to != 0 && orb->state == BMX_ALIVE)
      track_a_drift(orb, ds, ele->orientation, 0, NULL);

    if (set_hv1) {
      orb->vec[1] = orb->vec[1] + 0.5 * charge_dir * ele->hkick;
      orb->vec[3] = orb->vec[3] + 0.5 * charge_dir * ele->vkick;
    }
    if (set_t && key != BMX_KEY_SBEND && ele->tilt != 0) tilt_coords(ele->tilt, orb->vec);
    if (set_hv2) {
      if (key == BMX_KEY_HKICKER) {
        orb->vec[1] = orb->vec[1] + 0.5 * charge_dir * ele->kick;
      } else if (key == BMX_KEY_VKICKER) {
        orb->vec[3] = orb->vec[3] + 0.5 * charge_dir * ele->kick;
      } else {
        orb->vec[1] = orb->vec[1] + 0.5 * charge_dir * ele->hkick;
        orb->vec[3] = orb->vec[3] + 0.5 * charge_dir * ele->vkick;
      }
    }
  } else {
    if (set_hv2) {
      if (key == BMX_KEY_HKICKER) {
        orb->vec[1] = orb->vec[1] + 0.5 * charge_dir * ele->kick;
      } else if (key == BMX_KEY_VKICKER) {
        orb->vec[3] = orb->vec[3] + 0.5 * charge_dir * ele->kick;
      } else {
        orb->vec[1] = orb->vec[1] + 0.5 * charge_dir * ele->hkick;
        orb->vec[3] = orb->vec[3] + 0.5 * charge_dir * ele->vkick;
      }
    }
    if (set_t && key != BMX_KEY_SBEND) tilt_coords(-ele->tilt, orb->vec);
    if (set_hv1) {
      orb->vec[1] = orb->vec[1] + 0.5 * charge_dir * ele->hkick;
      orb->vec[3] = orb->vec[3] + 0.5 * charge_dir * ele->vkick;
    }
    if (is_misaligned) {
      position.r[0] = orb->vec[0];
      position.r[1] = orb->vec[2];
      position.r[2] = 0;
      mat3_unit(position.w);
      if (key == BMX_KEY_SBEND && (ele->g != 0 || ref_tilt != 0)) {
        position = bend_shift(position, ele->g, ds_center, 0, 0.0);
        if (ref_tilt != 0) {
          position = bend_shift(position, ele->g, -L_half, 0, 0.0);
          w_mat_for_tilt(ref_tilt, ws);
          mat3_vec(ws, position.r, position.r);
          mat3_mul(ws, position.w, position.w);
          position = bend_shift(position, ele->g, L_half, 1, ref_tilt);
        }
        ele_misalignment_L_S_calc(ele, L_mis, ws);
        mat3_vec(ws, position.r, position.r);
        for (i = 0; i < 3; i++) position.r[i] = position.r[i] + L_mis[i];
        mat3_mul(ws, position.w, position.w);
        if (drift_to != 0) {
          position = bend_shift(position, ele->g, sign_z_vel * orb->time_dir * L_half, 1, ref_tilt);
          position.r[2] = position.r[2] + (1 + sign_z_vel * orb->time_dir) * L_half;
        } else {
          if (ele->g == 0) ds = -ds_center;
          else ds = (1 / (real)(ele->g)) * atan2(position.r[2] * ele->g, 1.0 + position.r[0] * ele->g);
          position = bend_shift(position, ele->g, ds, 1, ref_tilt);
          position.r[2] = position.r[2] + L_half + ds;
        }
      } else {
        position.r[2] = position.r[2] - ds_center;
        floor_angles_to_w_mat(xp, yp, 0.0, ws, 0);
        mat3_vec(ws, position.r, position.r);
        mat3_mul(ws, position.w, position.w);
        position.r[0] = position.r[0] + x_off;
        position.r[1] = position.r[1] + y_off;
        position.r[2] = position.r[2] + (z_off + L_half);
      }
      if (ele->orientation == -1) position.r[2] = length - position.r[2];
      s_lab = position.r[2];
      pz = rel_p * rel_p - orb->vec[1] * orb->vec[1] - orb->vec[3] * orb->vec[3];
      if (pz <= 0) {
        orb->state = BMX_LOST_PZ;
      } else {
        p_vec0[0] = orb->vec[1];
        p_vec0[1] = orb->vec[3];
        p_vec0[2] = sign_z_vel * sqrt(pz);
        mat3_vec(position.w, p_vec0, p_vec);
        orb->vec[1] = p_vec[0];
        orb->vec[3] = p_vec[1];
        orb->vec[0] = position.r[0];
        orb->vec[2] = position.r[1];
      }
    } else {
      if (key == BMX_KEY_SBEND && ele->ref_tilt != 0) tilt_coords(-ele->ref_tilt, orb->vec);
      if (ele->orientation == 1) s_lab = s_pos0;
      else s_lab = length - s_pos0;
    }
    ds = s_target - s_lab;
    if (fabs(ds) > 1e-14 && drift_to != 0 && orb->state == BMX_ALIVE) track_a_drift(orb, ds, +1, 0, NULL);
  }
}

/* ------------------------------------------------------------------------ */
/* multipole kicks: bmad/modules/multipole_mod.f90                           */
/* ------------------------------------------------------------------------ */
/* ab_multipole_kick: multipole_mod.f90:314-401 (no dk) */
static void ab_multipole_kick(real a, real b, int n, int ref_species, int ele_orientation, const o_coord *coord,
                              real *kx, real *ky, int pole_type, int has_scale, real scale) {
  real f, a2, b2, x, y;
  int m;
  *kx = 0;
  *ky = 0;
  if (a == 0 && b == 0) return;
  if (pole_type == O_ELECTRIC) {
    f = (real)(charge_of(coord->species)) / (coord->beta * coord->p0c);
    a2 = a * f;
    b2 = -b * f;
  } else {
    if (ele_orientation == 0) {
      a2 = a;
      b2 = b;
    } else {
      f = coord->direction * ele_orientation * charge_to_mass_of(coord->species) / charge_to_mass_of(ref_species);
      a2 = a * f;
      b2 = b * f;
    }
  }
  if (has_scale) {
    a2 = scale * a2;
    b2 = scale * b2;
  }
  x = coord->vec[0];
  y = coord->vec[2];
  for (m = 0; m <= n; m += 2) {
    f = c_multi(n, m, 1) * mexp(x, n - m) * mexp(y, m);
    *kx = *kx + b2 * f;
    *ky = *ky - a2 * f;
  }
  for (m = 1; m <= n; m += 2) {
    f = c_multi(n, m, 1) * mexp(x, n - m) * mexp(y, m);
    *kx = *kx + a2 * f;
    *ky = *ky + b2 * f;
  }
}

/* ab_multipole_kicks: multipole_mod.f90:82-174 */
static void ab_multipole_kicks(const real *an, const real *bn, int ix_pole_max, const bmx_ele_rec *ele, int ref_species,
                               o_coord *orbit, int pole_type, real scale) {
  real kx, ky, kx_tot = 0, ky_tot = 0, r, alpha, dpz, mc;
  o_coord orb0 = *orbit;
  int n;
  r = ele->p0c / orbit->p0c;
  for (n = 0; n <= ix_pole_max; n++) {
    ab_multipole_kick(r * an[n], r * bn[n], n, ref_species, ele->orientation, orbit, &kx, &ky, pole_type, 1, scale);
    kx_tot = kx_tot + kx;
    ky_tot = ky_tot + ky;
  }
  orbit->vec[1] = orbit->vec[1] + kx_tot;
  orbit->vec[3] = orbit->vec[3] + ky_tot;
  if (pole_type == O_ELECTRIC) {
    alpha = (kx_tot * (2 * orb0.vec[1] + kx_tot) + ky_tot * (2 * orb0.vec[3] + ky_tot)) /
            ((1 + orb0.vec[5]) * (1 + orb0.vec[5]));
    if (alpha < -1) {
      orbit->state = BMX_LOST_PZ;
      return;
    }
    dpz = (1 + orb0.vec[5]) * sqrt_one(alpha);
    orbit->vec[5] = orb0.vec[5] + dpz;
    mc = mass_of(orbit->species) / orbit->p0c;
    orbit->beta = (1 + orbit->vec[5]) / sqrt((1 + orbit->vec[5]) * (1 + orbit->vec[5]) + mc * mc);
    orbit->vec[4] = orb0.vec[4] * orbit->beta / orb0.beta;
  }
}

/* multipole_kick (kt form): multipole_mod.f90:203-286 */
static void multipole_kick(real knl, real tilt, int n, int ref_species, int ele_orientation, o_coord *coord,
                           int ref_orb_offset) {
  real t, charge, sin_ang, cos_ang, x, y, x_terms[BMX_N_POLE], y_terms[BMX_N_POLE], x_value, y_value, x_vel, y_vel;
  int m;
  if (knl == 0) return;
  c_multi_init();
  t = tilt;
  charge = coord->direction * coord->time_dir * ele_orientation * charge_to_mass_of(coord->species) /
           charge_to_mass_of(ref_species);
  if (t == 0) {
    sin_ang = 0;
    cos_ang = 1;
    x = coord->vec[0];
    y = coord->vec[2];
  } else {
    sin_ang = sin(t);
    cos_ang = cos(t);
    x = coord->vec[0] * cos_ang + coord->vec[2] * sin_ang;
    y = -coord->vec[0] * sin_ang + coord->vec[2] * cos_ang;
  }
  if (n == 0 && ref_orb_offset) {
    coord->vec[1] = coord->vec[1] + charge * knl * cos_ang * coord->vec[5];
    coord->vec[3] = coord->vec[3] + charge * knl * sin_ang * coord->vec[5];
    coord->vec[4] = coord->vec[4] - charge * knl * (cos_ang * coord->vec[0] + sin_ang * coord->vec[2]);
    return;
  }
  x_terms[n] = 1.0;
  y_terms[0] = 1.0;
  for (m = 1; m <= n; m++) {
    x_terms[n - m] = x_terms[n - m + 1] * x;
    y_terms[m] = y_terms[m - 1] * y;
  }
  x_value = 0;
  for (m = 0; m <= n; m += 2) x_value = x_value + cm_tab[n][m] * x_terms[m] * y_terms[m];
  y_value = 0;
  for (m = 1; m <= n; m += 2) y_value = y_value + cm_tab[n][m] * x_terms[m] * y_terms[m];
  x_vel = charge * knl * x_value;
  y_vel = charge * knl * y_value;
  if (t == 0) {
    coord->vec[1] = coord->vec[1] + x_vel;
    coord->vec[3] = coord->vec[3] + y_vel;
  } else {
    coord->vec[1] = coord->vec[1] + x_vel * cos_ang - y_vel * sin_ang;
    coord->vec[3] = coord->vec[3] + x_vel * sin_ang + y_vel * cos_ang;
  }
}

/* ------------------------------------------------------------------------ */
/* energy kick: bmad/low_level/apply_energy_kick.f90:23-74                   */
/* ------------------------------------------------------------------------ */
static void apply_energy_kick(real dE, o_coord *orbit) {
  real mc2 = mass_of(orbit->species), p0c = orbit->p0c, pc, beta_old, E_old, E_new, pc_new, beta_new;
  pc = (1 + orbit->vec[5]) * p0c;
  beta_old = orbit->beta;
  E_old = pc / orbit->beta;
  E_new = E_old + dE;
  if (E_new < mc2) {
    orbit->vec[5] = -1;
    orbit->beta = 0;
    orbit->state = BMX_LOST_PZ;
    return;
  }
  orbit->vec[5] = orbit->vec[5] + (1 + orbit->vec[5]) * sqrt_one((2 * E_old * dE + dE * dE) / (pc * pc));
  pc_new = p0c * (1 + orbit->vec[5]);
  beta_new = pc_new / E_new;
  orbit->vec[4] = orbit->vec[4] * beta_new / beta_old;
  orbit->beta = beta_new;
}

/* ------------------------------------------------------------------------ */
/* edge kicks: bmad/modules/fringe_mod.f90                                   */
/* ------------------------------------------------------------------------ */
static int fringe_here(const bmx_ele_rec *ele, const o_coord *orb, int particle_at) {
  /* bmad/code/fringe_here.f90:15-35 */
  int physical_end = physical_ele_end(particle_at, orb, ele->orientation);
  int is_here = at_this_ele_end(physical_end, ele->fringe_at);
  if (ele->fringe_type == BMX_FRINGE_NONE) is_here = 0;
  return is_here;
}

/* hard_multipole_edge_kick: fringe_mod.f90:691-864 (quadrupole, sextupole) */
static void hard_multipole_edge_kick(const bmx_ele_rec *ele, int ref_species, int particle_at, o_coord *orbit) {
  int n, n_max;
  real rel_p, cn, x, y, px, py, denom, ap, bp, charge_dir;
  real fx = 0, dfx_dx = 0, dfx_dy = 0, fy = 0, dfy_dx = 0, dfy_dy = 0;
  real pr, pi_, p1r, p1i, tr, ti, cr, ci, dpxr, dpxi, dpyr, dpyi, xnr, xni, qr, qi;
  switch (ele->fringe_type) {
    case BMX_FRINGE_HARD_EDGE_ONLY:
      if (ele->key == BMX_KEY_SBEND) return;
      break;
    case BMX_FRINGE_FULL: break;
    default: return;
  }
  if (!fringe_here(ele, orbit, particle_at)) return;
  charge_dir = ele->orientation * orbit->direction;
  charge_dir = charge_dir * rel_tracking_charge_to_mass(orbit, ref_species);
  if (ele->key == BMX_KEY_QUADRUPOLE || ele->key == BMX_KEY_SBEND) n_max = 1;
  else if (ele->key == BMX_KEY_SEXTUPOLE) n_max = 2;
  else return;
  x = orbit->vec[0];
  y = orbit->vec[2];
  /* complex arithmetic written out: poly_n1 = 1, poly = x + i y */
  p1r = 1; p1i = 0;
  pr = x; pi_ = y;
  rel_p = 1 + orbit->vec[5];
  for (n = 1; n <= n_max; n++) {
    /* poly_n1 = poly; poly = poly * xy */
    p1r = pr; p1i = pi_;
    tr = pr * x - pi_ * y;
    ti = pr * y + pi_ * x;
    pr = tr; pi_ = ti;
    if (ele->key == BMX_KEY_SEXTUPOLE) {
      ap = 0;
      bp = ele->k2;
      if (n != 2) continue;
    } else {
      ap = 0;
      bp = ele->k1;
    }
    /* dpoly_dx = (n+1) poly_n1, dpoly_dy = i dpoly_dx */
    dpxr = (n + 1) * p1r; dpxi = (n + 1) * p1i;
    dpyr = -dpxi; dpyi = dpxr;
    /* cab = charge_dir * (bp + i ap) / (4 (n+2) rel_p) */
    cr = charge_dir * bp / (4 * (n + 2) * rel_p);
    ci = charge_dir * ap / (4 * (n + 2) * rel_p);
    if (particle_at == O_FIRST_EDGE) { cr = -cr; ci = -ci; }
    cn = (real)(n + 3) / (real)(n + 1);
    /* xny = x - i cn y */
    xnr = x; xni = -cn * y;
    /* fx += Re(cab * poly * xny) */
    qr = pr * xnr - pi_ * xni; qi = pr * xni + pi_ * xnr;
    fx = fx + (cr * qr - ci * qi);
    /* dfx_dx += Re(cab * (dpoly_dx * xny + poly)) */
    tr = dpxr * xnr - dpxi * xni + pr; ti = dpxr * xni + dpxi * xnr + pi_;
    dfx_dx = dfx_dx + (cr * tr - ci * ti);
    /* dfx_dy += Re(cab * (dpoly_dy * xny + poly * dxny_dy)), dxny_dy = -i cn */
    tr = dpyr * xnr - dpyi * xni + pi_ * cn; ti = dpyr * xni + dpyi * xnr - pr * cn;
    dfx_dy = dfx_dy + (cr * tr - ci * ti);
    /* xny = y + i cn x */
    xnr = y; xni = cn * x;
    qr = pr * xnr - pi_ * xni; qi = pr * xni + pi_ * xnr;
    fy = fy + (cr * qr - ci * qi);
    /* dfy_dx += Re(cab * (dpoly_dx * xny + poly * dxny_dx)), dxny_dx = i cn */
    tr = dpxr * xnr - dpxi * xni - pi_ * cn; ti = dpxr * xni + dpxi * xnr + pr * cn;
    dfy_dx = dfy_dx + (cr * tr - ci * ti);
    /* dfy_dy += Re(cab * (dpoly_dy * xny + poly)) */
    tr = dpyr * xnr - dpyi * xni + pr; ti = dpyr * xni + dpyi * xnr + pi_;
    dfy_dy = dfy_dy + (cr * tr - ci * ti);
  }
  px = orbit->vec[1];
  py = orbit->vec[3];
  denom = (1 - dfx_dx) * (1 - dfy_dy) - dfx_dy * dfy_dx;
  orbit->vec[0] = orbit->vec[0] - fx;
  orbit->vec[1] = px + ((1.0 - dfy_dy - denom) * px + dfy_dx * py) / denom;
  orbit->vec[2] = orbit->vec[2] - fy;
  orbit->vec[3] = py + (dfx_dy * px + (1.0 - dfx_dx - denom) * py) / denom;
  orbit->vec[4] = orbit->vec[4] + (orbit->vec[1] * fx + orbit->vec[3] * fy) / rel_p;
}

/* hwang_bend_edge_kick: fringe_mod.f90:240-418 (time_dir = +1) */
ORACLE_API void oracle_hwang_edge(const bmx_ele_rec *ele, int ref_species, int particle_at, double *vecio,
                                  int species, int direction);

static void hwang_bend_edge_kick(const bmx_ele_rec *ele, int ref_species, int particle_at, o_coord *orb) {
  real e, g_tot, fint_gap, gt, cos_e, sin_e, tan_e, sec_e, v[6], k1_tane, gt2, gs2, c_dir, k1, e_factor,
      fg_factor, dx, dpx, dy, dpy, dz;
  int element_end, i;
  c_dir = rel_tracking_charge_to_mass(orb, ref_species) * ele->orientation * orb->direction;
  element_end = physical_ele_end(particle_at, orb, ele->orientation);
  e_factor = 1 / (1 + orb->vec[5]);
  if (ele->is_on) {
    g_tot = ((real)(ele->g) + ele->dg) * c_dir;
    k1 = ele->k1;
  } else {
    g_tot = 0;
    k1 = 0;
  }
  if (element_end == BMX_ENTRANCE_END) {
    e = ele->e1;
    fint_gap = (real)(ele->fint) * ele->hgap;
  } else {
    e = ele->e2;
    fint_gap = (real)(ele->fintx) * ele->hgapx;
  }
  if (ele->fringe_type == BMX_FRINGE_HARD_EDGE_ONLY || ele->fringe_type == BMX_FRINGE_SAD_FULL) fint_gap = 0;
  cos_e = cos(e);
  sin_e = sin(e);
  tan_e = sin_e / cos_e;
  sec_e = 1 / cos_e;
  gt = g_tot * tan_e;
  gt2 = g_tot * tan_e * tan_e;
  gs2 = g_tot * sec_e * sec_e;
  k1_tane = k1 * tan_e * c_dir;
  fg_factor = 2 * fint_gap * gs2 * g_tot * sec_e * (1 + sin_e * sin_e);
  for (i = 0; i < 6; i++) v[i] = orb->vec[i];
  if (particle_at == O_FIRST_EDGE) { /* entering_element with time_dir = 1 */
    dx = (-gt2 * v[0] * v[0] + gs2 * v[2] * v[2]) * e_factor / 2;
    dpx = (gt * g_tot * (1 + 2 * tan_e * tan_e) * v[2] * v[2] / 2 + gt2 * (v[0] * v[1] - v[2] * v[3]) +
           k1_tane * (v[0] * v[0] - v[2] * v[2])) * e_factor;
    dy = gt2 * v[0] * v[2] * e_factor;
    dpy = (fg_factor * v[2] - gt2 * v[0] * v[3] - (g_tot + gt2) * v[1] * v[2] - 2 * k1_tane * v[0] * v[2]) * e_factor;
    dz = e_factor * e_factor * 0.5 *
         (v[2] * v[2] * fg_factor + v[0] * v[0] * v[0] * (4.0 * k1_tane - gt * gt2) / 6.0 +
          0.5 * v[0] * v[2] * v[2] * (-4.0 * k1_tane + gt * gs2) +
          (v[0] * v[0] * v[1] - 2.0 * v[0] * v[2] * v[3]) * gt2 - v[1] * v[2] * v[2] * gs2);
  } else {
    dx = (gt2 * v[0] * v[0] - gs2 * v[2] * v[2]) * e_factor / 2;
    dpx = (gt2 * (v[2] * v[3] - v[0] * v[1]) + k1_tane * (v[0] * v[0] - v[2] * v[2]) -
           gt * gt2 * (v[0] * v[0] + v[2] * v[2]) / 2) * e_factor;
    dy = -gt2 * v[0] * v[2] * e_factor;
    dpy = (fg_factor * v[2] + gt2 * v[0] * v[3] + (g_tot + gt2) * v[1] * v[2] + (gt * gs2 - 2 * k1_tane) * v[0] * v[2]) *
          e_factor;
    dz = e_factor * e_factor * 0.5 *
         (v[2] * v[2] * fg_factor + v[0] * v[0] * v[0] * (4.0 * k1_tane - gt * gt2) / 6.0 +
          0.5 * v[0] * v[2] * v[2] * (-4.0 * k1_tane + gt * gs2) -
          (v[0] * v[0] * v[1] - 2.0 * v[0] * v[2] * v[3]) * gt2 + v[1] * v[2] * v[2] * gs2);
  }
  orb->vec[0] = v[0] + dx;
  orb->vec[1] = v[1] + (dpx + gt * v[0]);
  orb->vec[2] = v[2] + dy;
  orb->vec[3] = v[3] + (dpy - gt * v[2]);
  orb->vec[4] = v[4] + dz;
}

/* linear_bend_edge_kick: fringe_mod.f90:129-190 */
static void linear_bend_edge_kick(const bmx_ele_rec *ele, int ref_species, int particle_at, o_coord *orb) {
  real e, g_tot, fint, hgap, ht_x, ht_y, cos_e, sin_e, tan_e, c_dir;
  int element_end;
  c_dir = rel_tracking_charge_to_mass(orb, ref_species) * ele->orientation * orb->direction;
  element_end = physical_ele_end(particle_at, orb, ele->orientation);
  g_tot = ele->is_on ? ((real)(ele->g) + ele->dg) * c_dir : (real)(0.0);
  if (element_end == BMX_ENTRANCE_END) { e = ele->e1; fint = ele->fint; hgap = ele->hgap; }
  else { e = ele->e2; fint = ele->fintx; hgap = ele->hgapx; }
  cos_e = cos(e); sin_e = sin(e); tan_e = sin_e / cos_e;
  ht_x = g_tot * tan_e;
  if (fint == 0) ht_y = -ht_x;
  else ht_y = -g_tot * tan(e - 2 * fint * g_tot * hgap * (1 + sin_e * sin_e) / cos_e);
  orb->vec[1] = orb->vec[1] + ht_x * orb->vec[0];
  orb->vec[3] = orb->vec[3] + ht_y * orb->vec[2];
}

/* bend_edge_kick: fringe_mod.f90:39-127 (fringe types hard_edge_only,
 * basic_bend, linear_edge, none; the entry points refuse full / sad_full /
 * soft_edge_only) */
static void bend_edge_kick(const bmx_ele_rec *ele, int ref_species, int particle_at, o_coord *orb) {
  if (!fringe_here(ele, orb, particle_at)) return;
  if (particle_at == O_FIRST_EDGE) hard_multipole_edge_kick(ele, ref_species, particle_at, orb);
  switch (ele->fringe_type) {
    case BMX_FRINGE_BASIC_BEND: case BMX_FRINGE_HARD_EDGE_ONLY:
      hwang_bend_edge_kick(ele, ref_species, particle_at, orb);
      break;
    case BMX_FRINGE_LINEAR_EDGE:
      linear_bend_edge_kick(ele, ref_species, particle_at, orb);
      break;
    default: break;
  }
  if (particle_at == O_SECOND_EDGE) hard_multipole_edge_kick(ele, ref_species, particle_at, orb);
}

/* electric_longitudinal_fringe: apply_element_edge_kick.f90:214-255 (bmad_standard field calc) */
static void electric_longitudinal_fringe(o_coord *orb, const bmx_ele_rec *ele, const real *a_el, const real *b_el,
                                         int ix_elec_max, real at_sign) {
  real ff, xr, xi, xor_, xoi, cr, ci, phi, Er0, Er1, tr, ti;
  int i;
  if (ele->l == 0) return;
  if (ix_elec_max < 0) return;
  ff = at_sign * charge_of(orb->species);
  xr = 1; xi = 0;
  phi = 0; Er0 = 0; Er1 = 0;
  for (i = 0; i <= ix_elec_max; i++) {
    xor_ = xr; xoi = xi;
    tr = xr * orb->vec[0] - xi * orb->vec[2];
    ti = xr * orb->vec[2] + xi * orb->vec[0];
    xr = tr; xi = ti;
    if (a_el[i] == 0 && b_el[i] == 0) continue;
    cr = b_el[i]; ci = -a_el[i];
    phi = phi - (cr * xr - ci * xi) / (i + 1);
    Er0 = Er0 + (cr * xor_ - ci * xoi);
    Er1 = Er1 - (cr * xoi + ci * xor_);
  }
  (void)Er0; (void)Er1;
  apply_energy_kick(-ff * phi, orb);
}

/* apply_element_edge_kick: bmad/low_level/apply_element_edge_kick.f90:37-278
 * (bmad_standard field calc; quadrupole, sextupole, sbend, and the electric
 * longitudinal fringe of any element with electric multipoles) */
static void apply_element_edge_kick(o_coord *orb, int particle_at, const bmx_ele_rec *ele, int ref_species) {
  int physical_end = physical_ele_end(particle_at, orb, ele->orientation), ix_elec_max;
  real at_sign, a_el[BMX_N_POLE], b_el[BMX_N_POLE];
  if (!at_this_ele_end(physical_end, ele->fringe_at)) return;
  at_sign = (particle_at == O_FIRST_EDGE) ? 1.0 : -1.0;
  multipole_ele_to_ab(ele, 0, O_ELECTRIC, 0, &ix_elec_max, a_el, b_el, NULL);
  if (particle_at == O_SECOND_EDGE) electric_longitudinal_fringe(orb, ele, a_el, b_el, ix_elec_max, at_sign);
  switch (ele->key) {
    case BMX_KEY_QUADRUPOLE:
      /* soft_quadrupole_edge_kick is a no-op for the supported fringe types */
      hard_multipole_edge_kick(ele, ref_species, particle_at, orb);
      break;
    case BMX_KEY_SEXTUPOLE:
      hard_multipole_edge_kick(ele, ref_species, particle_at, orb);
      break;
    case BMX_KEY_SBEND:
      bend_edge_kick(ele, ref_species, particle_at, orb);
      break;
    default: break;
  }
  if (particle_at == O_FIRST_EDGE) electric_longitudinal_fringe(orb, ele, a_el, b_el, ix_elec_max, at_sign);
}

/* init_fringe_info: bmad/low_level/init_fringe_info.f90 (has_fringe only) */
static int has_fringe(const bmx_ele_rec *ele) {
  if (ele->key == BMX_KEY_RFCAVITY) return 1;
  if (ele->key == BMX_KEY_DRIFT || ele->key == BMX_KEY_MARKER || ele->key == BMX_KEY_MULTIPOLE ||
      ele->key == BMX_KEY_AB_MULTIPOLE || ele->key == BMX_KEY_PIPE)
    return 0;
  if ((ele->fringe_type == BMX_FRINGE_NONE || ele->fringe_at == BMX_NO_END) && !ele->has_elec_poles) return 0;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* orbit_too_large: bmad/low_level/orbit_too_large.f90:26-120                */
/* ------------------------------------------------------------------------ */
static int orbit_too_large(o_coord *orb, double *unstable) {
  real rel_p, f_unstable, mal = o_com.max_aperture_limit;
  if (orb->vec[0] > mal) { orb->state = BMX_LOST_POS_X; if (unstable) *unstable = RV((fabs(orb->vec[0]) - mal) / mal); return 1; }
  if (-orb->vec[0] > mal) { orb->state = BMX_LOST_NEG_X; if (unstable) *unstable = RV((fabs(orb->vec[0]) - mal) / mal); return 1; }
  if (orb->vec[2] > mal) { orb->state = BMX_LOST_POS_Y; if (unstable) *unstable = RV((fabs(orb->vec[2]) - mal) / mal); return 1; }
  if (-orb->vec[2] > mal) { orb->state = BMX_LOST_NEG_Y; if (unstable) *unstable = RV((fabs(orb->vec[2]) - mal) / mal); return 1; }
  rel_p = 1 + orb->vec[5];
  if (rel_p < 0) { orb->state = BMX_LOST_PZ; if (unstable) *unstable = RV(fabs(rel_p)) + 1e-6; return 1; }
  f_unstable = orb->vec[1] * orb->vec[1] + orb->vec[3] * orb->vec[3] - rel_p * rel_p;
  if (f_unstable > 0) {
    if (unstable) *unstable = RV(sqrt(f_unstable));
    if (fabs(orb->vec[1]) > fabs(orb->vec[3])) orb->state = (orb->vec[1] > 0) ? BMX_LOST_POS_X : BMX_LOST_NEG_X;
    else orb->state = (orb->vec[3] > 0) ? BMX_LOST_POS_Y : BMX_LOST_NEG_Y;
    return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* check_aperture_limit: bmad/low_level/check_aperture_limit.f90:30-372      */
/*   (rectangular / elliptical apertures, collimator phase-space limits;     */
/*    offset_moves_aperture = false)                                         */
/* ------------------------------------------------------------------------ */
static void check_this_limit(o_coord *orb, real coord, real width2, real center, double *uf) {
  real d1, d2;
  if (width2 == 0) return;
  if (coord > center - width2 && coord < center + width2) return;
  orb->state = BMX_LOST;
  d1 = (center - width2) - coord;
  d2 = coord - (center + width2);
  *uf = RV(d1 > d2 ? d1 : d2);
}

static int elliptical_params_setup(real lim1, real lim2, real *pos, real *width2, real *center) {
  if ((lim1 == 0 && lim2 != 0) || (lim1 != 0 && lim2 == 0)) return 1;
  *width2 = (lim2 - lim1) / 2;
  *center = (lim2 + lim1) / 2;
  *pos = *pos - *center;
  if (*width2 == 0) *width2 = o_com.max_aperture_limit;
  if (*width2 < 0) *width2 = 0;
  return 0;
}

static void check_aperture_limit(o_coord *orb, const bmx_ele_rec *ele, int particle_at, double *unstable_out) {
  double uf = 0;
  real x1_lim, x2_lim, y1_lim, y2_lim, x_p, y_p, xw2, yw2, x0, y0, r, f;
  int physical_end;
  if (unstable_out) *unstable_out = 0;
  if (orbit_too_large(orb, unstable_out)) return;
  if (!o_com.aperture_limit_on) return;
  physical_end = physical_ele_end(particle_at, orb, ele->orientation);
  if (!at_this_ele_end(physical_end, ele->aperture_at)) return;
  if (ele->key == BMX_KEY_ECOLLIMATOR || ele->key == BMX_KEY_RCOLLIMATOR) {
    const double cl[8] = {ele->px_aperture_width2, ele->px_aperture_center, ele->py_aperture_width2,
                          ele->py_aperture_center, ele->z_aperture_width2, ele->z_aperture_center,
                          ele->pz_aperture_width2, ele->pz_aperture_center};
    if (ele->key == BMX_KEY_ECOLLIMATOR) {
      xw2 = cl[0]; yw2 = cl[2]; x0 = cl[1]; y0 = cl[3];
      if (xw2 == 0 || yw2 == 0) {
        check_this_limit(orb, orb->vec[1], xw2, x0, &uf);
        check_this_limit(orb, orb->vec[3], yw2, y0, &uf);
      } else {
        r = ((orb->vec[1] - x0) / xw2) * ((orb->vec[1] - x0) / xw2) + ((orb->vec[3] - y0) / yw2) * ((orb->vec[3] - y0) / yw2);
        if (r > 1) { orb->state = BMX_LOST; uf = RV(sqrt(r) - 1); }
      }
      xw2 = cl[4]; yw2 = cl[6]; x0 = cl[5]; y0 = cl[7];
      if (xw2 == 0 || yw2 == 0) {
        check_this_limit(orb, orb->vec[4], xw2, x0, &uf);
        check_this_limit(orb, orb->vec[5], yw2, y0, &uf);
      } else {
        r = ((orb->vec[4] - x0) / xw2) * ((orb->vec[4] - x0) / xw2) + ((orb->vec[5] - y0) / yw2) * ((orb->vec[5] - y0) / yw2);
        if (r > 1) { orb->state = BMX_LOST; uf = RV(sqrt(r) - 1); }
      }
    } else {
      check_this_limit(orb, orb->vec[1], cl[0], cl[1], &uf);
      check_this_limit(orb, orb->vec[3], cl[2], cl[3], &uf);
      check_this_limit(orb, orb->vec[4], cl[4], cl[5], &uf);
      check_this_limit(orb, orb->vec[5], cl[6], cl[7], &uf);
    }
  }
  x_p = orb->vec[0];
  y_p = orb->vec[2];
  x1_lim = -ele->x1_limit;
  x2_lim = ele->x2_limit;
  y1_lim = -ele->y1_limit;
  y2_lim = ele->y2_limit;
  if (ele->aperture_type == BMX_APERTURE_ELLIPTICAL) {
    if (elliptical_params_setup(x1_lim, x2_lim, &x_p, &xw2, &x0)) { if (unstable_out) *unstable_out = uf; return; }
    if (elliptical_params_setup(y1_lim, y2_lim, &y_p, &yw2, &y0)) { if (unstable_out) *unstable_out = uf; return; }
    r = (x_p / xw2) * (x_p / xw2) + (y_p / yw2) * (y_p / yw2);
    if (r < 1) { if (unstable_out) *unstable_out = uf; return; }
    if (fabs(x_p / xw2) > fabs(y_p / yw2)) orb->state = (x_p > 0) ? BMX_LOST_POS_X : BMX_LOST_NEG_X;
    else orb->state = (y_p > 0) ? BMX_LOST_POS_Y : BMX_LOST_NEG_Y;
    uf = RV(sqrt(r) - 1);
  } else {
    if (x1_lim != 0 && x_p < x1_lim) {
      f = fabs((x_p - x1_lim) / (x2_lim - x1_lim));
      if (f > uf) { orb->state = BMX_LOST_NEG_X; uf = RV(f); }
    }
    if (x2_lim != 0 && x_p > x2_lim) {
      f = fabs((x_p - x2_lim) / (x2_lim - x1_lim));
      if (f > uf) { orb->state = BMX_LOST_POS_X; uf = RV(f); }
    }
    if (y1_lim != 0 && y_p < y1_lim) {
      f = fabs((y_p - y1_lim) / (y2_lim - y1_lim));
      if (f > uf) { orb->state = BMX_LOST_NEG_Y; uf = RV(f); }
    }
    if (y2_lim != 0 && y_p > y2_lim) {
      f = fabs((y_p - y2_lim) / (y2_lim - y1_lim));
      if (f > uf) { orb->state = BMX_LOST_POS_Y; uf = RV(f); }
    }
  }
  if (unstable_out) *unstable_out = uf;
}

/* ------------------------------------------------------------------------ */
/* quadrupole: bmad/low_level/track_a_quadrupole.f90:17-148,                 */
/*   quad_mat2_calc.f90:22-81, low_energy_z_correction.f90:22-58             */
/* ------------------------------------------------------------------------ */
static void quad_mat2_calc(real k1, real length, real rel_p, real mat2[2][2], real *zc) {
  real sqrt_k, sk_l, k_l2, cx, sx;
  sqrt_k = sqrt(fabs(k1));
  sk_l = sqrt_k * length;
  if (fabs(sk_l) < 1e-10) {
    k_l2 = k1 * length * length;
    cx = 1 + k_l2 / 2;
    sx = (1 + k_l2 / 6) * length;
  } else if (k1 < 0) {
    cx = cos(sk_l);
    sx = sin(sk_l) / sqrt_k;
  } else {
    cx = cosh(sk_l);
    sx = sinh(sk_l) / sqrt_k;
  }
  mat2[0][0] = cx;
  mat2[0][1] = sx / rel_p;
  mat2[1][0] = k1 * sx * rel_p;
  mat2[1][1] = cx;
  zc[0] = k1 * (-cx * sx + length) / 4;
  zc[1] = -k1 * sx * sx / (2 * rel_p);
  zc[2] = -(cx * sx + length) / (4 * rel_p * rel_p);
}

static real low_energy_z_correction(const o_coord *orbit, const bmx_ele_rec *ele, real ds) {
  real mass = mass_of(orbit->species), p0c = orbit->p0c, e_tot = ele->e_tot, beta0, pz, f, me;
  beta0 = p0c / e_tot;
  pz = orbit->vec[5];
  if (mass * (beta0 * pz) * (beta0 * pz) < 3e-7 * e_tot) {
    me = mass / e_tot;
    f = beta0 * beta0 * (2 * beta0 * beta0 - me * me / 2);
    return ds * pz * (1 - 3 * (pz * beta0 * beta0) / 2 + pz * pz * f) * me * me;
  }
  return ds * (orbit->beta - beta0) / beta0;
}

static void track_a_quadrupole(o_coord *orbit, const bmx_ele_rec *ele, int ref_species) {
  o_coord start_orb = *orbit;
  real an[BMX_N_POLE], bn[BMX_N_POLE], ae[BMX_N_POLE], be[BMX_N_POLE], b1, rel_p, k1, kmx[2][2], kmy[2][2], dzx[3],
      dzy[3], rel_tracking_charge, charge_dir, length, r_step, step_len, v0, v1;
  int ix_mag_max, ix_elec_max, n_step, i, orientation, fringe;

  orientation = ele->orientation * start_orb.direction * start_orb.time_dir;
  rel_tracking_charge = rel_tracking_charge_to_mass(&start_orb, ref_species);
  charge_dir = rel_tracking_charge * orientation;
  length = orbit->time_dir * (real)(ele->l);

  multipole_ele_to_ab(ele, 0, O_MAGNETIC, 1, &ix_mag_max, an, bn, &b1);
  multipole_ele_to_ab(ele, 0, O_ELECTRIC, 0, &ix_elec_max, ae, be, NULL);

  n_step = 1;
  if (ix_mag_max > -1 || ix_elec_max > -1) {
    n_step = nint_r(fabs(length) / ds_step_of(ele));
    if (n_step < 1) n_step = 1;
  }
  r_step = (real)(orbit->time_dir) / n_step;
  step_len = (real)(ele->l) * r_step;
  if (length == 0) n_step = 0;

  if (has_misalign(ele)) offset_particle(ele, ref_species, 1, orbit, 1, 0, -1);

  fringe = has_fringe(ele);
  if (fringe) {
    apply_element_edge_kick(orbit, O_FIRST_EDGE, ele, ref_species);
    if (orbit->state != BMX_ALIVE) return;
  }

  if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, ele, ref_species, orbit, O_MAGNETIC, r_step / 2);
  if (ix_elec_max > -1) ab_multipole_kicks(ae, be, ix_elec_max, ele, ref_species, orbit, O_ELECTRIC, step_len / 2);

  for (i = 1; i <= n_step; i++) {
    rel_p = 1 + orbit->vec[5];
    k1 = charge_dir * b1 / (length * rel_p);
    quad_mat2_calc(-k1, step_len, rel_p, kmx, dzx);
    quad_mat2_calc(k1, step_len, rel_p, kmy, dzy);

    orbit->vec[4] = orbit->vec[4] +
                    (dzx[0] * orbit->vec[0] * orbit->vec[0] + dzx[1] * orbit->vec[0] * orbit->vec[1] +
                     dzx[2] * orbit->vec[1] * orbit->vec[1] + dzy[0] * orbit->vec[2] * orbit->vec[2] +
                     dzy[1] * orbit->vec[2] * orbit->vec[3] + dzy[2] * orbit->vec[3] * orbit->vec[3]);
    v0 = orbit->vec[0]; v1 = orbit->vec[1];
    orbit->vec[0] = kmx[0][0] * v0 + kmx[0][1] * v1;
    orbit->vec[1] = kmx[1][0] * v0 + kmx[1][1] * v1;
    v0 = orbit->vec[2]; v1 = orbit->vec[3];
    orbit->vec[2] = kmy[0][0] * v0 + kmy[0][1] * v1;
    orbit->vec[3] = kmy[1][0] * v0 + kmy[1][1] * v1;

    orbit->vec[4] = orbit->vec[4] + low_energy_z_correction(orbit, ele, step_len);

    if (orbit->direction == -1) orbit->vec[4] = orbit->vec[4] - 2.0 * O_C_LIGHT * orbit->beta * ele->delta_ref_time;

    if (i == n_step) {
      if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, ele, ref_species, orbit, O_MAGNETIC, r_step / 2);
      if (ix_elec_max > -1) ab_multipole_kicks(ae, be, ix_elec_max, ele, ref_species, orbit, O_ELECTRIC, step_len / 2);
    } else {
      if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, ele, ref_species, orbit, O_MAGNETIC, r_step);
      if (ix_elec_max > -1) ab_multipole_kicks(ae, be, ix_elec_max, ele, ref_species, orbit, O_ELECTRIC, step_len);
    }
  }

  if (fringe) {
    apply_element_edge_kick(orbit, O_SECOND_EDGE, ele, ref_species);
    if (orbit->state != BMX_ALIVE) return;
  }

  if (has_misalign(ele)) offset_particle(ele, ref_species, 0, orbit, 1, 0, -1);

  NOCOUNT_BEGIN;
  orbit->t = start_orb.t + orbit->direction * orbit->time_dir * ele->delta_ref_time +
             (start_orb.vec[4] - orbit->vec[4]) / (orbit->beta * O_C_LIGHT);
  NOCOUNT_END;
}

/* ------------------------------------------------------------------------ */
/* thick multipole: bmad/low_level/track_a_thick_multipole.f90:18-133        */
/*   (sextupole, octupole, thick_multipole, kickers, instruments, collimators */
/*    -- elseparator excluded)                                               */
/* ------------------------------------------------------------------------ */
static void track_a_thick_multipole(o_coord *orbit, const bmx_ele_rec *ele, int ref_species) {
  o_coord start_orb = *orbit;
  real an[BMX_N_POLE], bn[BMX_N_POLE], ae[BMX_N_POLE], be[BMX_N_POLE], length, r_step, step_len, ac_amp = 1;
  int ix_mag_max, ix_elec_max, n_step, i, fringe;

  length = (real)(ele->l) * start_orb.time_dir;
  multipole_ele_to_ab(ele, 0, O_MAGNETIC, 1, &ix_mag_max, an, bn, NULL);
  multipole_ele_to_ab(ele, 0, O_ELECTRIC, 0, &ix_elec_max, ae, be, NULL);

  if (ix_mag_max == -1 && ix_elec_max == -1) n_step = 1;
  else {
    n_step = nint_r(fabs(length / ds_step_of(ele)));
    if (n_step < 1) n_step = 1;
  }
  r_step = (real)(orbit->time_dir) / n_step;
  step_len = length / n_step;

  if (has_misalign(ele)) offset_particle(ele, ref_species, 1, orbit, 1, 0, -1);

  fringe = has_fringe(ele);
  if (fringe) {
    apply_element_edge_kick(orbit, O_FIRST_EDGE, ele, ref_species);
    if (orbit->state != BMX_ALIVE) return;
  }

  if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, ele, ref_species, orbit, O_MAGNETIC, ac_amp * r_step / 2);
  if (ix_elec_max > -1) ab_multipole_kicks(ae, be, ix_elec_max, ele, ref_species, orbit, O_ELECTRIC, ac_amp * step_len / 2);

  for (i = 1; i <= n_step; i++) {
    track_a_drift(orbit, step_len, 0, 1, NULL);
    if (i == n_step) {
      if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, ele, ref_species, orbit, O_MAGNETIC, ac_amp * r_step / 2);
      if (ix_elec_max > -1) ab_multipole_kicks(ae, be, ix_elec_max, ele, ref_species, orbit, O_ELECTRIC, ac_amp * step_len / 2);
    } else {
      if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, ele, ref_species, orbit, O_MAGNETIC, ac_amp * r_step);
      if (ix_elec_max > -1) ab_multipole_kicks(ae, be, ix_elec_max, ele, ref_species, orbit, O_ELECTRIC, ac_amp * step_len);
    }
  }

  if (fringe) {
    apply_element_edge_kick(orbit, O_SECOND_EDGE, ele, ref_species);
    if (orbit->state != BMX_ALIVE) return;
  }

  if (has_misalign(ele)) offset_particle(ele, ref_species, 0, orbit, 1, 0, -1);

  NOCOUNT_BEGIN;
  orbit->t = start_orb.t + orbit->direction * orbit->time_dir * ele->delta_ref_time +
             (start_orb.vec[4] - orbit->vec[4]) / (orbit->beta * O_C_LIGHT);
  NOCOUNT_END;
}

/* ------------------------------------------------------------------------ */
/* sbend: bmad/low_level/track_a_bend.f90:18-333, :338-471, :500-659         */
/* ------------------------------------------------------------------------ */
static void sbend_body_with_k1_map(const bmx_ele_rec *ele, real dg, real b1, int ref_species, int n_step, o_coord *orbit) {
  real g, length, k1, k_x, x_c, om_x, om_y, tau_x, tau_y, arg, s_x, c_x, s_y, c_y, r[6], z0, z1, z2, z11, z12, z22,
      z33, z34, z44, g_tot, rel_p, rel_p2, rel_charge_dir;
  int i;
  rel_charge_dir = rel_tracking_charge_to_mass(orbit, ref_species) * ele->orientation * orbit->direction;
  g = ele->g;
  length = orbit->time_dir * (real)(ele->l) / n_step;
  k1 = b1 / ele->l;
  g_tot = (g + dg) * rel_charge_dir;
  rel_p = 1 + orbit->vec[5];
  rel_p2 = rel_p * rel_p;
  k_x = k1 + g * g_tot;
  x_c = (g * rel_p - g_tot) / k_x;
  om_x = sqrt(fabs(k_x) / rel_p);
  om_y = sqrt(fabs(k1) / rel_p);
  tau_x = (k_x >= 0) ? -1.0 : 1.0;
  tau_y = (k1 >= 0) ? 1.0 : -1.0;
  arg = om_x * length;
  if (arg < 1e-6) {
    s_x = (1 + tau_x * arg * arg / 6) * length;
    c_x = 1 + tau_x * arg * arg / 2;
    z2 = g * length * length / (2 * rel_p);
  } else if (k_x > 0) {
    s_x = sin(arg) / om_x;
    c_x = cos(arg);
    z2 = tau_x * g * (1 - c_x) / (rel_p * om_x * om_x);
  } else {
    s_x = sinh(arg) / om_x;
    c_x = cosh(arg);
    z2 = tau_x * g * (1 - c_x) / (rel_p * om_x * om_x);
  }
  arg = om_y * length;
  if (arg < 1e-6) {
    s_y = (1 + tau_y * arg * arg / 6) * length;
    c_y = 1 + tau_y * arg * arg / 2;
  } else if (k1 < 0) {
    s_y = sin(om_y * length) / om_y;
    c_y = cos(om_y * length);
  } else {
    s_y = sinh(om_y * length) / om_y;
    c_y = cosh(om_y * length);
  }
  for (i = 0; i < 6; i++) r[i] = orbit->vec[i];
  r[0] = r[0] - x_c;
  z0 = -g * x_c * length;
  z1 = -g * s_x;
  z11 = tau_x * om_x * om_x * (length - c_x * s_x) / 4;
  z12 = -tau_x * om_x * om_x * s_x * s_x / (2 * rel_p);
  z22 = -(length + c_x * s_x) / (4 * rel_p2);
  z33 = tau_y * om_y * om_y * (length - c_y * s_y) / 4;
  z34 = -tau_y * om_y * om_y * s_y * s_y / (2 * rel_p);
  z44 = -(length + c_y * s_y) / (4 * rel_p2);
  orbit->vec[0] = c_x * r[0] + s_x * r[1] / rel_p + x_c;
  orbit->vec[1] = tau_x * om_x * om_x * rel_p * s_x * r[0] + c_x * r[1];
  orbit->vec[2] = c_y * r[2] + s_y * r[3] / rel_p;
  orbit->vec[3] = tau_y * om_y * om_y * rel_p * s_y * r[2] + c_y * r[3];
  orbit->vec[4] = r[4] + orbit->direction * ele->orientation * z0 + z1 * r[0] + z2 * r[1] + z11 * r[0] * r[0] +
                  z12 * r[0] * r[1] + z22 * r[1] * r[1] + z33 * r[2] * r[2] + z34 * r[2] * r[3] + z44 * r[3] * r[3];
}

/* apply_multipole_kicks (contained in track_a_bend.f90:338-471), exact_multipoles = off */
static void bend_apply_multipole_kicks(const bmx_ele_rec *ele, int ref_species, o_coord *orbit, real coef,
                                       real step_len, real r_step, int ix_mag_max, const real *an, const real *bn,
                                       int ix_elec_max, const real *ae, const real *be, real c_dir) {
  real B0 = 0, B1 = 0, E0 = 0, E1 = 0, f_p0c, kx, ky, f_coef, rel_p0, ps2, ps, Kx, Ky, alpha, charge, mc;
  real f;
  o_coord orb0;
  int i, m;
  charge = charge_of(orbit->species);
  f_p0c = r_step * ele->p0c / (O_C_LIGHT * charge_of(ref_species));
  for (i = 0; i <= ix_mag_max; i++) {
    if (an[i] == 0 && bn[i] == 0) continue;
    ab_multipole_kick(an[i], bn[i], i, ref_species, 0, orbit, &kx, &ky, O_MAGNETIC, 0, 0.0);
    B0 = B0 + f_p0c * ky;
    B1 = B1 - f_p0c * kx;
  }
  for (i = 0; i <= ix_elec_max; i++) {
    /* elec_multipole_field: bmad/code/elec_multipole_field.f90:20-70 */
    if (ae[i] == 0 && be[i] == 0) continue;
    for (m = 0; m <= i; m += 2) {
      f = c_multi(i, m, 1) * mexp(orbit->vec[0], i - m) * mexp(orbit->vec[2], m);
      E0 = E0 - be[i] * f;
      E1 = E1 - ae[i] * f;
    }
    for (m = 1; m <= i; m += 2) {
      f = c_multi(i, m, 1) * mexp(orbit->vec[0], i - m) * mexp(orbit->vec[2], m);
      E0 = E0 + ae[i] * f;
      E1 = E1 - be[i] * f;
    }
  }
  if (ix_mag_max > -1) {
    orb0 = *orbit;
    f_coef = coef * c_dir * (1 + (real)(ele->g) * orbit->vec[0]) * O_C_LIGHT / orb0.p0c;
    orbit->vec[1] = orbit->vec[1] - f_coef * B1;
    orbit->vec[3] = orbit->vec[3] + f_coef * B0;
  }
  if (ix_elec_max > -1) {
    orb0 = *orbit;
    rel_p0 = 1 + orb0.vec[5];
    ps2 = rel_p0 * rel_p0 - orb0.vec[1] * orb0.vec[1] - orb0.vec[3] * orb0.vec[3];
    ps = sqrt(ps2) / rel_p0;
    f_coef = step_len * coef * (1 + (real)(ele->g) * orb0.vec[0]) * charge / (ps * orb0.beta * orb0.p0c);
    Kx = f_coef * E0;
    Ky = f_coef * E1;
    alpha = (Kx * (2 * orb0.vec[1] + Kx) + Ky * (2 * orb0.vec[3] + Ky)) / (rel_p0 * rel_p0);
    orbit->vec[1] = orb0.vec[1] + Kx;
    orbit->vec[3] = orb0.vec[3] + Ky;
    orbit->vec[5] = orb0.vec[5] + rel_p0 * sqrt_one(alpha);
    mc = mass_of(orbit->species) / orbit->p0c;
    orbit->beta = (1 + orbit->vec[5]) / sqrt(ps2 + orbit->vec[1] * orbit->vec[1] + orbit->vec[3] * orbit->vec[3] + mc * mc);
    orbit->vec[4] = orb0.vec[4] * orbit->beta / orb0.beta;
  }
}

static void track_a_bend(o_coord *orbit, const bmx_ele_rec *ele, int ref_species) {
  o_coord start_orb = *orbit;
  real an[BMX_N_POLE], bn[BMX_N_POLE], ae[BMX_N_POLE], be[BMX_N_POLE], rel_charge_dir, c_dir, b1, length, g, dg,
      g_tot, r_step, step_len, angle, x, px, y, py, z, pz, rel_p, rel_p2, ll, cos_a, sin_a, sinc_a, cosc_a, gam2, m56,
      beta_ref, pt, g_p, phi_1, cos_plus, sin_plus, alpha, r, rad, denom, xi, L_u, L_v, L_c, angle_p, L_p, mc2;
  int ix_mag_max, ix_elec_max, n_step, n;

  if (has_misalign(ele)) offset_particle(ele, ref_species, 1, orbit, 1, 0, -1);

  rel_charge_dir = ele->orientation * orbit->direction * rel_tracking_charge_to_mass(orbit, ref_species);
  c_dir = ele->orientation * orbit->direction * charge_of(orbit->species);

  apply_element_edge_kick(orbit, O_FIRST_EDGE, ele, ref_species);
  if (orbit_too_large(orbit, NULL)) return;

  multipole_ele_to_ab(ele, 0, O_MAGNETIC, 1, &ix_mag_max, an, bn, &b1);
  b1 = b1 * rel_charge_dir;
  if (fabs(b1) < 1e-10) {
    bn[1] = b1;
    b1 = 0;
  }
  multipole_ele_to_ab(ele, 0, O_ELECTRIC, 0, &ix_elec_max, ae, be, NULL);

  length = orbit->time_dir * (real)(ele->l);
  g = ele->g;
  if (length == 0) {
    dg = 0;
  } else {
    dg = bn[0] / ele->l;
    bn[0] = 0;
  }
  g_tot = (g + dg) * rel_charge_dir;
  if (!ele->is_on) {
    dg = -g;
    g_tot = 0;
    b1 = 0;
  }

  n_step = 1;
  if (ix_mag_max > -1 || ix_elec_max > -1) {
    n_step = nint_r(fabs(length) / ds_step_of(ele));
    if (n_step < 1) n_step = 1;
  }
  r_step = (real)(orbit->time_dir) / n_step;
  step_len = length / n_step;
  angle = g * step_len;

  if (ix_mag_max > -1 || ix_elec_max > -1)
    bend_apply_multipole_kicks(ele, ref_species, orbit, 0.5, step_len, r_step, ix_mag_max, an, bn, ix_elec_max, ae, be, c_dir);

  for (n = 1; n <= n_step; n++) {
    if (b1 != 0) {
      sbend_body_with_k1_map(ele, dg, b1, ref_species, n_step, orbit);
      orbit->vec[4] = orbit->vec[4] + low_energy_z_correction(orbit, ele, step_len);
    } else if ((g == 0 && dg == 0) || step_len == 0) {
      track_a_drift(orbit, step_len, 0, 1, NULL);
    } else {
      x = orbit->vec[0];
      px = orbit->vec[1];
      y = orbit->vec[2];
      py = orbit->vec[3];
      z = orbit->vec[4];
      pz = orbit->vec[5];
      rel_p = 1 + pz;
      rel_p2 = rel_p * rel_p;
      if (dg == 0 && fabs(x * g) < 1e-9 && fabs(px) < 1e-9 && fabs(py) < 1e-9 && fabs(pz) < 1e-9) {
        ll = step_len;
        cos_a = cos(angle);
        sin_a = sin(angle);
        sinc_a = sinc(angle);
        cosc_a = cosc(angle);
        mc2 = mass_of(orbit->species);
        gam2 = mc2 * mc2 / ((rel_p * orbit->p0c) * (rel_p * orbit->p0c) + mc2 * mc2);
        m56 = ll * (gam2 - (g * ll) * (g * ll) * sincc(angle));
        orbit->vec[0] = cos_a * x + ll * sinc_a * px + g * ll * ll * cosc_a * pz;
        orbit->vec[1] = -g * sin_a * x + cos_a * px + g * ll * sinc_a * pz;
        orbit->vec[2] = y + ll * py;
        orbit->vec[4] = -g * ll * sinc_a * x - g * ll * ll * cosc_a * px + z + m56 * pz;
      } else {
        beta_ref = (real)(ele->p0c) / ele->e_tot;
        sinc_a = sinc(angle);
        pt = sqrt(rel_p2 - py * py);
        if (fabs(px) > pt) {
          orbit->state = BMX_LOST_PZ;
          return;
        }
        g_p = g_tot / pt;
        phi_1 = asin(px / pt);
        cos_a = cos(angle);
        sin_a = sin(angle);
        cosc_a = cosc(angle);
        cos_plus = cos(angle + phi_1);
        sin_plus = sin(angle + phi_1);
        alpha = 2 * (1 + g * x) * sin_plus * step_len * sinc_a -
                g_p * ((1 + g * x) * step_len * sinc_a) * ((1 + g * x) * step_len * sinc_a);
        r = cos_plus * cos_plus + g_p * alpha;
        if (r < 0 || (fabs(g_p) < 1e-5 && fabs(cos_plus) < 1e-5)) {
          orbit->state = BMX_LOST;
          orbit->vec[0] = 2 * o_com.max_aperture_limit;
          return;
        }
        rad = sqrt(r);
        if (cos_plus > 0) {
          denom = rad + cos_plus;
          xi = alpha / denom;
        } else {
          denom = rad - cos_plus;
          xi = denom / g_p;
        }
        orbit->vec[0] = x * cos_a - step_len * step_len * g * cosc_a + xi;
        if (fabs(orbit->vec[0]) > o_com.max_aperture_limit) {
          orbit->state = BMX_LOST;
          return;
        }
        L_u = xi;
        L_v = -orbit->time_dir * (step_len * sinc_a + x * sin_a);
        L_c = sqrt(L_v * L_v + L_u * L_u);
        angle_p = 2 * (angle + phi_1 - orbit->time_dir * atan2(L_u, -L_v));
        L_p = orbit->time_dir * L_c / sinc(angle_p / 2);
        orbit->vec[1] = pt * sin(phi_1 + angle - angle_p);
        orbit->vec[2] = y + py * L_p / pt;
        orbit->vec[4] = z + orbit->beta * step_len / beta_ref - rel_p * L_p / pt;
      }
    }
    if (ix_mag_max > -1 || ix_elec_max > -1) {
      if (n == n_step)
        bend_apply_multipole_kicks(ele, ref_species, orbit, 0.5, step_len, r_step, ix_mag_max, an, bn, ix_elec_max, ae, be, c_dir);
      else
        bend_apply_multipole_kicks(ele, ref_species, orbit, 1.0, step_len, r_step, ix_mag_max, an, bn, ix_elec_max, ae, be, c_dir);
    }
  }

  if (orbit_too_large(orbit, NULL)) return;
  apply_element_edge_kick(orbit, O_SECOND_EDGE, ele, ref_species);

  if (has_misalign(ele)) offset_particle(ele, ref_species, 0, orbit, 1, 0, -1);

  NOCOUNT_BEGIN;
  orbit->t = start_orb.t + orbit->direction * orbit->time_dir * ele->delta_ref_time +
             (start_orb.vec[4] - orbit->vec[4]) / (orbit->beta * O_C_LIGHT);
  NOCOUNT_END;
}

/* ------------------------------------------------------------------------ */
/* rfcavity: bmad/low_level/track_a_rfcavity.f90:17-146                      */
/*   particle_rf_time (bmad/code/particle_rf_time.f90:32-120),               */
/*   rf_coupler_kick (bmad/low_level/rf_coupler_kick.f90:20-106)             */
/* ------------------------------------------------------------------------ */
static void rf_coupler_kick(const bmx_ele_rec *ele, int particle_at, real phase, o_coord *orbit) {
  int physical_end = physical_ele_end(particle_at, orbit, ele->orientation);
  real ph, dp_coef, dp_x, dp_y, f2, dE, gradient;
  if (!at_this_ele_end(physical_end, ele->coupler_at)) return;
  if (ele->coupler_strength == 0) return;
  ph = phase;
  ph = O_PI / 2 - ph;
  ph = ph + O_TWOPI * ele->coupler_phase;
  gradient = (ele->l == 0) ? (real)(0.0) : (real)(ele->voltage) / ele->l;
  dp_coef = gradient * ele->coupler_strength;
  dp_x = dp_coef * cos(O_TWOPI * ele->coupler_angle);
  dp_y = dp_coef * sin(O_TWOPI * ele->coupler_angle);
  if (ele->coupler_at == BMX_BOTH_ENDS) {
    dp_x = dp_x / 2;
    dp_y = dp_y / 2;
  }
  orbit->vec[1] = orbit->vec[1] + orbit->time_dir * dp_x * cos(ph) / orbit->p0c;
  orbit->vec[3] = orbit->vec[3] + orbit->time_dir * dp_y * cos(ph) / orbit->p0c;
  f2 = orbit->time_dir * sin(ph) * O_TWOPI * ele->rf_frequency / O_C_LIGHT;
  dE = (dp_x * orbit->vec[0] + dp_y * orbit->vec[2]) * f2;
  apply_energy_kick(dE, orbit);
}

static void track_a_rfcavity(o_coord *orbit, const bmx_ele_rec *ele, int ref_species) {
  real an[BMX_N_POLE], bn[BMX_N_POLE], ae[BMX_N_POLE], be[BMX_N_POLE], length, r_step, step_len, beta_ref, dt_ref,
      charge_dir, voltage, phase0, phase, time, dE, z;
  int ix_mag_max, ix_elec_max, n_step, n_slice, i;
  if (ele->rf_frequency == 0 && ele->voltage != 0) {
    orbit->state = BMX_LOST;
    return;
  }
  length = (real)(ele->l) * orbit->time_dir;
  n_step = 1;
  r_step = (real)(orbit->time_dir) / n_step;
  step_len = length / n_step;
  beta_ref = (real)(ele->p0c) / ele->e_tot;
  n_slice = nint_r(fabs(length) / ds_step_of(ele));
  if (n_slice < 1) n_slice = 1;
  dt_ref = length / (O_C_LIGHT * beta_ref);
  charge_dir = rel_tracking_charge_to_mass(orbit, ref_species) * ele->orientation;

  multipole_ele_to_ab(ele, 0, O_MAGNETIC, 1, &ix_mag_max, an, bn, NULL);
  multipole_ele_to_ab(ele, 0, O_ELECTRIC, 0, &ix_elec_max, ae, be, NULL);

  offset_particle(ele, ref_species, 1, orbit, 1, 1, -1);

  if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, ele, ref_species, orbit, O_MAGNETIC, 0.5 * r_step);
  if (ix_elec_max > -1) ab_multipole_kicks(ae, be, ix_elec_max, ele, ref_species, orbit, O_ELECTRIC, 0.5 * step_len);

  voltage = (ele->is_on ? (real)(ele->voltage) : (real)(0.0)) * charge_dir;
  /* particle_rf_time */
  /* absolute_time_ref_shift = true: time relative to the element's ref_time_start */
  if (o_com.absolute_time_tracking) {
    /* t - ref_time_start and its modulo2 in the time's own precision */
    treal tl = orbit->t - ((treal)(ele->ref_time) - (treal)(ele->delta_ref_time));
    if (ele->rf_frequency != 0) {
      treal amp = (treal)(0.5) / (treal)(ele->rf_frequency), per = 2 * amp;
      if (!(tl >= -amp && tl < amp)) {
        tl = tl - per * floor_t(tl / per);
        if (tl >= amp) tl = tl - per;
      }
    }
    time = (real)(tl);
  } else {
    time = -orbit->vec[4] / (orbit->beta * O_C_LIGHT);
    if (ele->rf_frequency != 0) time = modulo2_qp(time, 0.5 / (real)(ele->rf_frequency));
  }
  /* rf_ref_time_offset = 0 for a non-slave element */
  phase0 = O_TWOPI * (ele->phi0 - time * ele->rf_frequency);
  if (!o_com.absolute_time_tracking) phase0 = phase0 + O_TWOPI * ele->phi0_multipass;
  if (o_com.absolute_time_tracking && ele->orientation * orbit->time_dir * orbit->direction == -1)
    phase0 = phase0 - O_TWOPI * ele->rf_frequency * dt_ref;
  phase = modulo2(phase0, O_PI);

  rf_coupler_kick(ele, O_FIRST_EDGE, phase, orbit);

  for (i = 0; i <= n_slice; i++) {
    dE = orbit->time_dir * voltage * sin(phase) / n_slice;
    if (i == 0 || i == n_slice) dE = dE / 2;
    apply_energy_kick(dE, orbit);
    if (orbit->vec[5] == -1) {
      orbit->state = BMX_LOST_PZ;
      return;
    }
    if (i != n_slice) {
      z = orbit->vec[4];
      track_a_drift(orbit, length / n_slice, 0, 1, NULL);
      phase = phase + O_TWOPI * ele->rf_frequency * (orbit->vec[4] - z) / (O_C_LIGHT * orbit->beta);
    }
  }

  rf_coupler_kick(ele, O_SECOND_EDGE, phase, orbit);
  orbit->phase1 = modulo2(phase, 0.5);

  if (ix_elec_max > -1) ab_multipole_kicks(ae, be, ix_elec_max, ele, ref_species, orbit, O_ELECTRIC, 0.5 * step_len);
  if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, ele, ref_species, orbit, O_MAGNETIC, 0.5 * r_step);

  offset_particle(ele, ref_species, 0, orbit, 1, 1, -1);
}

/* ------------------------------------------------------------------------ */
/* track1_bmad: bmad/low_level/track1_bmad.f90:22-283                        */
/* ------------------------------------------------------------------------ */
static int track1_bmad(o_coord *orbit, const bmx_ele_rec *ele, int ref_species) {
  int key = ele->key, ix_pole_max, n;
  real knl[BMX_N_POLE], tilt[BMX_N_POLE];
  if (!ele->is_on) {
    switch (key) {
      case BMX_KEY_AB_MULTIPOLE: case BMX_KEY_MULTIPOLE: return 0;
      case BMX_KEY_SBEND: break;
      default: key = BMX_KEY_DRIFT;
    }
  }
  switch (key) {
    case BMX_KEY_RCOLLIMATOR: case BMX_KEY_ECOLLIMATOR: case BMX_KEY_MONITOR: case BMX_KEY_INSTRUMENT:
    case BMX_KEY_PIPE: case BMX_KEY_KICKER: case BMX_KEY_HKICKER: case BMX_KEY_VKICKER:
    case BMX_KEY_OCTUPOLE: case BMX_KEY_THICK_MULTIPOLE: case BMX_KEY_SEXTUPOLE:
      track_a_thick_multipole(orbit, ele, ref_species);
      break;
    case BMX_KEY_DRIFT:
      track_a_drift(orbit, (real)(ele->l) * orbit->time_dir, 0, 1, NULL);
      break;
    case BMX_KEY_MARKER:
      orbit->t = orbit->t + ele->delta_ref_time;
      break;
    case BMX_KEY_MULTIPOLE: case BMX_KEY_AB_MULTIPOLE:
      if (has_misalign(ele)) offset_particle(ele, ref_species, 1, orbit, 0, 1, -1);
      multipole_ele_to_kt(ele, 1, &ix_pole_max, knl, tilt);
      if (ix_pole_max > -1) {
        for (n = 0; n < BMX_N_POLE; n++) {
          if (knl[n] == 0) continue;
          multipole_kick(knl[n], tilt[n], n, ref_species, ele->orientation, orbit, key == BMX_KEY_MULTIPOLE);
        }
      }
      if (has_misalign(ele)) offset_particle(ele, ref_species, 0, orbit, 0, 1, -1);
      break;
    case BMX_KEY_QUADRUPOLE:
      track_a_quadrupole(orbit, ele, ref_species);
      break;
    case BMX_KEY_RFCAVITY:
      track_a_rfcavity(orbit, ele, ref_species);
      break;
    case BMX_KEY_SBEND:
      track_a_bend(orbit, ele, ref_species);
      break;
    default:
      return -1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* radiation kicks: track1_radiation, bmad/modules/radiation_mod.f90:62-141  */
/* ------------------------------------------------------------------------ */
/* bmad_com%radiation_damping_on, %radiation_fluctuations_on,
 * %radiation_zero_average, %synch_rad_scale (bmad_struct.f90:2312-2330), the
 * maps of every element of the list oracle_track_bunch tracks (ele%rad_map
 * rm0 / rm1 in the bmx_rad_map layout, 84 doubles each, two per element) and
 * the fluctuation stream (Philox key and call number, as libbmx draws it). */
typedef struct o_rad_com_t {
  int damping_on, fluctuations_on, zero_average;
  double synch_rad_scale;
  uint32_t key0, key1, call;
  const double *maps;
} o_rad_com_t;
static o_rad_com_t o_rad = {0, 0, 0, 1.0, 0, 0, 0, NULL};

enum { O_RAD_MAP = 84, O_RAD_REF = 0, O_RAD_DAMP = 6, O_RAD_VEC = 42, O_RAD_STOC = 48 };

ORACLE_API void oracle_set_radiation(const double *maps, int damping_on, int fluctuations_on, int zero_average,
                                     double synch_rad_scale, uint64_t seed, uint64_t call) {
  o_rad.maps = maps;
  o_rad.damping_on = damping_on;
  o_rad.fluctuations_on = fluctuations_on;
  o_rad.zero_average = zero_average;
  o_rad.synch_rad_scale = synch_rad_scale;
  o_rad.key0 = (uint32_t)(seed & 0xffffffffu);
  o_rad.key1 = (uint32_t)(seed >> 32);
  o_rad.call = (uint32_t)call;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): the counter-based
 * generator libbmx draws its Gaussian deviates from.  The reference draws
 * from ran_gauss (sim_utils/math/random_mod.f90), a different stream, so
 * fluctuating tracks are compared with the reference in distribution and
 * with libbmx deviate by deviate. */
ORACLE_API void oracle_philox4x32_10(uint32_t *c, uint32_t k0, uint32_t k1) {
  int r;
  for (r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

/* Six standard normal deviates for particle pid at element ie, edge (0
 * entrance, 1 exit), turn: Box-Muller on three Philox blocks, u1 in (0, 1]
 * and u2 in [0, 1) from 53 bits each. */
ORACLE_API void oracle_rad_gauss6(uint32_t pid, int ie, int edge, uint32_t turn, uint32_t call, uint32_t k0,
                                  uint32_t k1, double *g) {
  int q;
  for (q = 0; q < 3; q++) {
    uint32_t c[4];
    double u1, u2, rad;
    c[0] = pid;
    c[1] = (uint32_t)(ie * 2 + edge);
    c[2] = turn;
    c[3] = call * 4u + (uint32_t)q;
    oracle_philox4x32_10(c, k0, k1);
    u1 = ((double)((((uint64_t)c[0]) << 21) | (c[1] >> 11)) + 1.0) * 0x1p-53;
    u2 = (double)((((uint64_t)c[2]) << 21) | (c[3] >> 11)) * 0x1p-53;
    rad = sqrt(-2.0 * log(u1));
    g[2 * q] = rad * cos(O_TWOPI * u2);
    g[2 * q + 1] = rad * sin(O_TWOPI * u2);
  }
}

/* Where a radiation step draws its deviates: element index, particle index, turn. */
typedef struct o_rad_at {
  int ie;
  uint32_t pid, turn;
} o_rad_at;

static void track1_radiation(o_coord *orbit, const bmx_ele_rec *ele, int edge, const o_rad_at *at) {
  const double *rm;
  real d[6], a, rr;
  double g[6];
  int i, j;
  if (!o_rad.damping_on && !o_rad.fluctuations_on) return;
  if (ele->l == 0) return;
  switch (ele->key) {
    case BMX_KEY_DRIFT: case BMX_KEY_PIPE: case BMX_KEY_MULTIPOLE: case BMX_KEY_AB_MULTIPOLE: case BMX_KEY_MARKER:
      return;
    default: break;
  }
  rm = o_rad.maps + ((size_t)at->ie * 2 + (edge == O_FIRST_EDGE ? 0 : 1)) * O_RAD_MAP;
  if (o_rad.damping_on) {
    rr = orbit->time_dir * o_rad.synch_rad_scale;
    for (j = 0; j < 6; j++) d[j] = orbit->vec[j] - rm[O_RAD_REF + j];
    for (i = 0; i < 6; i++) {
      a = rm[O_RAD_DAMP + i * 6] * d[0];
      for (j = 1; j < 6; j++) a = a + rm[O_RAD_DAMP + i * 6 + j] * d[j];
      orbit->vec[i] = orbit->vec[i] + rr * a;
    }
    if (!o_rad.zero_average)
      for (i = 0; i < 6; i++) orbit->vec[i] = orbit->vec[i] + rr * rm[O_RAD_VEC + i];
  }
  if (o_rad.fluctuations_on) {
    oracle_rad_gauss6(at->pid, at->ie, edge == O_FIRST_EDGE ? 0 : 1, at->turn, o_rad.call, o_rad.key0, o_rad.key1,
                      g);
    for (i = 0; i < 6; i++) {
      a = rm[O_RAD_STOC + i * 6] * g[0];
      for (j = 1; j < 6; j++) a = a + rm[O_RAD_STOC + i * 6 + j] * g[j];
      orbit->vec[i] = orbit->vec[i] + o_rad.synch_rad_scale * a;
    }
  }
  if (orbit->vec[5] < -1) orbit->state = BMX_LOST_PZ;
}

/* ------------------------------------------------------------------------ */
/* track1: bmad/code/track1.f90:43-342 (bmad_standard, no spin); radiation   */
/* (:198-201, :316-319) when `rad` is given                                  */
/* ------------------------------------------------------------------------ */
static int track1(o_coord *start_orb, const bmx_ele_rec *ele, int ref_species, o_coord *end_orb,
                  const o_rad_at *rad) {
  o_coord s2 = *start_orb;
  int err;
  int rad_on = rad && (o_rad.damping_on || o_rad.fluctuations_on) && ele->is_on;
  if (s2.direction * s2.time_dir == 1) s2.location = O_UPSTREAM;
  else s2.location = O_DOWNSTREAM;
  if (s2.state != BMX_ALIVE) {
    *end_orb = s2;
    return 0;
  }
  check_aperture_limit(&s2, ele, O_FIRST_EDGE, NULL);
  if (rad_on) track1_radiation(&s2, ele, O_FIRST_EDGE, rad);
  *end_orb = s2;
  if (end_orb->state != BMX_ALIVE) return 0;
  err = track1_bmad(end_orb, ele, ref_species);
  if (err) return err;
  if (orbit_too_large(end_orb, NULL)) return 0;
  if (end_orb->state != BMX_ALIVE) end_orb->location = O_INSIDE;
  else if (end_orb->direction * end_orb->time_dir == 1) end_orb->location = O_DOWNSTREAM;
  else end_orb->location = O_UPSTREAM;
  if (end_orb->state != BMX_ALIVE) return 0;
  if (rad_on) track1_radiation(end_orb, ele, O_SECOND_EDGE, rad);
  check_aperture_limit(end_orb, ele, O_SECOND_EDGE, NULL);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* init_coord: bmad/low_level/init_coord.f90:42-234 (beta and t)             */
/* ------------------------------------------------------------------------ */
static void init_coord(o_coord *orb, const double *vec, const bmx_ele_rec *ele, int species, int direction, int at_exit) {
  int i;
  real p0c = at_exit ? ele->p0c : ele->p0c_start;
  real e_tot = at_exit ? ele->e_tot : ele->e_tot_start;
  real mc2 = mass_of(species), pc;
  for (i = 0; i < 6; i++) orb->vec[i] = vec[i];
  orb->species = species;
  orb->direction = direction;
  orb->time_dir = 1;
  orb->state = BMX_ALIVE;
  orb->p0c = p0c;
  orb->s = 0;
  orb->phase1 = 0;
  orb->location = at_exit ? O_DOWNSTREAM : O_UPSTREAM;
  if (orb->vec[5] == 0) {
    orb->beta = p0c / e_tot;
  } else {
    pc = p0c * (1 + orb->vec[5]);
    orb->beta = pc / sqrt(pc * pc + mc2 * mc2); /* convert_pc_to */
  }
  orb->t = (at_exit ? (real)(ele->ref_time) : (real)(ele->ref_time) - ele->delta_ref_time) -
           orb->vec[4] / (orb->beta * O_C_LIGHT);
}

/* ======================================================================== */
/* exported entry points (ctypes from tests/ and bench.py)                   */
/* ======================================================================== */

/* Elements this restatement does not model: an sbend edge of fringe_type
 * full (exact_bend_edge_kick, fringe_mod.f90:1465-1686), sad_full or
 * soft_edge_only (sad_soft_bend_edge_kick, :892-996).  The entry points
 * refuse them (-2) rather than track past the missing edge map. */
static int unmodelled(const bmx_ele_rec *ele) {
  if (ele->key != BMX_KEY_SBEND || ele->fringe_at == BMX_NO_END) return 0;
  return ele->fringe_type == BMX_FRINGE_FULL || ele->fringe_type == BMX_FRINGE_SAD_FULL ||
         ele->fringe_type == BMX_FRINGE_SOFT_EDGE_ONLY;
}

static int any_unmodelled(const bmx_ele_rec *eles, int n_ele) {
  int ie;
  for (ie = 0; ie < n_ele; ie++)
    if (unmodelled(&eles[ie])) return 1;
  return 0;
}

/* One particle through one element (the tracking_method_test harness,
 * regression_tests/tracking_method_test/tracking_method_test.f90:188-199).
 * io: in = vec[6]; out = vec[6], t, beta, state.  Returns 0 or <0 on error. */
ORACLE_API int oracle_track1(const bmx_ele_rec *ele, int ref_species, int species, int direction,
                             const double *vec_in, double *vec_out, double *t_out, double *beta_out,
                             int *state_out, double t0_override) {
  o_coord a, b;
  int err, i;
  if (unmodelled(ele)) return -2;
  init_coord(&a, vec_in, ele, species, direction, direction == 1 ? 0 : 1);
  if (t0_override == t0_override && t0_override != 0) a.t = t0_override; /* NaN => keep */
  err = track1(&a, ele, ref_species, &b, NULL);
  for (i = 0; i < 6; i++) vec_out[i] = RV(b.vec[i]);
  if (t_out) { t_out[0] = RV(a.t); t_out[1] = RV(b.t); }
  if (beta_out) { beta_out[0] = RV(a.beta); beta_out[1] = RV(b.beta); }
  if (state_out) *state_out = b.state;
  return err;
}

/* Hwang edge map alone (regression_tests/fringe_test/fringe_test.f90). */
ORACLE_API void oracle_hwang_edge(const bmx_ele_rec *ele, int ref_species, int particle_at, double *vecio,
                                  int species, int direction) {
  o_coord a;
  int i;
  init_coord(&a, vecio, ele, species, direction, 0);
  hwang_bend_edge_kick(ele, ref_species, particle_at == 1 ? O_FIRST_EDGE : O_SECOND_EDGE, &a);
  for (i = 0; i < 6; i++) vecio[i] = RV(a.vec[i]);
}

/* check_aperture_limit at the second track edge (regression_tests/aperture_test). */
ORACLE_API int oracle_check_aperture(const bmx_ele_rec *ele, const double *vec, int species, double *unstable) {
  o_coord a;
  int i;
  for (i = 0; i < 6; i++) a.vec[i] = vec[i];
  a.species = species;
  a.direction = 1;
  a.time_dir = 1;
  a.state = BMX_ALIVE;
  a.p0c = ele->p0c;
  a.beta = 1;
  check_aperture_limit(&a, ele, O_SECOND_EDGE, unstable);
  return a.state;
}

/* multipole_ele_to_ab for tests (regression_tests/multipole_test). */
ORACLE_API int oracle_multipole_ele_to_ab(const bmx_ele_rec *ele, int use_ele_tilt, int pole_type, int include_kicks,
                                          double *a, double *b, double *b1) {
  real ra[BMX_N_POLE], rb[BMX_N_POLE], rb1;
  int ix, n;
  multipole_ele_to_ab(ele, use_ele_tilt, pole_type == 2 ? O_ELECTRIC : O_MAGNETIC, include_kicks, &ix, ra, rb,
                      b1 ? &rb1 : NULL);
  for (n = 0; n < BMX_N_POLE; n++) { a[n] = RV(ra[n]); b[n] = RV(rb[n]); }
  if (b1) *b1 = RV(rb1);
  return ix;
}

/*
 * Multi-particle tracking with the reference's loop structure: element outer,
 * OpenMP over particles inner (bmad/multiparticle/beam_utils.f90:69-74), turns
 * outermost (bsim/modules/lt_tracking_mod.f90:1631).  Elements ix_start+1 ..
 * n_ele of the array are tracked each turn (the full ring when ix_start = 0
 * is the beginning element and eles[0] is the first real element: here eles
 * is the element list itself and every element is tracked).
 * SoA in/out; beta in/out (NULL => initialised from pz as init_coord does).
 * ix_lost / turn_lost may be NULL.
 */
ORACLE_API int oracle_track_bunch(const bmx_ele_rec *eles, int n_ele, int ref_species, int n_turns, int64_t n,
                                  double *x, double *px, double *y, double *py, double *z, double *pz, double *beta,
                                  int *state, int *ix_lost, int *turn_lost, int n_threads) {
  int64_t i;
  int turn, ie;
  (void)n_threads;
  if (any_unmodelled(eles, n_ele)) return -2;
  c_multi_init(); /* lazily built table: build it before the parallel region */
  for (turn = 0; turn < n_turns; turn++) {
    for (int k = 0; k < n_ele; k++) {
      const bmx_ele_rec *ele;
      ie = (o_direction == 1) ? k : n_ele - 1 - k;
      ele = &eles[ie];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(n_threads > 0 ? n_threads : 1)
#endif
      for (i = 0; i < n; i++) {
        o_coord a, b;
        if (state[i] != BMX_ALIVE) continue;
        a.vec[0] = PDEP(x[i]); a.vec[1] = PDEP(px[i]); a.vec[2] = PDEP(y[i]); a.vec[3] = PDEP(py[i]);
        a.vec[4] = PDEP(z[i]); a.vec[5] = PDEP(pz[i]);
        a.species = O_TRACKED(ref_species);
        a.direction = o_direction;
        a.time_dir = 1;
        a.state = BMX_ALIVE;
        a.p0c = ele->p0c_start;
        a.beta = PDEP(beta[i]);
        a.t = o_times ? (treal)(o_times[i]) : (treal)(0.0);
        a.s = 0;
        a.phase1 = 0;
        a.location = O_UPSTREAM;
        {
          o_rad_at ra;
          ra.ie = ie;
          ra.pid = (uint32_t)i;
          ra.turn = (uint32_t)turn;
          track1(&a, ele, ref_species, &b, o_rad.maps ? &ra : NULL);
        }
        if (o_times) o_times[i] = RV(b.t);
        x[i] = RV(b.vec[0]); px[i] = RV(b.vec[1]); y[i] = RV(b.vec[2]); py[i] = RV(b.vec[3]);
        z[i] = RV(b.vec[4]); pz[i] = RV(b.vec[5]);
        beta[i] = RV(b.beta);
        state[i] = b.state;
        if (b.state != BMX_ALIVE) {
          if (ix_lost) ix_lost[i] = ie;
          if (turn_lost) turn_lost[i] = turn;
        }
      }
    }
  }
  return 0;
}

/* Particle-outer variant of the same tracking (one particle through all
 * elements and turns), used for the FLOP count and as a second CPU figure. */
ORACLE_API int oracle_track_particle_outer(const bmx_ele_rec *eles, int n_ele, int ref_species, int n_turns, int64_t n,
                                           double *x, double *px, double *y, double *py, double *z, double *pz,
                                           double *beta, int *state, int *ix_lost, int *turn_lost, int n_threads) {
  int64_t i;
  (void)n_threads;
  if (any_unmodelled(eles, n_ele)) return -2;
  c_multi_init(); /* lazily built table: build it before the parallel region */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(n_threads > 0 ? n_threads : 1)
#endif
  for (i = 0; i < n; i++) {
    o_coord a, b;
    int turn, ie;
    if (state[i] != BMX_ALIVE) continue;
    a.vec[0] = PDEP(x[i]); a.vec[1] = PDEP(px[i]); a.vec[2] = PDEP(y[i]); a.vec[3] = PDEP(py[i]);
    a.vec[4] = PDEP(z[i]); a.vec[5] = PDEP(pz[i]);
    a.species = O_TRACKED(ref_species);
    a.direction = 1;
    a.time_dir = 1;
    a.state = BMX_ALIVE;
    a.beta = PDEP(beta[i]);
    a.t = 0;
    a.s = 0;
    a.phase1 = 0;
    a.location = O_UPSTREAM;
    for (turn = 0; turn < n_turns && a.state == BMX_ALIVE; turn++) {
      for (ie = 0; ie < n_ele; ie++) {
        a.p0c = eles[ie].p0c_start;
        track1(&a, &eles[ie], ref_species, &b, NULL);
        a = b;
        if (a.state != BMX_ALIVE) {
          if (ix_lost) ix_lost[i] = ie;
          if (turn_lost) turn_lost[i] = turn;
          break;
        }
      }
    }
    x[i] = RV(a.vec[0]); px[i] = RV(a.vec[1]); y[i] = RV(a.vec[2]); py[i] = RV(a.vec[3]);
    z[i] = RV(a.vec[4]); pz[i] = RV(a.vec[5]);
    beta[i] = RV(a.beta);
    state[i] = a.state;
  }
  return 0;
}

/* The special functions alone, for the branch-switch checks of the tests:
 * which = 0 sinc, 1 cosc, 2 sincc, 3 sqrt_one. */
ORACLE_API double oracle_special(int which, double x) {
  switch (which) {
    case 0: return RV(sinc((real)(x)));
    case 1: return RV(cosc((real)(x)));
    case 2: return RV(sincc((real)(x)));
    case 3: return RV(sqrt_one((real)(x)));
  }
  return 0.0;
}

/* low_energy_z_correction alone (low_energy_z_correction.f90:22-58) for a
 * particle of `species` at pz (beta from convert_pc_to) in an element of
 * reference momentum p0c and energy e_tot, over the length ds. */
ORACLE_API double oracle_low_energy_z(int species, double p0c, double e_tot, double pz, double ds) {
  o_coord a;
  bmx_ele_rec e;
  real pc, mc2 = mass_of(species);
  memset(&e, 0, sizeof(e));
  e.p0c = p0c;
  e.e_tot = e_tot;
  memset(&a, 0, sizeof(a));
  a.vec[5] = pz;
  a.species = species;
  a.p0c = p0c;
  pc = (real)(p0c) * (1 + a.vec[5]);
  a.beta = pc / sqrt(pc * pc + mc2 * mc2);
  return RV(low_energy_z_correction(&a, &e, (real)(ds)));
}

#ifdef ORACLE_DUAL
/* ======================================================================== */
/* Radiation maps (dual-number build only): radiation_map_setup              */
/* (bmad/modules/radiation_mod.f90:160-232), tracking_rad_map_setup          */
/* (bmad/low_level/tracking_rad_map_setup.f90:19-84) and                     */
/* rad1_damp_and_stoc_mats with calc_rad_at_pt and add_bend_edge_to_damp_mat */
/* (bmad/modules/rad_6d_mod.f90:406-718).  Transfer matrices (make_mat6) are */
/* the Jacobians of this oracle's track1 by forward-mode differentiation;    */
/* element slices follow create_element_slice / makeup_super_slave1          */
/* (bmad/code/create_element_slice.f90, bmad/modules/bookkeeper_mod.f90);    */
/* fields are em_field_calc's bmad_standard fields (bmad/code/              */
/* em_field_calc.f90:323-720) through g_bending_strength_from_em_field      */
/* (bmad/code/g_bending_strength_from_em_field.f90:24-76).                  */
/* ======================================================================== */
static const double O_R_E = 2.8179403227e-15;                                   /* physical_constants.f90:35 */
static const double O_H_BAR = 4.135667696e-15 / (2 * 3.141592653589793238462643383279); /* :38-39 */
static const double O_SQRT3 = 1.732050807568877293527446341505872;             /* :16 */

/* classical_radius (bmad/code/classical_radius.f90:14-24) */
static double o_classical_radius(int species) {
  double q = charge_of(species);
  return O_R_E * O_M_ELECTRON * q * q / mass_of(species);
}

static int o_nint(double x) { return (int)(x < 0 ? -floor(-x + 0.5) : floor(x + 0.5)); }

/* A slice of `lord` of length l_slice starting `offset` into it, with its
 * upstream / downstream ends as asked (create_element_slice ->
 * makeup_super_slave1, bookkeeper_mod.f90): fringe_at and the bend edges cut
 * to the ends kept, kicks, voltage and integrated multipoles scaled by the
 * length ratio, offsets moved along the pitch, num_steps from ds_step
 * (attribute_bookkeeper.f90:390-401).  No aperture (the slices carry the
 * reference orbit).  no_misalignment: also drop offsets, pitches and tilts
 * (the slices of calc_rad_at_pt, dispatch = no_misalignment$, which
 * offset_particle.f90:176 honours).  Returns -1 for a pitched or misaligned
 * bend, which this restatement does not slice. */
static int o_slice(const bmx_ele_rec *lord, double l_slice, double offset, int inc_up, int inc_down,
                   int no_misalignment, bmx_ele_rec *sl) {
  double len_lord = lord->l, coef, ds_step, s_del;
  int inc_ent, inc_exit, ifr, n;
  *sl = *lord;
  coef = (fabs(l_slice) >= fabs(len_lord)) ? 1.0 : l_slice / len_lord;
  ds_step = fabs(len_lord) / (lord->num_steps > 0 ? lord->num_steps : 1);
  sl->l = l_slice;
  n = (ds_step > 0) ? o_nint(fabs(l_slice / ds_step)) : 1;
  sl->num_steps = n > 1 ? n : 1;
  inc_ent = lord->orientation == 1 ? inc_up : inc_down;
  inc_exit = lord->orientation == 1 ? inc_down : inc_up;
  if (lord->key != BMX_KEY_MULTIPOLE && lord->key != BMX_KEY_AB_MULTIPOLE) {
    ifr = lord->fringe_at;
    if (inc_ent && inc_exit) {
    } else if (inc_ent) {
      sl->fringe_at = (ifr == BMX_ENTRANCE_END || ifr == BMX_BOTH_ENDS) ? BMX_ENTRANCE_END : BMX_NO_END;
    } else if (inc_exit) {
      sl->fringe_at = (ifr == BMX_EXIT_END || ifr == BMX_BOTH_ENDS) ? BMX_EXIT_END : BMX_NO_END;
    } else {
      sl->fringe_at = BMX_NO_END;
    }
  }
  if (lord->key == BMX_KEY_HKICKER || lord->key == BMX_KEY_VKICKER) {
    sl->kick = lord->kick * coef;
  } else if (has_hkick_attributes(lord->key)) {
    sl->hkick = lord->hkick * coef;
    sl->vkick = lord->vkick * coef;
  }
  if (lord->key == BMX_KEY_RFCAVITY) sl->voltage = lord->voltage * coef;
  if (lord->key != BMX_KEY_MULTIPOLE && lord->key != BMX_KEY_AB_MULTIPOLE)
    for (n = 0; n < BMX_N_POLE; n++) {
      sl->a[n] = lord->a[n] * coef;
      sl->b[n] = lord->b[n] * coef;
    }
  if (has_orientation_attributes(lord->key)) {
    if (lord->key == BMX_KEY_SBEND && lord->g != 0) {
      if (lord->x_offset != 0 || lord->y_offset != 0 || lord->z_offset != 0 || lord->x_pitch != 0 ||
          lord->y_pitch != 0)
        return -1;
    } else {
      s_del = offset + l_slice / 2 - len_lord / 2;
      sl->x_offset = lord->x_offset + s_del * sin(lord->x_pitch) * cos(lord->y_pitch);
      sl->y_offset = lord->y_offset + s_del * sin(lord->y_pitch);
      sl->z_offset = lord->z_offset + s_del * (cos(lord->x_pitch) * cos(lord->y_pitch) - 1);
    }
  }
  sl->x1_limit = sl->x2_limit = sl->y1_limit = sl->y2_limit = 0;
  sl->px_aperture_width2 = sl->py_aperture_width2 = sl->z_aperture_width2 = sl->pz_aperture_width2 = 0;
  sl->delta_ref_time = lord->delta_ref_time * coef;
  if (lord->key == BMX_KEY_SBEND) {
    ifr = sl->fringe_at;
    if (ifr == BMX_NO_END || ifr == BMX_EXIT_END) sl->e1 = sl->fint = sl->hgap = 0;
    if (ifr == BMX_NO_END || ifr == BMX_ENTRANCE_END) sl->e2 = sl->fintx = sl->hgapx = 0;
  }
  if (no_misalignment) {
    sl->x_offset = sl->y_offset = sl->z_offset = sl->x_pitch = sl->y_pitch = 0;
    sl->tilt = 0;
    sl->ref_tilt = 0;
  }
  return 0;
}

/* A particle at vec (values) as the reference orbit: init_coord's beta from
 * pz (carried as a function of pz), coordinates seeded with unit derivatives. */
static void o_seed(o_coord *a, const double *vec, const bmx_ele_rec *ele, int species) {
  int i;
  real pc, mc2 = mass_of(species);
  for (i = 0; i < 6; i++) {
    a->vec[i] = real(vec[i]);
    a->vec[i].d[i] = 1.0;
  }
  a->species = species;
  a->direction = 1;
  a->time_dir = 1;
  a->state = BMX_ALIVE;
  a->p0c = ele->p0c_start;
  a->s = 0;
  a->phase1 = 0;
  a->location = O_UPSTREAM;
  a->t = 0;
  pc = real(ele->p0c_start) * (1 + a->vec[5]);
  a->beta = pc / sqrt(pc * pc + mc2 * mc2);
}

static void o_take(const o_coord *b, double *vout, double *M) {
  int i, j;
  for (i = 0; i < 6; i++) {
    if (vout) vout[i] = b->vec[i].v;
    if (M)
      for (j = 0; j < 6; j++) M[i * 6 + j] = b->vec[i].d[j];
  }
}

/* make_mat6 (bmad/code/make_mat6.f90:23): the element's map and Jacobian at vin. */
static int o_make_mat6(const bmx_ele_rec *ele, int ref_species, const double *vin, double *vout, double *M) {
  o_coord a, b;
  int i;
  if (ele->l == 0 && ele->key != BMX_KEY_MULTIPOLE && ele->key != BMX_KEY_AB_MULTIPOLE) {
    /* a zero-length slice: no body, no ends */
    for (i = 0; i < 6; i++) vout[i] = vin[i];
    for (i = 0; i < 36; i++) M[i] = (i % 7 == 0) ? 1.0 : 0.0;
    return 0;
  }
  o_seed(&a, vin, ele, ref_species);
  track1(&a, ele, ref_species, &b, NULL);
  if (b.state != BMX_ALIVE) return -1;
  o_take(&b, vout, M);
  return 0;
}

static void m6_mul(const double *A, const double *B, double *C) {
  double t[36];
  int i, j, k;
  for (i = 0; i < 6; i++)
    for (j = 0; j < 6; j++) {
      double a = 0;
      for (k = 0; k < 6; k++) a += A[i * 6 + k] * B[k * 6 + j];
      t[i * 6 + j] = a;
    }
  memcpy(C, t, sizeof(t));
}

static void m6_vec(const double *A, const double *v, double *out) {
  double t[6];
  int i, k;
  for (i = 0; i < 6; i++) {
    double a = 0;
    for (k = 0; k < 6; k++) a += A[i * 6 + k] * v[k];
    t[i] = a;
  }
  memcpy(out, t, sizeof(t));
}

static void m6_transpose(const double *A, double *T) {
  double t[36];
  int i, j;
  for (i = 0; i < 6; i++)
    for (j = 0; j < 6; j++) t[j * 6 + i] = A[i * 6 + j];
  memcpy(T, t, sizeof(t));
}

/* mat_symp_conj (sim_utils/matrix/mat_symp_conj.f90): -S M^T S, S the
 * block-diagonal [[0, 1], [-1, 0]]. */
static void m6_symp_conj(const double *M, double *C) {
  double S[36] = {0}, T[36];
  int i;
  for (i = 0; i < 6; i += 2) {
    S[i * 6 + i + 1] = 1;
    S[(i + 1) * 6 + i] = -1;
  }
  m6_transpose(M, T);
  m6_mul(S, T, T);
  m6_mul(T, S, C);
  for (i = 0; i < 36; i++) C[i] = -C[i];
}

/* super_polint (sim_utils/math/super_recipes_mod.f90): Neville's polynomial
 * extrapolation to x of the n points (xa, ya); y and its error estimate dy. */
static void o_super_polint(const double *xa, const double *ya, int n, double x, double *y, double *dy) {
  double c[8], d[8], den[8], ho[8];
  int m, i, ns = 0;
  double best = 1e300;
  for (i = 0; i < n; i++) {
    c[i] = ya[i];
    d[i] = ya[i];
    ho[i] = xa[i] - x;
    if (fabs(x - xa[i]) < best) { best = fabs(x - xa[i]); ns = i; }
  }
  *y = ya[ns];
  ns = ns - 1; /* 0-based: the reference's ns - 1 with ns 1-based */
  *dy = 0;
  for (m = 1; m <= n - 1; m++) {
    for (i = 0; i < n - m; i++) den[i] = ho[i] - ho[i + m];
    for (i = 0; i < n - m; i++) den[i] = (c[i + 1] - d[i]) / den[i];
    for (i = 0; i < n - m; i++) {
      d[i] = ho[i + m] * den[i];
      c[i] = ho[i] * den[i];
    }
    /* reference: if (2*ns < n-m) dy = c(ns+1) else dy = d(ns), ns = ns-1 (ns 1-based after its -1) */
    if (2 * (ns + 1) < n - m) {
      *dy = c[ns + 1];
    } else {
      *dy = d[ns];
      ns = ns - 1;
    }
    *y = *y + *dy;
  }
}

/* em_field_calc's bmad_standard B field of the element at (x, y) of the
 * local orbit (em_field_calc.f90:323-720 for this path's keys), with its x
 * and y derivatives carried by dual numbers (dual slot 0: d/dx, slot 2:
 * d/dy).  Returns -1 for what it does not model (rf cavity fields,
 * electric multipoles). */
static int o_field(const bmx_ele_rec *ele, int ref_species, const double *v, real B[3]) {
  o_coord lo;
  real x, y, kx, ky, a[BMX_N_POLE], b[BMX_N_POLE];
  double f_p0c, ref_charge = charge_of(ref_species), tilt;
  int i, ix, add_kicks;
  memset(&lo, 0, sizeof(lo));
  for (i = 0; i < 6; i++) lo.vec[i] = real(v[i]);
  lo.vec[0].d[0] = 1.0;
  lo.vec[2].d[2] = 1.0;
  lo.species = ref_species;
  lo.direction = 1;
  lo.time_dir = 1;
  lo.beta = 1;
  lo.p0c = ele->p0c;
  x = lo.vec[0];
  y = lo.vec[2];
  f_p0c = (ref_charge == 0) ? 0.0 : ele->p0c / (O_C_LIGHT * ref_charge);
  B[0] = B[1] = B[2] = real(0.0);
  if (ele->has_elec_poles) return -1;
  switch (ele->key) {
    case BMX_KEY_DRIFT: case BMX_KEY_ECOLLIMATOR: case BMX_KEY_RCOLLIMATOR: case BMX_KEY_INSTRUMENT:
    case BMX_KEY_MONITOR: case BMX_KEY_PIPE: case BMX_KEY_MARKER: case BMX_KEY_MULTIPOLE: case BMX_KEY_AB_MULTIPOLE:
    case BMX_KEY_THICK_MULTIPOLE:
      break;
    case BMX_KEY_HKICKER: B[1] = real(-ele->kick * f_p0c / ele->l); break;
    case BMX_KEY_VKICKER: B[0] = real(ele->kick * f_p0c / ele->l); break;
    case BMX_KEY_KICKER:
      B[0] = real(ele->vkick * f_p0c / ele->l);
      B[1] = real(-ele->hkick * f_p0c / ele->l);
      break;
    case BMX_KEY_OCTUPOLE:
      B[0] = -(y * y * y - 3.0 * y * x * x) / 6.0 * ele->k3 * f_p0c;
      B[1] = (x * x * x - 3.0 * x * y * y) / 6.0 * ele->k3 * f_p0c;
      break;
    case BMX_KEY_QUADRUPOLE:
      B[0] = y * ele->k1 * f_p0c;
      B[1] = x * ele->k1 * f_p0c;
      break;
    case BMX_KEY_SEXTUPOLE:
      B[0] = x * y * ele->k2 * f_p0c;
      B[1] = (x * x - y * y) / 2.0 * ele->k2 * f_p0c;
      break;
    case BMX_KEY_SBEND: B[1] = real(ele->g * f_p0c); break;
    default: return -1; /* rf cavity fields are not restated */
  }
  /* multipoles (:661-707): a bend's include its kicks, dg, k1 and k2 */
  if (ele->key == BMX_KEY_SBEND) {
    multipole_ele_to_ab(ele, 0, O_MAGNETIC, 1, &ix, a, b, NULL);
    add_kicks = 0;
  } else {
    multipole_ele_to_ab(ele, 0, O_MAGNETIC, 0, &ix, a, b, NULL);
    add_kicks = 1;
  }
  if (ix > -1) {
    if (ele->l == 0) return -1;
    for (i = 0; i <= ix; i++) {
      if (a[i] == 0 && b[i] == 0) continue;
      ab_multipole_kick(a[i], b[i], i, ref_species, 0, &lo, &kx, &ky, O_MAGNETIC, 0, 1.0);
      B[0] = B[0] + f_p0c * ky / ele->l;
      B[1] = B[1] - f_p0c * kx / ele->l;
    }
  }
  /* hkick / vkick of other elements (:1404-1416 of the field tail) */
  if (add_kicks && has_hkick_attributes(ele->key) && (ele->hkick != 0 || ele->vkick != 0) &&
      ele->key != BMX_KEY_KICKER && ele->key != BMX_KEY_HKICKER && ele->key != BMX_KEY_VKICKER) {
    tilt = (ele->key == BMX_KEY_SBEND) ? ele->ref_tilt : ele->tilt;
    if (tilt == 0) {
      B[0] = B[0] + ele->vkick * f_p0c / ele->l;
      B[1] = B[1] - ele->hkick * f_p0c / ele->l;
    } else {
      B[0] = B[0] + (ele->vkick * cos(tilt) - ele->hkick * sin(tilt)) * f_p0c / ele->l;
      B[1] = B[1] - (ele->hkick * cos(tilt) + ele->vkick * sin(tilt)) * f_p0c / ele->l;
    }
  }
  return 0;
}

/* g_from_field (g_bending_strength_from_em_field.f90:62-74) for a magnetic field */
static void o_g_from_field(const double *B, const double *vel, double beta, double charge, double fact, double *g) {
  double force[3], vb[3], dot;
  int i;
  vb[0] = vel[1] * B[2] - vel[2] * B[1];
  vb[1] = vel[2] * B[0] - vel[0] * B[2];
  vb[2] = vel[0] * B[1] - vel[1] * B[0];
  for (i = 0; i < 3; i++) force[i] = (0.0 + vb[i] * beta * O_C_LIGHT) * charge;
  dot = force[0] * vel[0] + force[1] * vel[1] + force[2] * vel[2];
  for (i = 0; i < 3; i++) g[i] = -(force[i] - vel[i] * dot) * fact;
}

/* The integrand of rad1_damp_and_stoc_mats at z_pos (calc_rad_at_pt,
 * rad_6d_mod.f90:590-686): ele is the half slice, orb0 / v0 its body-frame
 * entrance orbit. */
typedef struct o_rad_stash {
  double gamma0, kd_coef, kf_coef;
} o_rad_stash;

static int o_calc_rad_at_pt(const bmx_ele_rec *ele, int ref_species, const o_rad_stash *st, int opening,
                            const double *v0, double z_pos, double *D1, double *dv1, double *S1, int save,
                            double *orb_save, double *mat_save) {
  bmx_ele_rec runt;
  double vz[6], M[36], mb[36], Bv[3], dBx[3], dBy[3], g[3], dgx[3], dgy[3], vel[3], fact, rel_p, g1, g2, g3,
      dg2_dx, dg2_dy, v[6], kf, kv, gg;
  real B[3];
  int i, j;
  if (o_slice(ele, z_pos, 0.0, 0, 0, 1, &runt) < 0) return -1;
  if (o_make_mat6(&runt, ref_species, v0, vz, M) < 0) return -1;
  if (save) {
    memcpy(orb_save, vz, sizeof(vz));
    memcpy(mat_save, M, sizeof(M));
  }
  m6_symp_conj(M, mb); /* matrix from z_pos back to 0 */
  if (o_field(ele, ref_species, vz, B) < 0) return -1;
  for (i = 0; i < 3; i++) {
    Bv[i] = B[i].v;
    dBx[i] = B[i].d[0];
    dBy[i] = B[i].d[2];
  }
  memcpy(v, vz, sizeof(v));
  rel_p = 1 + v[5];
  vel[0] = v[1] / rel_p;
  vel[1] = v[3] / rel_p;
  vel[2] = sqrt(1 - vel[0] * vel[0] - vel[1] * vel[1]);
  {
    double pc = ele->p0c * rel_p, mc2 = mass_of(ref_species), beta = pc / sqrt(pc * pc + mc2 * mc2);
    double charge = charge_of(O_TRACKED(ref_species));
    fact = 1 / (beta * ele->p0c * rel_p);
    o_g_from_field(Bv, vel, beta, charge, fact, g);
    o_g_from_field(dBx, vel, beta, charge, fact, dgx);
    o_g_from_field(dBy, vel, beta, charge, fact, dgy);
  }
  for (i = 0; i < 3; i++) {
    g[i] = g[i] * rel_p;
    dgx[i] = dgx[i] * rel_p;
    dgy[i] = dgy[i] * rel_p;
  }
  g1 = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  g2 = g1 * g1;
  g3 = g1 * g2;
  dg2_dx = 2 * (g[0] * dgx[0] + g[1] * dgx[1] + g[2] * dgx[2]);
  dg2_dy = 2 * (g[0] * dgy[0] + g[1] * dgy[1] + g[2] * dgy[2]);
  if (ele->key == BMX_KEY_SBEND) {
    gg = 1 + ele->g * v0[0]; /* the reference takes the entrance orbit's x here */
    g1 = g1 * gg;
    g2 = g2 * gg;
    g3 = g3 * gg;
    dg2_dx = dg2_dx * gg + g2 * ele->g;
  }
  /* damping matrix, columns (:645-650) */
  for (i = 0; i < 6; i++) {
    const double m2 = mb[i * 6 + 1], m4 = mb[i * 6 + 3], m6 = mb[i * 6 + 5], kd = st->kd_coef;
    D1[i * 6 + 0] = -kd * (m2 * dg2_dx * v[1] * rel_p + m4 * dg2_dx * v[3] * rel_p + m6 * dg2_dx * rel_p * rel_p);
    D1[i * 6 + 1] = -kd * (m2 * g2 * rel_p);
    D1[i * 6 + 2] = -kd * (m2 * dg2_dy * v[1] * rel_p + m4 * dg2_dy * v[3] * rel_p + m6 * dg2_dy * rel_p * rel_p);
    D1[i * 6 + 3] = -kd * (m4 * g2 * rel_p);
    D1[i * 6 + 4] = 0;
    D1[i * 6 + 5] = -kd * (m2 * g2 * v[1] + m4 * g2 * v[3] + m6 * 2 * g2 * rel_p);
  }
  /* damping vector (:654) */
  {
    double w[6] = {0.0, v[1] * g2 * rel_p, 0.0, v[3] * g2 * rel_p, 0.0, g2 * rel_p * rel_p};
    m6_vec(mb, w, dv1);
    for (i = 0; i < 6; i++) dv1[i] = -st->kd_coef * dv1[i];
  }
  /* stochastic variance (:658-674) */
  kf = st->kf_coef * rel_p * rel_p * g3;
  for (i = 0; i < 6; i++)
    for (j = 0; j < 6; j++) {
      double b2i = mb[i * 6 + 1], b4i = mb[i * 6 + 3], b6i = mb[i * 6 + 5];
      double b2j = mb[j * 6 + 1], b4j = mb[j * 6 + 3], b6j = mb[j * 6 + 5];
      S1[i * 6 + j] = kf * (rel_p * rel_p * b6i * b6j + v[1] * v[1] * b2i * b2j + v[3] * v[3] * b4i * b4j +
                            rel_p * v[1] * (b2i * b6j + b6i * b2j) + rel_p * v[3] * (b4i * b6j + b6i * b4j) +
                            v[1] * v[3] * (b2i * b4j + b4i * b2j));
    }
  if (opening) {
    kv = st->kf_coef * 13.0 * g1 / (55.0 * st->gamma0 * st->gamma0);
    for (i = 0; i < 6; i++)
      for (j = 0; j < 6; j++) {
        double b2i = mb[i * 6 + 1], b4i = mb[i * 6 + 3], b2j = mb[j * 6 + 1], b4j = mb[j * 6 + 3];
        S1[i * 6 + j] = S1[i * 6 + j] + kv * (g[1] * g[1] * b2i * b2j + g[0] * g[0] * b4i * b4j -
                                              g[0] * g[1] * (b2i * b4j + b4i * b2j));
      }
  }
  return 0;
}

/* add_bend_edge_to_damp_mat (rad_6d_mod.f90:691-716) */
static void o_add_bend_edge(double *D, const bmx_ele_rec *ele, double e_edge, const double *orb, const double *mat,
                            double kd) {
  double dm[36] = {0}, dg2_dx, rel_p, ci[36];
  if (e_edge == 0) return;
  dg2_dx = -tan(e_edge) * (ele->g + ele->dg) * (ele->g + ele->dg);
  rel_p = 1 + orb[5];
  dm[1 * 6 + 0] = -kd * dg2_dx * orb[1] * rel_p;
  dm[1 * 6 + 2] = -kd * -dg2_dx * orb[1] * rel_p;
  dm[3 * 6 + 0] = -kd * dg2_dx * orb[3] * rel_p;
  dm[3 * 6 + 2] = -kd * -dg2_dx * orb[3] * rel_p;
  dm[5 * 6 + 0] = -kd * dg2_dx * rel_p * rel_p;
  dm[5 * 6 + 2] = -kd * -dg2_dx * rel_p * rel_p;
  if (mat) {
    m6_symp_conj(mat, ci);
    m6_mul(ci, dm, dm);
    m6_mul(dm, mat, dm);
  }
  for (int i = 0; i < 36; i++) D[i] += dm[i];
}

/* rad1_damp_and_stoc_mats (rad_6d_mod.f90:406-588) for the half slice `ele`
 * whose map is ele_mat6 about orb_in -> orb_out; include_opening_angle. */
static int o_rad1(const bmx_ele_rec *ele, int ref_species, const double *orb_in, const double *orb_out,
                  const double *ele_mat6, double g2_tol, double g3_tol, double *ref, double *D, double *dv,
                  double *S) {
  double length = ele->l, mc2 = mass_of(ref_species), v0[6], mat0[36], mat0_inv[36], orb_end[6], mat_end[36],
         eps_damp, eps_stoc, del_z, l_ref, z_pos, D1[36], dv1[6], S1[36], Dsum[36], dvsum[6], Ssum[36], ddamp[36],
         ddvec[6], dstoc[36], d_max, T[36], xa[5];
  struct { double h, dv[6], D[36], S[36]; } qi[5];
  o_rad_stash st;
  int j, j1, n, n_pts, i, k, j_min_test = 3, j_max = 10;
  const double cd = 2.0 / 3.0, cf = 55.0 * O_H_BAR * O_C_LIGHT / (24.0 * O_SQRT3);
  for (i = 0; i < 6; i++) ref[i] = orb_out[i];
  for (i = 0; i < 36; i++) D[i] = S[i] = 0;
  for (i = 0; i < 6; i++) dv[i] = 0;
  if (length == 0) return 0;
  if (orb_out[1] == orb_in[1] && orb_out[3] == orb_in[3] && ele->key != BMX_KEY_SBEND) return 0;
  if (ele->p0c_start < 2.0 * mc2 || ele->p0c < 2.0 * mc2) return 0;
  /* offsets and fringes at the upstream end */
  {
    o_coord a;
    o_seed(&a, orb_in, ele, ref_species);
    offset_particle(ele, ref_species, 1, &a, 1, 0, -1);
    if (has_fringe(ele)) apply_element_edge_kick(&a, O_FIRST_EDGE, ele, ref_species);
    o_take(&a, v0, mat0);
  }
  {
    double pc = ele->p0c, e = sqrt(pc * pc + mc2 * mc2);
    st.gamma0 = e / mc2; /* convert_pc_to (p0c, ref_species, gamma) */
  }
  {
    double radi = o_classical_radius(ref_species);
    st.kd_coef = cd * radi * st.gamma0 * st.gamma0 * st.gamma0;
    st.kf_coef = cf * radi * pow(st.gamma0, 5) / mc2;
  }
  eps_damp = st.kd_coef * g2_tol / length;
  eps_stoc = st.kf_coef * g3_tol / length;
  memset(qi, 0, sizeof(qi));
  qi[0].h = 4;
  for (j = 1; j <= j_max; j++) {
    if (j == 1) {
      n_pts = 2;
      del_z = length;
      l_ref = 0;
    } else {
      n_pts = 1 << (j - 2);
      del_z = ele->l / n_pts;
      l_ref = del_z / 2;
    }
    memset(Dsum, 0, sizeof(Dsum));
    memset(dvsum, 0, sizeof(dvsum));
    memset(Ssum, 0, sizeof(Ssum));
    for (n = 1; n <= n_pts; n++) {
      z_pos = l_ref + (n - 1) * del_z;
      int save = (j == 1 && n == 1);
      if (o_calc_rad_at_pt(ele, ref_species, &st, 1, v0, z_pos, D1, dv1, S1, save, orb_end, mat_end) < 0)
        return -1;
      for (i = 0; i < 36; i++) { Dsum[i] += D1[i]; Ssum[i] += S1[i]; }
      for (i = 0; i < 6; i++) dvsum[i] += dv1[i];
    }
    j1 = j < 4 ? j : 4;
    if (j > 4)
      for (k = 0; k < 4; k++) qi[k] = qi[k + 1];
    qi[j1].h = 0.25 * qi[j1 - 1].h;
    for (i = 0; i < 36; i++) {
      qi[j1].D[i] = 0.5 * (qi[j1 - 1].D[i] + del_z * Dsum[i]);
      qi[j1].S[i] = 0.5 * (qi[j1 - 1].S[i] + del_z * Ssum[i]);
    }
    for (i = 0; i < 6; i++) qi[j1].dv[i] = 0.5 * (qi[j1 - 1].dv[i] + del_z * dvsum[i]);
    if (j < j_min_test) continue;
    for (k = 0; k < j1; k++) xa[k] = qi[k + 1].h;
    for (i = 0; i < 6; i++) {
      double ya[5];
      for (k = 0; k < j1; k++) ya[k] = qi[k + 1].dv[i];
      o_super_polint(xa, ya, j1, 0.0, &dv1[i], &ddvec[i]);
    }
    for (i = 0; i < 36; i++) {
      double ya[5];
      for (k = 0; k < j1; k++) ya[k] = qi[k + 1].D[i];
      o_super_polint(xa, ya, j1, 0.0, &D1[i], &ddamp[i]);
      for (k = 0; k < j1; k++) ya[k] = qi[k + 1].S[i];
      o_super_polint(xa, ya, j1, 0.0, &S1[i], &dstoc[i]);
    }
    d_max = 0;
    for (i = 0; i < 36; i++) {
      if (eps_damp * fabs(ddamp[i]) > d_max) d_max = eps_damp * fabs(ddamp[i]);
      if (eps_stoc * fabs(dstoc[i]) > d_max) d_max = eps_stoc * fabs(dstoc[i]);
    }
    for (i = 0; i < 6; i++)
      if (eps_damp * fabs(ddvec[i]) > d_max) d_max = eps_damp * fabs(ddvec[i]);
    if (d_max < 1) break;
  }
  /* bend edges */
  if (ele->key == BMX_KEY_SBEND) {
    if (ele->orientation == 1) {
      o_add_bend_edge(D1, ele, ele->e1, v0, NULL, st.kd_coef);
      o_add_bend_edge(D1, ele, ele->e2, orb_end, mat_end, st.kd_coef);
    } else {
      o_add_bend_edge(D1, ele, ele->e2, v0, NULL, st.kd_coef);
      o_add_bend_edge(D1, ele, ele->e1, orb_end, mat_end, st.kd_coef);
    }
  }
  /* referred to the downstream end (:568-577) */
  m6_symp_conj(mat0, mat0_inv);
  m6_mul(ele_mat6, mat0_inv, T);
  m6_vec(T, dv1, dv);
  m6_mul(T, D1, D);
  m6_mul(D, mat0, D);
  m6_mul(mat0_inv, S1, S1);
  m6_transpose(mat0_inv, T);
  m6_mul(S1, T, S1);
  m6_mul(ele_mat6, S1, S);
  m6_transpose(ele_mat6, T);
  m6_mul(S, T, S);
  return 0;
}

/* dpotrf 'L': S = L L^T in place (lower triangle), upper set to zero; -1 when not positive definite */
static int o_cholesky_lower(double *A) {
  double L[36] = {0};
  int i, j, k;
  for (j = 0; j < 6; j++) {
    double s = A[j * 6 + j];
    for (k = 0; k < j; k++) s -= L[j * 6 + k] * L[j * 6 + k];
    if (!(s > 0)) return -1;
    L[j * 6 + j] = sqrt(s);
    for (i = j + 1; i < 6; i++) {
      double t = A[i * 6 + j];
      for (k = 0; k < j; k++) t -= L[i * 6 + k] * L[j * 6 + k];
      L[i * 6 + j] = t / L[j * 6 + j];
    }
  }
  memcpy(A, L, sizeof(L));
  return 0;
}

/* tracking_rad_map_setup for one half slice (tracking_rad_map_setup.f90:19-84):
 * out = ref_orb[6], damp_dmat[36], xfer_damp_vec[6], stoc_mat[36]. */
static int o_tracking_rad_map_setup(const bmx_ele_rec *runt, int ref_species, const double *orb0, const double *orb1,
                                    const double *mat6, double tol, double g2_int, double g3_int, int upstream,
                                    double *out) {
  double *ref = out, *D = out + 6, *dv = out + 42, *S = out + 48, m_inv[36], T[36];
  int i;
  memset(out, 0, 84 * sizeof(double));
  for (i = 0; i < 6; i++) ref[i] = -1; /* rad_map_struct default */
  if (orb0[1] == orb1[1] && orb0[3] == orb1[3] && runt->key != BMX_KEY_SBEND) return 0;
  if (o_rad1(runt, ref_species, orb0, orb1, mat6, tol * g2_int, tol * g3_int, ref, D, dv, S) < 0) return -1;
  if (upstream) {
    m6_symp_conj(mat6, m_inv);
    m6_mul(m_inv, D, D);
    m6_vec(m_inv, dv, dv);
    m6_mul(m_inv, S, S);
    m6_transpose(m_inv, T);
    m6_mul(S, T, S);
    for (i = 0; i < 6; i++) ref[i] = orb0[i];
  } else {
    for (i = 0; i < 6; i++) ref[i] = orb1[i];
  }
  if (o_cholesky_lower(S) < 0) memset(S, 0, 36 * sizeof(double));
  return 0;
}

/* radiation_map_setup (radiation_mod.f90:160-232) about the reference orbit
 * orb_in at the element's entrance: rm[0 .. 83] = rm0 (entrance half),
 * rm[84 .. 167] = rm1 (exit half), in the bmx_rad_map layout.  total_length,
 * g2_integral and g3_integral are the branch's (branch%param): they set the
 * Romberg tolerance (1e-4 / total_length of each integral).  Returns 0, -1
 * when an element or its reference orbit is outside what this restatement
 * models, or -2 when the reference particle is lost in the element. */
ORACLE_API int oracle_rad_map_setup(const bmx_ele_rec *ele, int ref_species, const double *orb_in,
                                    double total_length, double g2_integral, double g3_integral, double *rm) {
  bmx_ele_rec runt;
  double orb_out[6], orb1[6], orb2[6], M[36], tol;
  int i;
  memset(rm, 0, 2 * 84 * sizeof(double));
  for (i = 0; i < 6; i++) rm[i] = rm[84 + i] = -1;
  if (ele->l == 0) return 0;
  if (o_make_mat6(ele, ref_species, orb_in, orb_out, M) < 0) return -2;
  if (orb_in[1] == orb_out[1] && orb_in[3] == orb_out[3] && ele->key != BMX_KEY_SBEND) {
    for (i = 0; i < 6; i++) {
      rm[i] = orb_in[i];
      rm[84 + i] = orb_out[i];
    }
    return 0;
  }
  tol = 1e-4 / total_length;
  /* first half: the upstream end */
  if (o_slice(ele, 0.5 * ele->l, 0.0, 1, 0, 0, &runt) < 0) return -1;
  if (o_make_mat6(&runt, ref_species, orb_in, orb1, M) < 0) return -2;
  if (o_tracking_rad_map_setup(&runt, ref_species, orb_in, orb1, M, tol, g2_integral, g3_integral, 1, rm) < 0)
    return -1;
  /* second half: the downstream end, its map referred to the element's exit orbit */
  if (o_slice(ele, 0.5 * ele->l, 0.5 * ele->l, 0, 1, 0, &runt) < 0) return -1;
  if (o_make_mat6(&runt, ref_species, orb1, orb2, M) < 0) return -2;
  if (o_tracking_rad_map_setup(&runt, ref_species, orb1, orb_out, M, tol, g2_integral, g3_integral, 0, rm + 84) < 0)
    return -1;
  return 0;
}

/* The element's map and Jacobian at vin (make_mat6), for tests. */
ORACLE_API int oracle_make_mat6(const bmx_ele_rec *ele, int ref_species, const double *vin, double *vout, double *M) {
  return o_make_mat6(ele, ref_species, vin, vout, M);
}

/* A slice of an element (create_element_slice), for tests. */
ORACLE_API int oracle_ele_slice(const bmx_ele_rec *ele, double l_slice, double offset, int inc_up, int inc_down,
                                bmx_ele_rec *out) {
  return o_slice(ele, l_slice, offset, inc_up, inc_down, 0, out);
}
#endif

#ifdef ORACLE_DUAL
/* One particle through n_turns of the element list with dual numbers: returns
 * the final coordinates and the 6x6 Jacobian J[i*6+j] = d out_i / d in_j. */
ORACLE_API int oracle_jacobian(const bmx_ele_rec *eles, int n_ele, int ref_species, int n_turns, const double *vec,
                               double beta, double *out, double *J) {
  o_coord a, b;
  int turn, ie, i, j;
  if (any_unmodelled(eles, n_ele)) return -2;
  for (i = 0; i < 6; i++) {
    a.vec[i] = real(vec[i]);
    a.vec[i].d[i] = 1.0;
  }
  a.species = O_TRACKED(ref_species);
  a.direction = 1;
  a.time_dir = 1;
  a.state = BMX_ALIVE;
  a.t = 0;
  a.s = 0;
  a.phase1 = 0;
  a.location = O_UPSTREAM;
  {
    /* beta as a function of pz: convert_pc_to, so its derivative is carried */
    real pc = real(eles[0].p0c_start) * (1 + a.vec[5]);
    real mc2 = mass_of(ref_species);
    a.beta = (vec[5] == 0) ? real(beta) : pc / sqrt(pc * pc + mc2 * mc2);
  }
  for (turn = 0; turn < n_turns; turn++) {
    for (ie = 0; ie < n_ele; ie++) {
      a.p0c = eles[ie].p0c_start;
      track1(&a, &eles[ie], ref_species, &b, NULL);
      a = b;
      if (a.state != BMX_ALIVE) return a.state;
    }
  }
  for (i = 0; i < 6; i++) {
    out[i] = a.vec[i].v;
    for (j = 0; j < 6; j++) J[i * 6 + j] = a.vec[i].d[j];
  }
  return 0;
}
#endif

#ifdef ORACLE_COUNT_FLOPS
ORACLE_API long long oracle_flops_read_reset(void) {
  long long v = oracle_flop_counter;
  oracle_flop_counter = 0;
  return v;
}
#endif
