"""Parity of the HIP tracker (through the libbmx C ABI) with the CPU oracle.

Tolerance: after one turn every coordinate of every surviving particle agrees
to 1e-12 relative to that coordinate's largest magnitude over the bunch -- for
z relative to max(max|z|, 1e-3 x the distance tracked), because z is formed as
a difference of path lengths of that size (the oracle itself moves by up to
8.5e-11 of max|z| between two FP-legal evaluation orders on the 4000-element
ring; tests/test_fp_conditioning.py); loss states, loss elements and loss turns
are identical.  Multi-turn survival flags are compared outside the chaotic band:
particles whose oracle survival changes under a 1e-12 relative perturbation
of their initial coordinates are excluded and counted.
"""
import math

import numpy as np
import pytest

from bmx import abi, beam, lattice, tracker
import oracle
from helpers import coord_errors

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _coords(n, seed, scale):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((6, n)) * np.array(scale)[:, None]
    return c


def _gpu_track(ctx, eles, species, coords, n_turns, tracked_species=None, direction=1, ix_start=0, ix_end=0, t=None,
               absolute_time=False):
    lat = ctx.lattice(eles, species, tracked_species=tracked_species, absolute_time_tracking=absolute_time)
    b = tracker.Bunch(coords, lat.p0c, abi.MASS[species], e_tot=lat.e_tot, t=t)
    tracker.track_bunch(lat, b, n_turns=n_turns, ix_start=ix_start, ix_end=ix_end, direction=direction)
    out = b.numpy()
    lat.close()
    return out


def _path(eles, n_turns):
    return n_turns * sum(e.l for e in eles)


def _ref_times(eles):
    """Reference times of the element boundaries: [entrance of element 0, exit of each element]."""
    return np.array([eles[0].ref_time - eles[0].delta_ref_time] + [e.ref_time for e in eles])


def _init_times(eles, coords, species, edge, jitter=0.0, seed=5):
    """init_coord's particle time at a boundary of the element list
    (init_coord.f90:222: t = ref_time - z / (beta c)), plus a seeded offset
    of up to `jitter` seconds per particle."""
    p0c = eles[0].p0c
    beta = oracle.beta_of(coords[5], p0c, abi.MASS[species])
    beta = np.where(coords[5] == 0, p0c / eles[0].e_tot, beta)
    t = _ref_times(eles)[edge] - coords[4] / (beta * abi.C_LIGHT)
    if jitter:
        t = t + jitter * np.random.default_rng(seed).uniform(-1.0, 1.0, coords.shape[1])
    return t


def _compare_times(g, r, tol=1e-12):
    """Particle times of the surviving particles: relative to the largest |t|."""
    alive = r["state"] == abi.ALIVE
    if not alive.any():
        return 0.0
    err = float(np.max(np.abs(g["t"][alive] - r["t"][alive])) / np.max(np.abs(r["t"][alive])))
    print(f"parity: t {err:.2e}", flush=True)
    assert err <= tol, err
    return err


def _compare(g, r, path_length, tol=TOL):
    assert np.array_equal(g["state"], r["state"]), (
        f"state mismatch at {np.flatnonzero(g['state'] != r['state'])[:10]}")
    lost = r["state"] != abi.ALIVE
    assert np.array_equal(g["ix_lost"][lost], r["ix_lost"][lost])
    assert np.array_equal(g["turn_lost"][lost], r["turn_lost"][lost])
    alive = ~lost
    errs = coord_errors(g, r, alive, path_length)
    print("parity:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    worst = max(errs.values())
    assert worst <= tol, f"relative error {errs} > {tol}"
    return worst


SCALE = [2e-3, 2e-4, 2e-3, 2e-4, 2e-3, 2e-3]

ELEMENT_CASES = {
    "drift": [lattice.ele("drift", l=1.3)],
    "quad_f": [lattice.ele("quadrupole", l=0.5, k1=0.8, num_steps=1)],
    "quad_d": [lattice.ele("quadrupole", l=0.5, k1=-0.8, num_steps=1)],
    "quad_tiny_k": [lattice.ele("quadrupole", l=0.5, k1=1e-22, num_steps=1)],
    "quad_mult_offset_tilt": [lattice.ele("quadrupole", l=0.6, k1=1.0, tilt=math.pi / 4, num_steps=10, x_offset=0.05,
                                          y_pitch=0.001, a={0: 0.01, 1: 0.1}, vkick=-1e-4)],
    "sbend_hard_edge": [lattice.ele("sbend", l=2.0, angle=2 * math.pi / 200, fringe_type=abi.FRINGE_HARD_EDGE_ONLY,
                                    num_steps=1)],
    "sbend_e1e2_fint": [lattice.ele("sbend", l=0.6, g=0.01, e1=0.1, e2=0.2, fint=0.5, hgap=0.02, num_steps=1)],
    "sbend_k1_k2": [lattice.ele("sbend", l=0.6, g=0.05, k1=0.3, k2=3.0, num_steps=4)],
    "sbend_linear_edge": [lattice.ele("sbend", l=1.0, g=0.03, e1=0.05, e2=-0.04, fringe_type=abi.FRINGE_LINEAR_EDGE,
                                      num_steps=1)],
    "thin_sextupole": [lattice.ele("multipole", knl={2: 1.5})],
    "thin_multipole_tilted": [lattice.ele("multipole", knl={1: 1.0, 3: 4.5}, tn={1: math.pi / 4, 3: 0.31 * math.pi},
                                          x_offset=0.001)],
    "thin_dipole_ref": [lattice.ele("multipole", knl={0: 1e-3, 2: 0.5})],
    "ab_multipole": [lattice.ele("ab_multipole", a={2: 10.0}, b={2: 10.0}, x_offset=0.002, tilt=0.1)],
    "thick_sextupole": [lattice.ele("sextupole", l=0.2, k2=7.5, num_steps=4)],
    "thick_sextupole_offset": [lattice.ele("sextupole", l=0.2, k2=-7.5, num_steps=4, x_offset=5e-5, y_offset=-3e-5)],
    "octupole_tilt_pitch": [lattice.ele("octupole", l=0.5, k3=100.0, tilt=math.pi / 8, num_steps=5, x_offset=0.001,
                                        y_pitch=0.001)],
    "thick_multipole": [lattice.ele("thick_multipole", l=0.8, a={2: 1.0}, b={3: 20.0}, x_offset=0.002, tilt=0.1,
                                    num_steps=7)],
    "kickers": [lattice.ele("hkicker", l=0.5, kick=1e-4, num_steps=3), lattice.ele("vkicker", l=0.5, kick=-2e-4,
                                                                                  num_steps=2),
                lattice.ele("kicker", l=0.4, hkick=1e-4, vkick=2e-4, num_steps=2, tilt=0.2)],
    "instrument_monitor": [lattice.ele("instrument", l=0.3), lattice.ele("monitor", l=0.2)],
    "rfcavity": [lattice.ele("rfcavity", l=1.0, voltage=1e6, rf_frequency=500e6, phi0=0.1, num_steps=10)],
    "rfcavity_offset_tilt": [lattice.ele("rfcavity", l=0.5, voltage=5e5, rf_frequency=500e6, phi0=0.0, num_steps=5,
                                         x_offset=1e-3, tilt=0.1, a={2: 0.02})],
    "quad_hard_edge": [lattice.ele("quadrupole", l=0.5, k1=0.8, num_steps=1, fringe_type=abi.FRINGE_HARD_EDGE_ONLY)],
    "quad_full_fringe_entrance_tilt": [lattice.ele("quadrupole", l=0.4, k1=-1.2, num_steps=3, tilt=0.2, x_offset=1e-3,
                                                   fringe_type=abi.FRINGE_FULL, fringe_at=abi.ENTRANCE_END,
                                                   b={2: 0.5})],
    "quad_soft_edge_only": [lattice.ele("quadrupole", l=0.5, k1=0.8, num_steps=1, fringe_type=abi.FRINGE_SOFT_EDGE_ONLY)],
    "sext_hard_edge_exit": [lattice.ele("sextupole", l=0.2, k2=-20.0, num_steps=4, fringe_type=abi.FRINGE_FULL,
                                        fringe_at=abi.EXIT_END, y_offset=2e-4)],
    "sext_hard_edge_both": [lattice.ele("sextupole", l=0.3, k2=30.0, num_steps=2,
                                        fringe_type=abi.FRINGE_HARD_EDGE_ONLY)],
    "sbend_offset_pitch": [lattice.ele("sbend", l=2.0, angle=0.05, num_steps=1, x_offset=1e-3, y_offset=-5e-4,
                                       z_offset=2e-3, x_pitch=2e-4, y_pitch=-3e-4)],
    "sbend_ref_tilt": [lattice.ele("sbend", l=1.5, angle=0.03, ref_tilt=0.4, num_steps=1)],
    "sbend_ref_tilt_offset_k1": [lattice.ele("sbend", l=1.0, g=0.02, k1=0.2, ref_tilt=-0.3, x_offset=5e-4,
                                             y_pitch=1e-4, e1=0.01, e2=0.02, num_steps=4)],
    "sbend_zero_g_offset": [lattice.ele("sbend", l=0.8, g=0.0, x_offset=1e-3, x_pitch=1e-4, num_steps=1)],
    "quad_electric": [lattice.ele("quadrupole", l=0.6, k1=0.5, num_steps=4, a_elec={1: 2e4}, b_elec={0: -2e3},
                                  tilt=0.3, x_offset=1e-3)],
    "quad_electric_hard_edge": [lattice.ele("quadrupole", l=0.6, k1=-1.0, num_steps=3, b_elec={1: 5e4, 2: 1e6},
                                            r0_elec=0.05, fringe_type=abi.FRINGE_FULL)],
    "sext_electric_entrance": [lattice.ele("sextupole", l=0.3, k2=10.0, num_steps=3, b_elec={2: 2e6}, a_elec={0: 3e3},
                                           fringe_at=abi.ENTRANCE_END)],
    "instrument_electric": [lattice.ele("instrument", l=0.4, b_elec={1: 3e4})],
    "rfcavity_electric": [lattice.ele("rfcavity", l=0.5, voltage=5e5, rf_frequency=500e6, phi0=0.05, num_steps=5,
                                      a_elec={1: 1e4})],
    "sbend_electric_edges": [lattice.ele("sbend", l=0.6, g=0.01, e1=0.1, e2=0.2, num_steps=4, a_elec={1: 2e4},
                                         b_elec={0: -2e3}, r0_elec=3.0, x_offset=5e-4, ref_tilt=0.1)],
    "sbend_k1_electric": [lattice.ele("sbend", l=0.6, g=0.02, k1=0.3, num_steps=5, b_elec={1: 3e4, 2: 2e5},
                                      fringe_at=abi.EXIT_END)],
    "rfcavity_hvkick_tilt_offset": [lattice.ele("rfcavity", l=0.8, voltage=4e5, rf_frequency=500e6, phi0=0.02,
                                                num_steps=4, hkick=1e-4, vkick=-2e-4, tilt=0.2, x_offset=1e-3)],
    "rfcavity_coupler_both": [lattice.ele("rfcavity", l=1.0, voltage=1e6, rf_frequency=500e6, phi0=0.1, num_steps=5,
                                          coupler_strength=0.05, coupler_angle=0.1, coupler_phase=0.2,
                                          coupler_at=abi.BOTH_ENDS)],
    "rfcavity_coupler_entrance": [lattice.ele("rfcavity", l=0.6, voltage=5e5, rf_frequency=1.3e9, phi0=-0.05,
                                              num_steps=3, coupler_strength=0.1, coupler_angle=0.3,
                                              coupler_at=abi.ENTRANCE_END)],
    "rcollimator_phase_space": [lattice.ele("rcollimator", l=0.5, x_limit=3e-3, y_limit=2e-3,
                                            px_aperture_width2=3e-4, py_aperture_width2=4e-4, pz_aperture_width2=3e-3,
                                            z_aperture_width2=4e-3, z_aperture_center=1e-4)],
    "ecollimator_phase_space_both_ends": [lattice.ele("ecollimator", l=0.4, x_limit=4e-3, y_limit=3e-3,
                                                      aperture_at=abi.BOTH_ENDS, px_aperture_width2=4e-4,
                                                      py_aperture_width2=5e-4, px_aperture_center=2e-5,
                                                      pz_aperture_width2=3e-3)],
    "rcollimator": [lattice.ele("rcollimator", l=0.5, x_limit=3e-3, y_limit=2e-3)],
    "ecollimator": [lattice.ele("ecollimator", l=0.5, x_limit=3e-3, y_limit=2e-3)],
    "aperture_entrance": [lattice.ele("drift", l=0.5, x_limit=3e-3, y_limit=3e-3, aperture_at=abi.ENTRANCE_END,
                                      aperture_type=abi.APERTURE_ELLIPTICAL)],
}


@pytest.mark.parametrize("name", sorted(ELEMENT_CASES))
@pytest.mark.parametrize("e_tot", [1e9, 3e6])
def test_element_parity(ctx, name, e_tot):
    eles = lattice.set_reference([lattice.EleRec_copy(e) for e in ELEMENT_CASES[name]], abi.ELECTRON, e_tot=e_tot)
    coords = _coords(4096, 11, SCALE)
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 1)
    r = oracle.track_bunch(eles, abi.ELECTRON, 1, coords)
    _compare(g, r, _path(eles, 1))


@pytest.mark.parametrize("name", sorted(ELEMENT_CASES))
@pytest.mark.parametrize("variant", ["reversed", "reversed_anti", "anti"])
def test_element_parity_reversed_and_antiparticle(ctx, name, variant):
    """Every element case with orientation -1 (sector bends included), tracked as the reference species or as its
    antiparticle (a positron through an electron lattice), against the
    oracle's offset_particle / fringe_here / charge_dir logic for those."""
    eles = lattice.set_reference([lattice.EleRec_copy(e) for e in ELEMENT_CASES[name]], abi.ELECTRON, e_tot=1e9)
    if variant.startswith("reversed"):
        for e in eles:
            e.orientation = -1
    tracked = abi.POSITRON if variant.endswith("anti") else None
    coords = _coords(4096, 13, SCALE)
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 1, tracked_species=tracked)
    r = oracle.track_bunch(eles, abi.ELECTRON, 1, coords, tracked_species=tracked or 0)
    _compare(g, r, _path(eles, 1))


@pytest.mark.parametrize("name", sorted(ELEMENT_CASES))
@pytest.mark.parametrize("variant", ["backward", "backward_reversed_anti"])
def test_element_parity_backward(ctx, name, variant):
    """Every element case tracked backward (direction = -1): the backward
    drift, the quadrupole's 2 c beta delta_ref_time, the k1 bend's z0 sign,
    physical ends and the sign_z_vel = -1 frame changes, against the oracle's
    backward loop; as the reference species (orientation +1) and as the
    antiparticle through reversed elements."""
    eles = lattice.set_reference([lattice.EleRec_copy(e) for e in ELEMENT_CASES[name]], abi.ELECTRON, e_tot=1e9)
    tracked = None
    if variant == "backward_reversed_anti":
        for e in eles:
            e.orientation = -1
        tracked = abi.POSITRON
    coords = _coords(4096, 17, SCALE)
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 1, tracked_species=tracked, direction=-1)
    r = oracle.track_bunch(eles, abi.ELECTRON, 1, coords, tracked_species=tracked or 0, direction=-1)
    _compare(g, r, _path(eles, 1))


@pytest.mark.parametrize("name", sorted(n for n in ELEMENT_CASES if n.startswith("rfcavity")))
@pytest.mark.parametrize("direction", [1, -1])
def test_rf_parity_absolute_time(ctx, name, direction):
    """Every rf cavity case with absolute-time phases (particle_rf_time.f90:
    72-74), forward and backward, from init_coord's times at the start edge
    plus per-particle offsets of up to 0.3 ns (a sizeable fraction of an rf
    period), against the oracle carrying coord_struct%t: coordinates, states
    and the particle times."""
    eles = lattice.set_reference([lattice.EleRec_copy(e) for e in ELEMENT_CASES[name]], abi.ELECTRON, e_tot=1e9)
    coords = _coords(4096, 19, SCALE)
    edge = 0 if direction == 1 else len(eles)
    t0 = _init_times(eles, coords, abi.ELECTRON, edge, jitter=3e-10)
    ix = 0 if direction == 1 else len(eles)
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 1, direction=direction, ix_start=ix, ix_end=ix, t=t0,
                   absolute_time=True)
    r = oracle.track_bunch(eles, abi.ELECTRON, 1, coords, direction=direction, t=t0, absolute_time=True)
    _compare(g, r, _path(eles, 1))
    _compare_times(g, r)


def _spread_tol(eles, n_turns, coords, species=abi.ELECTRON, **kw):
    """Parity bar over several turns: 4 x the oracle's own spread between two
    FP-legal evaluation orders (plain and FMA-contracted builds) on the same
    particles, never below the one-turn bar."""
    from helpers import coord_errors
    a = oracle.track_bunch(eles, species, n_turns, coords, **kw)
    b = oracle.track_bunch(eles, species, n_turns, coords, fma=True, **kw)
    m = (a["state"] == abi.ALIVE) & (b["state"] == abi.ALIVE)
    spread = max(coord_errors(a, b, m, _path(eles, n_turns)).values())
    print(f"parity: oracle FP-order spread over {n_turns} turns {spread:.2e}", flush=True)
    return max(TOL, 4.0 * spread)


def test_ring_absolute_time_off_harmonic(ctx):
    """A FODO ring with two 1 MV cavities whose 500 MHz is not a harmonic of
    the revolution frequency: in absolute time the rf phase of the reference
    particle slips from turn to turn (whole lattice periods enter the phase as
    k frac(f T)), which relative time would not show.  One turn at the 1e-12
    bar; five turns forward from init_coord's times, then three backward from
    a mid-ring element, at 4 x the oracle's own FP-order spread (relative time
    alongside, as the control), against the oracle carrying coord_struct%t."""
    eles = lattice.fodo_ring_c0(n_cells=8)
    for i in (10, 38):  # two 1.0 m drifts become cavities
        e = eles[i]
        eles[i] = lattice.ele("rfcavity", l=e.l, voltage=1e6, rf_frequency=500e6, phi0=0.02, num_steps=5,
                              x_limit=0.01, y_limit=0.01, aperture_type=abi.APERTURE_ELLIPTICAL)
    lattice.set_reference(eles, abi.ELECTRON, e_tot=1e9)
    f_t = 500e6 * lattice.revolution_time(eles)
    assert abs(f_t - round(f_t)) > 0.05, f_t  # off harmonic
    coords, _ = beam.c0_bunch(eles[:7], n=4096, seed=23)
    t0 = _init_times(eles, coords, abi.ELECTRON, 0)
    for n_turns in (1, 5):
        kw = dict(t=t0, absolute_time=True)
        g = _gpu_track(ctx, eles, abi.ELECTRON, coords, n_turns, **kw)
        r = oracle.track_bunch(eles, abi.ELECTRON, n_turns, coords, **kw)
        _compare(g, r, _path(eles, n_turns), tol=TOL if n_turns == 1 else _spread_tol(eles, n_turns, coords, **kw))
        _compare_times(g, r)
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 5)
    r_rel = oracle.track_bunch(eles, abi.ELECTRON, 5, coords)
    _compare(g, r_rel, _path(eles, 5), tol=_spread_tol(eles, 5, coords))
    assert np.max(np.abs(r_rel["pz"] - r["pz"])) > 1e-6  # the slip matters
    # backward from the exit of element 20 (Bmad index), three turns
    ix = 20
    tb = _init_times(eles, coords, abi.ELECTRON, ix)
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 3, direction=-1, ix_start=ix, ix_end=ix, t=tb, absolute_time=True)
    rot = eles[ix:] + eles[:ix]  # the oracle's backward loop starts at the end of its list
    r = oracle.track_bunch(rot, abi.ELECTRON, 3, coords, direction=-1, t=tb, absolute_time=True)
    r["ix_lost"] = np.where(r["ix_lost"] >= 0, (r["ix_lost"] + ix) % len(eles), -1)
    _compare(g, r, _path(eles, 3), tol=_spread_tol(rot, 3, coords, direction=-1, t=tb, absolute_time=True))
    _compare_times(g, r)


def test_ring_backward_turns_and_element_ranges(ctx):
    """A positron bunch circulating backward through an electron FODO ring
    (20 cells, misaligned and tilted elements absent): 3 full turns, then
    track_bunch's two backward ranges (ix_start > ix_end, and the wrap
    through element 1 and the last element), loss elements included."""
    eles = lattice.fodo_ring_c0(n_cells=20)
    coords, _ = beam.c0_bunch(eles[:7], n=4096, seed=23)
    coords[0] *= 5.0  # large enough for some losses on the 1 cm aperture
    coords[2] *= 5.0
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 3, tracked_species=abi.POSITRON, direction=-1)
    r = oracle.track_bunch(eles, abi.ELECTRON, 3, coords, n_threads=8, tracked_species=abi.POSITRON, direction=-1)
    n_lost = int(np.sum(r["state"] != abi.ALIVE))
    print("backward ring: lost", n_lost, "of", coords.shape[1], flush=True)
    _compare(g, r, _path(eles, 3))
    n = len(eles)
    for ix_start, ix_end in ((n - 5, 17), (30, 60)):
        g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 1, tracked_species=abi.POSITRON, direction=-1,
                       ix_start=ix_start, ix_end=ix_end)
        idx = list(range(ix_end, ix_start)) if ix_start > ix_end else list(range(ix_end, n)) + list(range(ix_start))
        r = oracle.track_bunch([eles[i] for i in idx], abi.ELECTRON, 1, coords, n_threads=8,
                               tracked_species=abi.POSITRON, direction=-1)
        lost = r["state"] != abi.ALIVE
        r["ix_lost"] = np.where(lost, np.array(idx)[np.where(lost, r["ix_lost"], 0)], -1)
        _compare(g, r, sum(eles[i].l for i in idx))


def test_zero_particle_and_reference_orbit(ctx):
    """Exact reference particle (all zero) stays on the reference orbit; the
    sbend near-zero branch (track_a_bend.f90:136-159) is taken."""
    eles = lattice.fodo_ring_c0(n_cells=2)
    coords = np.zeros((6, 64))
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 5)
    r = oracle.track_bunch(eles, abi.ELECTRON, 5, coords)
    _compare(g, r, _path(eles, 5), tol=0.0 if np.all(r["x"] == 0) else TOL)


def test_c0_ring_one_turn(ctx):
    eles = lattice.fodo_ring_c0()
    coords, _ = beam.c0_bunch(eles[:7], n=10_000, seed=20261203)
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 1)
    r = oracle.track_bunch(eles, abi.ELECTRON, 1, coords, n_threads=8)
    _compare(g, r, _path(eles, 1))


def test_c0_ring_large_amplitude_losses(ctx):
    """Amplitudes up to the 1 cm aperture: losses on the aperture and in the
    sextupoles; states, loss element and turn must agree exactly."""
    eles = lattice.fodo_ring_c0(n_cells=20)
    coords, _ = beam.c0_bunch(eles[:7], n=4096, seed=5)
    coords *= 25.0
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 3)
    r = oracle.track_bunch(eles, abi.ELECTRON, 3, coords, n_threads=8)
    assert np.sum(r["state"] != abi.ALIVE) > 0
    _compare(g, r, _path(eles, 3), tol=1e-11)


def test_c1_ring_with_rf_one_turn(ctx):
    eles = lattice.ring_c1(n_cells=500)
    tw = beam.periodic_twiss(lattice.ring_c1(n_cells=500, n_rf=0)[:8])
    coords = beam.matched_gaussian(2048, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed=3)
    g = _gpu_track(ctx, eles, abi.ELECTRON, coords, 1)
    r = oracle.track_bunch(eles, abi.ELECTRON, 1, coords, n_threads=8)
    # the oracle's own spread between two FP-legal evaluation orders (plain and
    # FMA-contracted) on the same particles: the floor of this metric
    f = oracle.track_bunch(eles, abi.ELECTRON, 1, coords, n_threads=8, fma=True)
    floor = coord_errors(f, r, r["state"] == abi.ALIVE, _path(eles, 1))
    print("oracle plain vs fma:", {k: f"{v:.2e}" for k, v in floor.items()}, flush=True)
    _compare(g, r, _path(eles, 1))


@pytest.mark.parametrize("ring", ["c0", "c1"])
def test_one_turn_jacobian_matches_symplectic_oracle(ctx, ring):
    """The device's one-turn map has the oracle's Jacobian: central
    differences of device tracking, Richardson-extrapolated over the steps h
    and h/2 ((4 D(h/2) - D(h)) / 3, truncation O(h^4); the octupoles of the
    large ring make O(h^2) too coarse), against the oracle's exact
    dual-number Jacobian of the same map, which is symplectic to 1e-12
    (test_oracle_invariants).  What remains is the difference formula's
    truncation and rounding, and the device Jacobian is symplectic to the
    same order.  The step is large (2e-5) because z after a turn is a small
    difference of path lengths of metres, rounded to ~1e-12 m: the oracle's
    own map differenced the same way agrees with its exact Jacobian to
    2e-8 at this step and to 2e-6 at h = 2e-7."""
    if ring == "c0":
        eles = lattice.fodo_ring_c0()
        coords, _ = beam.c0_bunch(eles[:7], n=3, seed=9)
    else:
        eles = lattice.ring_c1()
        tw = beam.periodic_twiss(lattice.ring_c1(n_rf=0)[:8])
        coords = beam.matched_gaussian(3, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed=9)
    h = 2e-5
    for i in range(coords.shape[1]):
        v = coords[:, i]
        _, J = oracle.jacobian(eles, abi.ELECTRON, 1, v)
        pts = np.repeat(v[:, None], 24, axis=1)
        for k in range(6):
            for j, step in enumerate((h, h / 2)):
                pts[k, 12 * j + 2 * k] += step
                pts[k, 12 * j + 2 * k + 1] -= step
        g = _gpu_track(ctx, eles, abi.ELECTRON, pts, 1)
        assert np.all(g["state"] == abi.ALIVE)
        out = np.array([g[c] for c in ("x", "px", "y", "py", "z", "pz")])
        d_h = (out[:, 0:12:2] - out[:, 1:12:2]) / (2 * h)
        d_h2 = (out[:, 12::2] - out[:, 13::2]) / h
        Jd = (4 * d_h2 - d_h) / 3
        err = np.max(np.abs(Jd - J)) / np.max(np.abs(J))
        S = np.zeros((6, 6))
        for k in range(3):
            S[2 * k, 2 * k + 1], S[2 * k + 1, 2 * k] = 1.0, -1.0
        symp = np.max(np.abs(Jd.T @ S @ Jd - S))
        print(f"jacobian {ring} particle {i}: |J_dev - J_oracle| / max|J| = {err:.2e}, "
              f"symplectic error oracle {oracle.symp_error(J):.2e} device {symp:.2e}", flush=True)
        assert oracle.symp_error(J) <= 1e-12
        assert err <= 1e-7
        assert symp <= 1e-6


def test_one_turn_matrix_and_twiss_from_tracking(ctx):
    """bmx_one_turn_matrix (central differences of device tracking, steps
    d_orb = 1e-5 as twiss_from_tracking takes them), for three ring variants
    of different length in one call: against the oracle differenced with the
    same 12 probes (the same matrix to rounding), and against the oracle's
    exact one-turn Jacobian to the truncation of a central difference at that
    step.  Then the periodic Twiss from the device matrix (twiss_from_mat2 of
    each plane): equal to rounding to the Twiss of the oracle's same-probe
    matrix, and within the central difference's truncation (5e-6 on this
    ring, from the sextupoles) of the exact one, which the linear-optics
    model of the cell reproduces to 1e-14."""
    vs = lattice.da_variants(n_var=3, min_cells=6, max_cells=14)
    offs = [0]
    for v in vs:
        offs.append(offs[-1] + len(v))
    lat = ctx.lattice([e for v in vs for e in v], abi.ELECTRON, variant_offsets=offs)
    m, status = tracker.one_turn_matrix(lat)
    assert np.all(status == 0)
    d = 1e-5
    probes = np.zeros((6, 12))
    for q in range(6):
        probes[q, 2 * q], probes[q, 2 * q + 1] = d, -d
    for v, eles in enumerate(vs):
        _, J = oracle.jacobian(eles, abi.ELECTRON, 1, np.zeros(6))
        r = oracle.track_bunch(eles, abi.ELECTRON, 1, probes)
        o = np.array([r[c] for c in ("x", "px", "y", "py", "z", "pz")])
        m_fd = (o[:, 0::2] - o[:, 1::2]) / (2 * d)
        scale = np.max(np.abs(J))
        err_fd = np.max(np.abs(m[v] - m_fd)) / scale
        err = np.max(np.abs(m[v] - J)) / scale
        print(f"one-turn matrix variant {v}: vs oracle same probes {err_fd:.2e}, vs exact Jacobian {err:.2e}, "
              f"symplectic error {oracle.symp_error(m[v]):.2e}", flush=True)
        assert err_fd <= 1e-9
        assert err <= 1e-6
        assert oracle.symp_error(m[v]) <= 1e-5
    lat.close()
    eles = lattice.fodo_ring_c0()
    lat = ctx.lattice(eles, abi.ELECTRON)
    tw, M = beam.twiss_from_tracking(lat)
    lin = beam.periodic_twiss(eles)
    r = oracle.track_bunch(eles, abi.ELECTRON, 1, probes)
    o = np.array([r[c] for c in ("x", "px", "y", "py", "z", "pz")])
    m_fd = (o[:, 0::2] - o[:, 1::2]) / (2 * d)
    tw_fd = [*beam.twiss_from_mat2(m_fd[0:2, 0:2])[:2], *beam.twiss_from_mat2(m_fd[2:4, 2:4])[:2]]
    print("twiss from tracking:", np.round(tw, 8), "linear model:", np.round(lin, 8), flush=True)
    np.testing.assert_allclose(tw[:4], tw_fd, rtol=1e-9)
    np.testing.assert_allclose(tw, lin, rtol=1e-5)
    lat.close()


def test_host_buffer_path(ctx):
    eles = lattice.fodo_ring_c0(n_cells=5)
    coords = _coords(1000, 4, [1e-3, 1e-4, 1e-3, 1e-4, 1e-3, 1e-3])
    lat = ctx.lattice(eles, abi.ELECTRON)
    g = tracker.track_bunch_host(lat, coords, n_turns=2)
    r = oracle.track_bunch(eles, abi.ELECTRON, 2, coords)
    _compare(g, r, _path(eles, 2))
    # the backward program from host buffers, on a lattice that already ran forward
    lat_p = ctx.lattice(eles, abi.ELECTRON, tracked_species=abi.POSITRON)
    tracker.track_bunch_host(lat_p, coords, n_turns=1)
    g = tracker.track_bunch_host(lat_p, coords, n_turns=2, direction=-1)
    r = oracle.track_bunch(eles, abi.ELECTRON, 2, coords, tracked_species=abi.POSITRON, direction=-1)
    _compare(g, r, _path(eles, 2))


def test_grid_matches_oracle(ctx):
    """bmx_track_grid on three variants of uneven length vs the oracle tracking
    the same polar-grid initial conditions.  Survival turns must be identical
    outside the chaotic band, defined as the initial conditions whose oracle
    survival differs between two FP-legal evaluation orders (plain / FMA)."""
    vs = lattice.da_variants(n_var=3, min_cells=6, max_cells=14)
    offs = [0]
    for v in vs:
        offs.append(offs[-1] + len(v))
    flat = [e for v in vs for e in v]
    lat = ctx.lattice(flat, abi.ELECTRON, variant_offsets=offs)
    n_turns = 30
    surv, area, _ = tracker.track_grid(lat, n_turns, n_angle=8, n_amp=16, n_pz=3, amp_max_x=0.02, amp_ratio_y=0.5,
                                       pz_min=-0.01, pz_max=0.01)
    n_band = 0
    for iv, v in enumerate(vs):
        ics = []
        for ipz in range(3):
            pz = -0.01 + 0.01 * ipz
            for ia in range(8):
                th = math.pi * ia / 7
                for k in range(16):
                    amp = 0.02 * (k + 1) / 16
                    ics.append([amp * math.cos(th), 0, amp * 0.5 * math.sin(th), 0, 0, pz])
        c = np.array(ics).T
        r = oracle.track_bunch(v, abi.ELECTRON, n_turns, c, n_threads=8)
        r2 = oracle.track_bunch(v, abi.ELECTRON, n_turns, c, n_threads=8, fma=True)
        s_ref = np.where(r["state"] == abi.ALIVE, n_turns, r["turn_lost"])
        s_ref2 = np.where(r2["state"] == abi.ALIVE, n_turns, r2["turn_lost"])
        band = s_ref != s_ref2
        n_band += int(band.sum())
        s_gpu = surv[iv].reshape(-1)
        mism = np.flatnonzero((s_ref != s_gpu) & ~band)
        assert len(mism) == 0, f"variant {iv}: survival mismatch outside the chaotic band at {mism[:10]}"
        assert 0 < np.sum(s_ref < n_turns) < len(s_ref), "grid should straddle the aperture"
    print(f"grid: chaotic band {n_band} of {3 * c.shape[1]} initial conditions", flush=True)
    assert np.all(area > 0)


@pytest.mark.parametrize("ring", ["c1", "c1_no_rf", "variants"])
def test_cell_kernels_match_interpreter(ctx, ring, monkeypatch):
    """The runtime-compiled cell kernels (the lattice's repeating cell tracked
    straight through, op types, flags and counts as constants) against the op
    interpreter (BMX_JIT=0): the same op functions in the same order, so loss
    states, elements and turns are identical and coordinates differ only by
    the rounding of a different FMA contraction of the unrolled code, held to
    the 1e-12 bar of the oracle parity tests (scales as there).  Bunch
    tracking on the 4000-element ring at 30x the matched beam size (a large
    part of the particles lost), with and without rf, one turn (a second turn
    with rf at this amplitude already amplifies last-digit differences to
    ~1e-12); survival turns of the polar grid on three error-seeded variants,
    whose cells fold the offset unset into the element bodies."""
    if ring == "variants":
        vs = lattice.da_variants(n_var=3, min_cells=40, max_cells=60, seed=5)
        offs = np.cumsum([0] + [len(v) for v in vs]).tolist()
        eles = [e for v in vs for e in v]
        grid = dict(n_angle=8, n_amp=16, n_pz=2, amp_max_x=0.012, amp_ratio_y=0.5, pz_min=-1e-3, pz_max=1e-3)
        runs = {}
        for mode in ("1", "0"):
            monkeypatch.setenv("BMX_JIT", mode)
            lat = ctx.lattice(eles, abi.ELECTRON, variant_offsets=offs)
            cells = lat.cells()
            print(f"BMX_JIT={mode}: cells {cells}", flush=True)
            assert (cells[0] > 0) == (mode == "1")
            runs[mode] = tracker.track_grid(lat, n_turns=64, **grid)[0]
            lat.close()
        assert np.array_equal(runs["1"], runs["0"])
        return
    eles = lattice.ring_c1(n_cells=500, n_rf=0 if ring == "c1_no_rf" else 8)
    tw = beam.periodic_twiss(lattice.ring_c1(n_cells=500, n_rf=0)[:8])
    coords = beam.matched_gaussian(100_000, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed=11) * 30.0
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BMX_JIT", mode)
        lat = ctx.lattice(eles, abi.ELECTRON)
        cells = lat.cells()
        print(f"BMX_JIT={mode}: cells {cells}", flush=True)
        assert (cells[0] > 0) == (mode == "1")
        b = tracker.Bunch(coords, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
        tracker.track_bunch(lat, b, n_turns=1)
        out[mode] = b.numpy()
        lat.close()
    r = out["0"]
    n_lost = int(np.sum(r["state"] != abi.ALIVE))
    print(f"lost {n_lost} of {coords.shape[1]}", flush=True)
    assert 0 < n_lost < coords.shape[1]
    for k in ("state", "ix_lost", "turn_lost"):
        assert np.array_equal(out["1"][k], r[k]), k
    errs = coord_errors(out["1"], r, r["state"] == abi.ALIVE, _path(eles, 1))
    lost = r["state"] != abi.ALIVE
    errs_lost = coord_errors(out["1"], r, lost, _path(eles, 1))
    print("cell vs interpreter, alive:", {k: f"{v:.1e}" for k, v in errs.items()},
          "lost:", {k: f"{v:.1e}" for k, v in errs_lost.items()}, flush=True)
    assert max(errs.values()) <= TOL and max(errs_lost.values()) <= TOL


@pytest.mark.parametrize("case", ["near_zero_orbit", "k1_small_argument", "low_energy_z", "sincc_cutoff"])
def test_branch_switches_on_the_device(ctx, case):
    """The device on both sides of the reference's branch switches, against
    the oracle (whose two sides tests/test_oracle_invariants.py pins against
    the mathematics): the sbend's near-zero-orbit map (track_a_bend.f90:
    136-137), sbend_body_with_k1_map's small-argument branch (:546-559, with
    its z2 sign), low_energy_z_correction's series (low_energy_z_correction.f90:
    44, protons at 1 GeV) and the sincc cut-off (sincc.f90, bend angle per
    step across 0.1)."""
    species, e_tot = abi.ELECTRON, 1e9
    if case == "near_zero_orbit":
        variants = [lattice.ele("sbend", l=1.0, g=0.01, fringe_type=abi.FRINGE_NONE, num_steps=1)]
        rng = np.random.default_rng(31)
        coords = rng.uniform(-1.0, 1.0, (6, 4096)) * np.array([1e-9, 2e-9, 1e-9, 2e-9, 1e-9, 2e-9])[:, None]
    elif case == "k1_small_argument":
        variants = [lattice.ele("sbend", l=1.0, g=0.01, k1=-1e-4 + 1e-12 * f, fringe_type=abi.FRINGE_NONE,
                                num_steps=1) for f in (0.999, 1.001)]
        coords = _coords(4096, 37, SCALE)
        coords[5] = 0.0
    elif case == "low_energy_z":
        species = abi.PROTON
        variants = [lattice.ele("quadrupole", l=0.5, k1=0.3, num_steps=1)]
        thr = 1.6347376651828182e-3
        rng = np.random.default_rng(41)
        coords = _coords(4096, 41, [1e-3, 1e-4, 1e-3, 1e-4, 1e-3, 0.0])
        coords[5] = thr * rng.uniform(0.98, 1.02, 4096)
    else:
        variants = [lattice.ele("sbend", l=1.0, g=a, num_steps=1, fringe_type=abi.FRINGE_NONE) for a in (0.0999, 0.1001)]
        coords = _coords(4096, 43, SCALE)
    for e in variants:
        eles = lattice.set_reference([e], species, e_tot=e_tot)
        g = _gpu_track(ctx, eles, species, coords, 1)
        r = oracle.track_bunch(eles, species, 1, coords)
        # 1e-9 orbits: the exact body (taken by the larger of them) works in
        # units of 1/g, so its rounding is absolute, not relative to the orbit;
        # the bar there is the oracle's own FP-order spread
        tol = _spread_tol(eles, 1, coords) if case == "near_zero_orbit" else TOL
        _compare(g, r, _path(eles, 1), tol=tol)
