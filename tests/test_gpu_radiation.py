"""Radiation damping and fluctuations on the device (OP_RAD, the radiation
steps of track1: bmad/code/track1.f90:198-201, :316-319; track1_radiation,
bmad/modules/radiation_mod.f90:62-141) against the oracle's restatement,
with the maps of the oracle's radiation_map_setup (pinned in
test_oracle_radiation.py).

* damping alone on the 100-cell FODO ring: every coordinate of 4096
  particles within the parity bar after 1 and 20 turns, states identical;
* fluctuations with damping: the device draws the same Philox deviates as
  the oracle (particle, element and edge, turn, call), so the tracks agree
  deviate by deviate to the same bar; successive calls draw fresh deviates
  and bmx_lattice_set_rad_stream replays a call;
* the runtime-compiled cell kernels against the op interpreter, the
  no-allocation entry point, multi-variant tracking;
* the equilibrium of a 6 GeV rf ring tracked for 60,000 turns: the vertical
  emittance (photon opening angle) and the energy spread against the
  equilibrium of the one-turn map with damping and the maps' stochastic
  covariance (statistical, 16,384 particles);
* closed_orbit_test's SLS-2 rings: bmx_closed_orbit in six dimensions from
  the zero orbit with damping, maps from bmx.radiation, at the z the test
  records, and the synchrotron tune of bmx_transfer_matrix about it;
* what is refused: backward tracking, fluctuations in the survival scans.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

from bmx import abi, beam, lattice, tracker
import oracle
from helpers import NAMES, coord_errors, ring_from_fixture

pytestmark = pytest.mark.gpu


def _no_apertures(eles):
    for e in eles:
        e.x1_limit = e.x2_limit = e.y1_limit = e.y2_limit = 0.0
    return eles


@pytest.fixture(scope="module")
def c0():
    eles = lattice.fodo_ring_c0(n_cells=100)
    maps = oracle.rad_maps(eles, abi.ELECTRON, np.zeros((len(eles), 6)))
    coords, _ = beam.c0_bunch(eles[:7], n=4096, seed=11)
    return eles, maps, coords * 3.0


def _gpu(ctx, eles, maps, coords, n_turns, rad, calls=1):
    lat = ctx.lattice(eles, abi.ELECTRON, rad_maps=maps, radiation=rad)
    b = tracker.Bunch(coords, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    for _ in range(calls):
        tracker.track_bunch(lat, b, n_turns=n_turns)
    out = b.numpy()
    return lat, out


def _bar(r, f, mask, path):
    """max(1e-12, 4x the oracle's own plain-vs-FMA spread) per coordinate."""
    spread = coord_errors(f, r, mask, path)
    return {k: max(1e-12, 4.0 * spread[k]) for k in NAMES}, spread


@pytest.mark.parametrize("n_turns", [1, 20])
def test_damping_matches_oracle(ctx, c0, n_turns):
    eles, maps, coords = c0
    print(f"damping: c0 ring {len(eles)} elements, 4096 particles, {n_turns} turns", flush=True)
    lat, g = _gpu(ctx, eles, maps, coords, n_turns, dict(damping_on=True))
    r = oracle.track_bunch_rad(eles, abi.ELECTRON, n_turns, coords, maps, n_threads=16)
    f = oracle.track_bunch_rad(eles, abi.ELECTRON, n_turns, coords, maps, n_threads=16, fma=True)
    assert np.array_equal(g["state"], r["state"])
    alive = r["state"] == abi.ALIVE
    path = sum(e.l for e in eles) * n_turns
    bar, spread = _bar(r, f, alive, path)
    err = coord_errors(g, r, alive, path)
    print(f"damping {n_turns} turns: err {err} bar {bar}", flush=True)
    for k in NAMES:
        assert err[k] <= bar[k], (k, err[k], bar[k])
    # the radiation is there: the bunch lost energy against tracking without it
    plain = oracle.track_bunch(eles, abi.ELECTRON, n_turns, coords, n_threads=16)
    assert np.mean(g["pz"][alive] - plain["pz"][alive]) < -0.5 * n_turns * 6.9e-7
    lat.close()


@pytest.mark.parametrize("n_turns", [1, 5])
def test_fluctuations_match_oracle(ctx, c0, n_turns):
    eles, maps, coords = c0
    seed = 0x1234567890ABCDEF
    print(f"fluctuations: c0 ring, 4096 particles, {n_turns} turns", flush=True)
    lat, g = _gpu(ctx, eles, maps, coords, n_turns, dict(damping_on=True, fluctuations_on=True, seed=seed))
    r = oracle.track_bunch_rad(eles, abi.ELECTRON, n_turns, coords, maps, fluctuations_on=True, seed=seed, call=0,
                               n_threads=16)
    f = oracle.track_bunch_rad(eles, abi.ELECTRON, n_turns, coords, maps, fluctuations_on=True, seed=seed, call=0,
                               n_threads=16, fma=True)
    assert np.array_equal(g["state"], r["state"])
    alive = r["state"] == abi.ALIVE
    path = sum(e.l for e in eles) * n_turns
    bar, spread = _bar(r, f, alive, path)
    err = coord_errors(g, r, alive, path)
    print(f"fluctuations {n_turns} turns: err {err} bar {bar}", flush=True)
    for k in NAMES:
        assert err[k] <= bar[k], (k, err[k], bar[k])
    # and the kicks are not the damping alone: another seed gives other tracks
    r2 = oracle.track_bunch_rad(eles, abi.ELECTRON, n_turns, coords, maps, fluctuations_on=True, seed=seed + 1,
                                call=0, n_threads=16)
    assert np.max(np.abs(r2["pz"] - r["pz"])) > 1e3 * err["pz"] * np.max(np.abs(r["pz"]))
    lat.close()


def test_stream_advances_per_call_and_replays(ctx, c0):
    eles, maps, coords = c0
    seed = 77
    lat = ctx.lattice(eles, abi.ELECTRON, rad_maps=maps,
                      radiation=dict(damping_on=False, fluctuations_on=True, seed=seed))
    outs = []
    for _ in range(2):
        b = tracker.Bunch(coords, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
        tracker.track_bunch(lat, b, n_turns=1)
        outs.append(b.numpy())
    lat.set_rad_stream(seed, 1)
    b = tracker.Bunch(coords, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    tracker.track_bunch(lat, b, n_turns=1)
    again = b.numpy()
    assert not np.array_equal(outs[0]["pz"], outs[1]["pz"])
    for k in NAMES:
        assert np.array_equal(again[k], outs[1][k]), k
    r1 = oracle.track_bunch_rad(eles, abi.ELECTRON, 1, coords, maps, damping_on=False, fluctuations_on=True, seed=seed,
                                call=1, n_threads=16)
    alive = r1["state"] == abi.ALIVE
    err = coord_errors(outs[1], r1, alive, sum(e.l for e in eles))
    assert max(err.values()) < 1e-9
    lat.close()


def test_cell_kernels_and_interpreter_agree(ctx, c0):
    eles, maps, coords = c0
    rad = dict(damping_on=True, fluctuations_on=True, seed=5)
    lat_c, g_c = _gpu(ctx, eles, maps, coords, 3, rad)
    cells = lat_c.cells()
    old = os.environ.get("BMX_JIT")
    os.environ["BMX_JIT"] = "0"
    try:
        lat_i, g_i = _gpu(ctx, eles, maps, coords, 3, rad)
    finally:
        if old is None:
            del os.environ["BMX_JIT"]
        else:
            os.environ["BMX_JIT"] = old
    assert cells[0] > 0 and lat_i.cells() == (0, 0)
    assert np.array_equal(g_c["state"], g_i["state"])
    alive = g_c["state"] == abi.ALIVE
    err = coord_errors(g_c, g_i, alive, 3 * sum(e.l for e in eles))
    print(f"cell kernels vs interpreter with radiation: {err}", flush=True)
    assert max(err.values()) < 1e-12
    lat_c.close()
    lat_i.close()


def test_noalloc_and_multi_variant(ctx, c0):
    eles, maps, coords = c0
    rad = dict(damping_on=True, fluctuations_on=True, seed=9)
    lat = ctx.lattice(eles, abi.ELECTRON, rad_maps=maps, radiation=rad)
    b1 = tracker.Bunch(coords, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    tracker.track_bunch(lat, b1, n_turns=2)
    lat.set_rad_stream(9, 0)
    b2 = tracker.Bunch(coords, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    import torch
    torch.cuda.synchronize()
    p = b2.particles_struct()
    st = abi.Stats()
    rc = ctx.lib.bmx_track_bunch_noalloc(ctx.h, lat.h, 0, 0, 2, 1, C.byref(p), C.byref(st))
    assert rc == 0, ctx.lib.bmx_last_error(ctx.h)
    o1, o2 = b1.numpy(), b2.numpy()
    for k in NAMES:
        assert np.array_equal(o1[k], o2[k]), k
    lat.close()
    # two variants (the ring and a copy) in one launch with damping: each as the ring alone
    n = len(eles)
    both = list(eles) + [lattice.EleRec_copy(e) for e in eles]
    lat2 = ctx.lattice(both, abi.ELECTRON, variant_offsets=[0, n, 2 * n],
                       rad_maps=np.concatenate([maps, maps]), radiation=dict(damping_on=True))
    res = tracker.track_multi(lat2, [coords[:, :1000], coords[:, 1000:3000]], n_turns=3)
    r = oracle.track_bunch_rad(eles, abi.ELECTRON, 3, coords[:, :3000], maps, n_threads=16)
    for v, sl in ((0, slice(0, 1000)), (1, slice(1000, 3000))):
        rv = {k: r[k][sl] for k in list(NAMES) + ["state"]}
        assert np.array_equal(res[v]["state"], rv["state"])
        alive = rv["state"] == abi.ALIVE
        err = coord_errors(res[v], rv, alive, 3 * sum(e.l for e in eles))
        assert max(err.values()) < 1e-10, err
    lat2.close()


def test_refusals(ctx, c0):
    eles, maps, coords = c0
    lat = ctx.lattice(eles, abi.ELECTRON, rad_maps=maps, radiation=dict(damping_on=True, fluctuations_on=True))
    b = tracker.Bunch(coords[:, :64], lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    with pytest.raises(RuntimeError, match="forward only"):
        tracker.track_bunch(lat, b, n_turns=1, direction=-1)
    with pytest.raises(RuntimeError, match="fluctuations"):
        tracker.track_grid(lat, 4, n_angle=2, n_amp=4, n_pz=1, amp_max_x=1e-3)
    lat.close()
    # damping alone is accepted by the survival scans
    lat = ctx.lattice(eles, abi.ELECTRON, rad_maps=maps, radiation=dict(damping_on=True))
    surv, area, _ = tracker.track_grid(lat, 4, n_angle=2, n_amp=4, n_pz=1, amp_max_x=1e-3)
    assert np.all(np.asarray(surv) == 4)
    lat.close()


def test_equilibrium_emittance_and_energy_spread(ctx):
    """A 6 GeV ring with 8 rf cavities, radiation scaled by 0.1 (energy loss
    0.72 MeV against 8 MV of rf; the equilibrium emittances scale with it,
    the damping times inversely), 16,384 particles from the axis for 60,000
    turns (3.6 vertical damping times): the vertical emittance and the
    energy spread of the bunch against the equilibrium Sigma = M Sigma M^T +
    B of the one-turn map with damping and the maps' stochastic covariance."""
    scale = 0.1
    eles = _no_apertures(lattice.ring_c1(n_cells=50, n_rf=8))
    maps = oracle.rad_maps(eles, abi.ELECTRON, np.zeros((len(eles), 6)))
    I = np.eye(6)
    Md, B = I.copy(), np.zeros((6, 6))
    for i, e in enumerate(eles):
        M = oracle.make_mat6(e, abi.ELECTRON, [0.0] * 6)[1]
        D0, D1 = scale * maps[i, 0, 6:42].reshape(6, 6), scale * maps[i, 1, 6:42].reshape(6, 6)
        S0, S1 = scale * maps[i, 0, 48:].reshape(6, 6), scale * maps[i, 1, 48:].reshape(6, 6)
        B = (I + D0) @ B @ (I + D0).T + S0 @ S0.T
        B = M @ B @ M.T
        B = (I + D1) @ B @ (I + D1).T + S1 @ S1.T
        Md = (I + D1) @ M @ (I + D0) @ Md
    Sig = sla.solve_discrete_lyapunov(Md, B)
    eps_y_eq = np.sqrt(np.linalg.det(Sig[2:4, 2:4]))
    sig_pz_eq = np.sqrt(Sig[5, 5])
    lat = ctx.lattice(eles, abi.ELECTRON, rad_maps=maps,
                      radiation=dict(damping_on=True, fluctuations_on=True, synch_rad_scale=scale, seed=2026))
    b = tracker.Bunch(np.zeros((6, 16384)), lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    for k in range(12):
        tracker.track_bunch(lat, b, n_turns=5000)
        print(f"equilibrium run: {5000 * (k + 1)} turns", flush=True)
    o = b.numpy()
    assert np.all(o["state"] == abi.ALIVE)
    cy = np.cov(np.vstack([o["y"], o["py"]]))
    eps_y = np.sqrt(np.linalg.det(cy))
    sig_pz = np.std(o["pz"])
    print(f"eps_y {eps_y:.4e} (equilibrium {eps_y_eq:.4e}), sigma_pz {sig_pz:.4e} ({sig_pz_eq:.4e})", flush=True)
    assert eps_y == pytest.approx(eps_y_eq, rel=0.06)
    assert sig_pz == pytest.approx(sig_pz_eq, rel=0.04)
    lat.close()


def test_design_orbit_maps_equal_the_oracles(ctx):
    """bmx.radiation.design_orbit_maps (the library's own transfer matrices)
    against the oracle's radiation_map_setup on the design orbit: the FODO
    ring's bends, the 6 GeV rf ring's, and gradient bends with Hwang and
    linear edges at e1, e2 != 0."""
    from bmx import radiation
    extra = [
        lattice.ele("drift", l=0.3),
        lattice.ele("sbend", l=3.0, angle=np.pi / 8, e1=0.07, e2=0.07, fringe_type=abi.FRINGE_BASIC_BEND, num_steps=1),
        lattice.ele("quadrupole", l=0.5, k1=0.6, num_steps=1),
        lattice.ele("sbend", l=2.0, angle=0.1, k1=-0.2, e1=0.05, e2=-0.03, fringe_type=abi.FRINGE_LINEAR_EDGE,
                    fint=0.5, hgap=0.02, num_steps=4),
        lattice.ele("sbend", l=1.5, angle=0.08, e1=0.02, fringe_type=abi.FRINGE_HARD_EDGE_ONLY,
                    fringe_at=abi.ENTRANCE_END, num_steps=1),
    ]
    lattice.set_reference(extra, abi.ELECTRON, e_tot=3e9)
    for name, eles in (("c0", lattice.fodo_ring_c0(n_cells=20)), ("rf ring", lattice.ring_c1(n_cells=50, n_rf=8)),
                       ("edges", extra)):
        mp = radiation.design_orbit_maps(ctx, eles, abi.ELECTRON)
        mo = oracle.rad_maps(eles, abi.ELECTRON, np.zeros((len(eles), 6)))
        n_bends = 0
        for i, e in enumerate(eles):
            for h in range(2):
                for lo, hi, what in ((6, 42, "damp_dmat"), (42, 48, "xfer_damp_vec"), (48, 84, "stoc_mat")):
                    a, b = mp[i, h, lo:hi], mo[i, h, lo:hi]
                    scale = np.max(np.abs(b))
                    if e.key != abi.KEY_SBEND:
                        assert np.all(a == 0.0) and scale == 0.0, (name, i, what)
                        continue
                    assert scale > 0, (name, i, h, what)
                    assert np.max(np.abs(a - b)) <= 1e-11 * scale, (name, i, h, what, np.max(np.abs(a - b)) / scale)
            n_bends += e.key == abi.KEY_SBEND
        print(f"design-orbit maps {name}: {n_bends} bends equal to the oracle's", flush=True)


with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "closed_orbit_ad05f.json")) as _f:
    CO = json.load(_f)


@pytest.mark.parametrize("name", sorted(CO["lattices"]))
def test_closed_orbit_test_rings(ctx, name):
    """closed_orbit_test (closed_orbit_test.f90:45-80, output.correct:4-13)
    on the device: the 3120-element SLS-2 ring with radiation damping, its
    maps from bmx.radiation.design_orbit_maps (the library's own transfer
    matrices), closed in six dimensions by bmx_closed_orbit from the zero
    orbit as closed_orbit_calc starts (for ad05f_pos_RF the nearest fixed
    point is the unstable one; the longitudinal stability search moves it to
    the stable one).  z matches the recorded value to its ABS 5e-8.  The
    synchrotron tune comes from bmx_transfer_matrix over one turn about that
    orbit: its derivatives leave the radiation kicks out, as make_mat6's
    element matrices do, so the one-turn matrix is transfer_matrix_calc's
    t1 (calc_z_tune.f90:36-59)."""
    from bmx import radiation
    eles = ring_from_fixture(CO, name)
    g2 = sum(e.g ** 2 * e.l for e in eles if e.key == abi.KEY_SBEND)
    g3 = sum(e.g ** 3 * e.l for e in eles if e.key == abi.KEY_SBEND)
    print(f"closed_orbit_test {name}: {len(eles)} elements", flush=True)
    maps = radiation.design_orbit_maps(ctx, eles, abi.ELECTRON, g2, g3)
    lat = ctx.lattice(eles, abi.ELECTRON, rad_maps=maps, radiation=dict(damping_on=1))
    orb, status = tracker.closed_orbit(lat, n_dim=6)
    assert status[0] == 0, status
    rows = CO["rows"]
    for tag in ("1-B", "2-A", "2-B"):
        r = rows[f"{name}-{tag} Z"]
        assert abs(orb[0, 4] - r["value"]) <= r["tol"], (tag, orb[0, 4], r["value"])
    out, mat, state = tracker.transfer_matrix(lat, orb[0])
    assert state[0] == abi.ALIVE
    assert np.abs(out[0] - orb[0]).max() < 1e-12
    a = mat[0]
    denom = 2 * (a[4, 4] * a[5, 5] - a[4, 5] * a[5, 4])
    assert (a[4, 4] + a[5, 5]) / denom - 1 < -1e-9
    ang = np.abs(np.angle(np.linalg.eigvals(a)))
    tune = np.sign(a[4, 5]) * ang.min() / (2 * np.pi)
    for tag in ("1-B", "2-B"):
        r = rows[f"{name}-{tag} Z_tune"]
        assert abs(tune - r["value"]) <= r["tol"], (tag, tune, r["value"])
    print(f"closed_orbit_test {name}: z {orb[0, 4]:.10f}, synchrotron tune {tune:.10f}", flush=True)
    lat.close()
