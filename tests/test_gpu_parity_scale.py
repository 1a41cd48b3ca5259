"""Parity with the oracle at the sizes the workloads run at, over many turns.

* the 100-cell FODO ring with 10,000 matched particles for its full 100
  turns, and the same bunch at 8x its size, where about 45 % is lost:
  states, loss elements and loss turns identical outside the chaotic band
  (particles whose oracle survival differs between the plain and the
  FMA-contracted oracle builds, counted), coordinates within 4x the oracle's
  own plain-vs-FMA spread;
* the batched grid kernel on a 4000-element error-seeded variant with 8 rf
  cavities for 1024 turns, on a sub-sampled polar grid (8 angles x 16
  amplitudes up to 12 sigma_x, pz -2, 0, +2 %): survival turns identical
  outside the chaotic band.  The variant has the configured quadrupole k1
  and sextupole k2 errors with 5 um offsets: with 50 um offsets every
  4000-element variant loses every initial condition within two turns;
  with 5 um the survival turns spread over the whole 1024;
* z on the rf ring against the oracle's spread measured relative to max|z|
  itself (no path-length floor).
"""
import numpy as np
import pytest

from bmx import abi, beam, lattice, tracker
import oracle
from helpers import coord_errors

pytestmark = pytest.mark.gpu


def _track(ctx, eles, coords, n_turns):
    lat = ctx.lattice(eles, abi.ELECTRON)
    b = tracker.Bunch(coords, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    tracker.track_bunch(lat, b, n_turns=n_turns)
    out = b.numpy()
    lat.close()
    return out


def _survival(r, n_turns):
    return np.where(r["state"] == abi.ALIVE, n_turns, r["turn_lost"])


@pytest.mark.parametrize("scale", [1.0, 8.0])
def test_c0_ring_100_turns(ctx, scale):
    eles = lattice.fodo_ring_c0()
    coords, _ = beam.c0_bunch(eles[:7], n=10_000, seed=20261203)
    coords *= scale
    n_turns = 100
    print(f"c0 x{scale}: {len(eles)} elements, 10000 particles, {n_turns} turns", flush=True)
    g = _track(ctx, eles, coords, n_turns)
    r = oracle.track_bunch(eles, abi.ELECTRON, n_turns, coords, n_threads=16)
    f = oracle.track_bunch(eles, abi.ELECTRON, n_turns, coords, n_threads=16, fma=True)
    band = ((r["state"] != f["state"]) | (_survival(r, n_turns) != _survival(f, n_turns))
            | (r["ix_lost"] != f["ix_lost"]))
    out = ~band
    print(f"c0 x{scale}: lost {np.sum(r['state'] != abi.ALIVE)}, chaotic band {band.sum()}", flush=True)
    assert np.array_equal(g["state"][out], r["state"][out])
    lost = out & (r["state"] != abi.ALIVE)
    assert np.array_equal(g["ix_lost"][lost], r["ix_lost"][lost])
    assert np.array_equal(g["turn_lost"][lost], r["turn_lost"][lost])
    if scale > 1:
        assert lost.sum() > 1000
    alive = out & (r["state"] == abi.ALIVE) & (g["state"] == abi.ALIVE)
    path = n_turns * sum(e.l for e in eles)
    errs = coord_errors(g, r, alive, path)
    spread = coord_errors(f, r, alive, path)
    tol = max(1e-12, 4 * max(spread.values()))
    print("c0 parity:", {k: f"{v:.1e}" for k, v in errs.items()}, f"bar {tol:.1e}", flush=True)
    assert max(errs.values()) <= tol


def test_grid_4000_element_variant_1024_turns(ctx):
    v = lattice.da_variants(n_var=1, min_cells=500, max_cells=500, seed=20261205, offsets_scale=0.1)[0]
    assert len(v) == 4000 and sum(e.key == abi.KEY_RFCAVITY for e in v) == 8
    tw = beam.periodic_twiss(lattice.ring_c1(n_cells=500, n_rf=0)[:8])
    amp_max = 12.0 * (10e-9 * tw[0]) ** 0.5
    n_angle, n_amp, n_pz, n_turns = 8, 16, 3, 1024
    lat = ctx.lattice(v, abi.ELECTRON)
    print(f"grid: 4000 elements, {n_angle}x{n_amp}x{n_pz} ICs, {n_turns} turns", flush=True)
    surv, area, _ = tracker.track_grid(lat, n_turns, n_angle, n_amp, n_pz, amp_max, 1.0, -0.02, 0.02)
    lat.close()
    th = np.linspace(0.0, np.pi, n_angle)
    a = amp_max * (np.arange(n_amp) + 1) / n_amp
    pz = np.linspace(-0.02, 0.02, n_pz)
    c = np.zeros((6, n_pz, n_angle, n_amp))
    c[0] = np.cos(th)[None, :, None] * a[None, None, :]
    c[2] = np.sin(th)[None, :, None] * a[None, None, :]
    c[5] = pz[:, None, None]
    c = c.reshape(6, -1)
    r = oracle.track_bunch(v, abi.ELECTRON, n_turns, c, n_threads=16)
    f = oracle.track_bunch(v, abi.ELECTRON, n_turns, c, n_threads=16, fma=True)
    s_ref, s_fma = _survival(r, n_turns), _survival(f, n_turns)
    band = s_ref != s_fma
    s_gpu = surv.reshape(-1)
    mism = np.flatnonzero((s_gpu != s_ref) & ~band)
    print(f"grid: {len(np.unique(s_ref))} distinct survival turns, up to {s_ref.max()}; chaotic band {band.sum()}",
          flush=True)
    assert len(mism) == 0, (mism[:10], s_gpu[mism[:10]], s_ref[mism[:10]])
    assert len(np.unique(s_ref)) > 50 and np.sum(s_ref >= 400) > 0


def test_c1_z_against_oracle_spread(ctx):
    """z after one turn of the rf ring, relative to max|z| itself: within 4x
    the oracle's own plain-vs-FMA spread on the same scale (and 1e-12)."""
    eles = lattice.ring_c1(n_cells=500)
    tw = beam.periodic_twiss(lattice.ring_c1(n_cells=500, n_rf=0)[:8])
    coords = beam.matched_gaussian(2048, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed=3)
    g = _track(ctx, eles, coords, 1)
    r = oracle.track_bunch(eles, abi.ELECTRON, 1, coords, n_threads=16)
    f = oracle.track_bunch(eles, abi.ELECTRON, 1, coords, n_threads=16, fma=True)
    alive = r["state"] == abi.ALIVE
    zmax = np.max(np.abs(r["z"][alive]))
    err = np.max(np.abs(g["z"][alive] - r["z"][alive])) / zmax
    spread = np.max(np.abs(f["z"][alive] - r["z"][alive])) / zmax
    print(f"z relative to max|z|: device {err:.2e}, oracle spread {spread:.2e}", flush=True)
    assert err <= max(1e-12, 4 * spread)
