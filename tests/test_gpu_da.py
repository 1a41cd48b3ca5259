"""GPU tests of the multi-variant entry points (bmx_track_multi, bmx_track_ics,
bmx_closed_orbit, bmx_track_grid with reference orbits) and of the batched
dynamic-aperture bisection, against the CPU oracle on the same inputs.

Coordinates after one turn: 1e-12 relative per coordinate as in test_gpu_parity.  Survival
turns and DA points: identical outside the chaotic band (initial conditions
whose oracle survival differs between the plain and FMA oracle builds).
Closed orbits: the oracle's one-turn map returns the GPU's closed orbit to
itself within 1e-13 m (1e-13 in px, py).
"""
import math

import numpy as np
import pytest

from bmx import abi, beam, da, lattice, tracker
import oracle
from helpers import coord_errors

pytestmark = pytest.mark.gpu


def _variants(n_var=3, min_cells=6, max_cells=14, n_rf=0, **kw):
    vs = lattice.da_variants(n_var=n_var, min_cells=min_cells, max_cells=max_cells, n_rf=n_rf, **kw)
    offs = [0]
    for v in vs:
        offs.append(offs[-1] + len(v))
    return vs, offs


def _upload(ctx, vs, offs):
    return ctx.lattice([e for v in vs for e in v], abi.ELECTRON, variant_offsets=offs)


def _oracle_survival(v, ics, n_turns, fma=False):
    r = oracle.track_bunch(v, abi.ELECTRON, n_turns, ics, n_threads=8, fma=fma)
    return np.where(r["state"] == abi.ALIVE, n_turns, r["turn_lost"])


def test_track_multi_matches_oracle(ctx):
    # variant lengths from the lower end of the batched-scan range (63+ cells)
    vs, offs = _variants(n_var=3, min_cells=63, max_cells=90)
    lat = _upload(ctx, vs, offs)
    rng = np.random.default_rng(12)
    coords = [rng.standard_normal((6, n)) * np.array([1e-3, 1e-4, 1e-3, 1e-4, 1e-3, 1e-3])[:, None]
              for n in (300, 0, 517)]
    res = tracker.track_multi(lat, coords, n_turns=1)
    for v, (c, g) in enumerate(zip(coords, res)):
        if c.shape[1] == 0:
            assert len(g["x"]) == 0
            continue
        r = oracle.track_bunch(vs[v], abi.ELECTRON, 1, c, n_threads=8)
        assert np.array_equal(g["state"], r["state"])
        alive = r["state"] == abi.ALIVE
        errs = coord_errors(g, r, alive, sum(e.l for e in vs[v]))
        print(f"variant {v}:", {k: f"{e:.1e}" for k, e in errs.items()}, flush=True)
        assert max(errs.values()) <= 1e-12
    lat.close()


@pytest.mark.parametrize("n_dim,n_rf", [(4, 0), (6, 2)])
def test_closed_orbit_is_fixed_point_of_oracle_map(ctx, n_dim, n_rf):
    vs, offs = _variants(n_var=3, min_cells=20, max_cells=40, n_rf=n_rf, offsets_scale=0.2)
    lat = _upload(ctx, vs, offs)
    guess = np.zeros((3, 6))
    if n_dim == 4:
        guess[:, 5] = 1e-3
    orb, status = tracker.closed_orbit(lat, n_dim, guess)
    print("closed orbits:", orb, status, flush=True)
    assert np.all(status == 0)
    assert np.max(np.abs(orb[:, 0])) > 1e-6, "misaligned variants should have a non-zero closed orbit"
    for v in range(3):
        r = oracle.track_bunch(vs[v], abi.ELECTRON, 1, orb[v][:, None])
        back = np.array([r[k][0] for k in ("x", "px", "y", "py", "z", "pz")])
        res = np.abs(back[:n_dim] - orb[v][:n_dim])
        print(f"variant {v}: |T(x) - x| = {res}", flush=True)
        assert np.all(res < 1e-13)
    lat.close()


def test_track_ics_matches_oracle(ctx):
    vs, offs = _variants(n_var=2)
    lat = _upload(ctx, vs, offs)
    rng = np.random.default_rng(4)
    ics = []
    for v in range(2):
        c = np.zeros((6, 400))
        a = rng.uniform(0, 0.02, 400)
        th = rng.uniform(0, math.pi, 400)
        c[0], c[2], c[5] = a * np.cos(th), 0.5 * a * np.sin(th), rng.uniform(-0.01, 0.01, 400)
        ics.append(c)
    n_turns = 40
    surv = tracker.track_ics(lat, ics, n_turns, turns_per_block=7)
    n_band = 0
    for v in range(2):
        s1 = _oracle_survival(vs[v], ics[v], n_turns)
        s2 = _oracle_survival(vs[v], ics[v], n_turns, fma=True)
        band = s1 != s2
        n_band += band.sum()
        assert np.array_equal(surv[v][~band], s1[~band])
        assert 0 < np.sum(s1 < n_turns) < len(s1)
    print(f"chaotic band {n_band} of 800", flush=True)
    lat.close()


def test_grid_with_reference_orbit(ctx):
    vs, offs = _variants(n_var=2, min_cells=30, max_cells=40, offsets_scale=0.2)
    lat = _upload(ctx, vs, offs)
    ref, status = da.reference_orbits(lat, [-0.001, 0.0, 0.001], rf_on=False)
    assert np.all(status == 0)
    n_turns = 25
    surv, area, _ = tracker.track_grid(lat, n_turns, n_angle=5, n_amp=12, n_pz=3, amp_max_x=0.02, amp_ratio_y=0.5,
                                       pz_min=-0.001, pz_max=0.001, ref_orbit=ref)
    for v in range(2):
        ics = []
        for ipz in range(3):
            for ia in range(5):
                th = math.pi * ia / 4
                for k in range(12):
                    a = 0.02 * (k + 1) / 12
                    o = ref[v, ipz].copy()
                    o[0] += a * math.cos(th)
                    o[2] += a * 0.5 * math.sin(th)
                    ics.append(o)
        c = np.array(ics).T
        s1 = _oracle_survival(vs[v], c, n_turns)
        s2 = _oracle_survival(vs[v], c, n_turns, fma=True)
        band = s1 != s2
        g = surv[v].reshape(-1)
        assert np.array_equal(g[~band], s1[~band])
    lat.close()


def test_dynamic_aperture_scan_matches_oracle_bisection(ctx):
    """The batched bisection with GPU survival gives the same DA points as the
    same bisection driven by the oracle, ray by ray, outside rays whose
    decisions hit the chaotic band."""
    vs, offs = _variants(n_var=2, min_cells=30, max_cells=40, offsets_scale=0.2)
    lat = _upload(ctx, vs, offs)
    ref, status = da.reference_orbits(lat, [0.0, 0.001], rf_on=False)
    assert np.all(status == 0)
    n_turn = 30

    def oracle_track(fma):
        def tr(ics, idx_var):
            out = np.zeros(ics.shape[1], dtype=np.int64)
            for v in range(len(vs)):
                sel = idx_var == v
                if sel.any():
                    out[sel] = _oracle_survival(vs[v], ics[:, sel], n_turn, fma=fma)
            return out
        return tr

    kw = dict(n_turn=n_turn, n_angle=5, rel_accuracy=1e-2, abs_accuracy=1e-5)
    g = da.dynamic_aperture_scan(lat, ref, **kw)
    o1 = da.dynamic_aperture_scan(lat, ref, track=oracle_track(False), **kw)
    o2 = da.dynamic_aperture_scan(lat, ref, track=oracle_track(True), **kw)
    band = (o1["x"] != o2["x"]) | (o1["y"] != o2["y"])
    print(f"DA points: chaotic rays {band.sum()} of {band.size}", flush=True)
    print("x:", g["x"][0, 0], "y:", g["y"][0, 0], flush=True)
    for k in ("x", "y", "i_turn"):
        assert np.array_equal(g[k][~band], o1[k][~band]), k
    assert np.all(g["x"][:, :, 0] > 1e-4)
    lat.close()


def test_graph_turn_blocks_match_eager(ctx):
    """Survival from hipGraph-captured turn blocks (several replays, with
    re-capture as the live count falls) equals block-by-block launches, for
    several schedules, on the grid and on explicit initial conditions."""
    vs, offs = _variants(n_var=3, min_cells=6, max_cells=14)
    lat = _upload(ctx, vs, offs)
    kw = dict(n_angle=8, n_amp=16, n_pz=3, amp_max_x=0.02, amp_ratio_y=0.5, pz_min=-0.005, pz_max=0.005)
    saved = ctx.schedule()
    results = []
    for tpb, depth in [(4, 1), (4, 8), (16, 8), (32, 2), (7, 4)]:
        ctx.set_schedule(tpb, depth)
        surv, area, _ = tracker.track_grid(lat, 60, turns_per_block=0, **kw)
        results.append((surv, area))
        print(f"schedule ({tpb}, {depth}): alive {np.sum(surv >= 60)}", flush=True)
    for surv, area in results[1:]:
        assert np.array_equal(surv, results[0][0])
        assert np.array_equal(area, results[0][1])
    rng = np.random.default_rng(8)
    ics = [np.vstack([rng.uniform(-0.01, 0.01, 300), np.zeros(300), rng.uniform(-0.005, 0.005, 300), np.zeros(300),
                      np.zeros(300), rng.uniform(-0.005, 0.005, 300)]) for _ in range(3)]
    ctx.set_schedule(5, 1)
    eager = tracker.track_ics(lat, ics, 45)
    ctx.set_schedule(5, 6)
    graph = tracker.track_ics(lat, ics, 45)
    for a, b in zip(eager, graph):
        assert np.array_equal(a, b)
    ctx.set_schedule(saved["turns_per_block"], saved["graph_depth"])
    lat.close()


def test_tune_grid_picks_and_stores_a_schedule(ctx):
    vs, offs = _variants(n_var=2, min_cells=6, max_cells=10)
    lat = _upload(ctx, vs, offs)
    kw = dict(n_angle=8, n_amp=16, n_pz=3, amp_max_x=0.02, amp_ratio_y=0.5, pz_min=-0.005, pz_max=0.005)
    saved = ctx.schedule()
    best = tracker.tune_grid(lat, 40, **kw)
    print("tuned:", best, flush=True)
    assert best["turns_per_block"] in (4, 8, 16, 32, 64, 128) and best["graph_depth"] in (1, 8)
    assert best["probe_pet_per_s"] > 0
    assert ctx.schedule()["turns_per_block"] == best["turns_per_block"]
    tuned, _, _ = tracker.track_grid(lat, 40, turns_per_block=0, **kw)
    ctx.set_schedule(32, 1)
    ref, _, _ = tracker.track_grid(lat, 40, turns_per_block=0, **kw)
    assert np.array_equal(tuned, ref)
    ctx.set_schedule(saved["turns_per_block"], saved["graph_depth"])
    lat.close()


def test_scaled_upload_equals_built_variants(ctx):
    """bmx_lattice_upload_scaled (knob variants scaled in the library) tracks
    exactly as bmx_lattice_upload of the variants built one by one
    (surrogate.knob_lattice on the error-seeded rf ring), over 20 turns."""
    from bmx import surrogate
    spec = surrogate.ScanSpec(n_cells=16, n_rf=8, error_seed=77)
    knobs = np.random.default_rng(12).uniform(-0.2, 0.2, (5, surrogate.N_KNOBS))
    lat_a = surrogate.upload_variants(ctx, knobs, spec)
    built = [surrogate.knob_lattice(k, spec) for k in knobs]
    offs = [0]
    for b in built:
        offs.append(offs[-1] + len(b))
    lat_b = ctx.lattice([e for b in built for e in b], abi.ELECTRON, variant_offsets=offs)
    rng = np.random.default_rng(13)
    coords = [rng.standard_normal((6, 512)) * np.array([2e-3, 2e-4, 2e-3, 2e-4, 2e-3, 5e-3])[:, None]
              for _ in knobs]
    a = tracker.track_multi(lat_a, coords, n_turns=20)
    b = tracker.track_multi(lat_b, coords, n_turns=20)
    lat_a.close()
    lat_b.close()
    for va, vb in zip(a, b):
        for k in ("x", "px", "y", "py", "z", "pz", "state", "ix_lost", "turn_lost"):
            assert np.array_equal(va[k], vb[k]), k
