"""Host logic of the surrogate-screened DA scan (bmx.surrogate.surrogate_scan),
with the FP64 tracker and the network replaced by CPU stand-ins: a synthetic
survival field whose true boundary is known, and predictions off by a chosen
number of grid steps.  The band search with its extension fallback must
recover the true surviving-prefix length of every ray however far off the
prediction is, and track fewer initial conditions than the full grid when the
prediction is close.  (The GPU end-to-end run is in test_gpu_surrogate_da.py.)"""
import numpy as np
import pytest

from bmx import surrogate


class _FakeLat:
    def close(self):
        pass


def _run(monkeypatch, truth, pred_shift, islands=None, refine=0, **kw):
    spec = surrogate.ScanSpec(n_turns=100, n_angle=6, n_amp=40, n_pz=3)
    n_lat = truth.shape[0]
    N = spec.n_turns
    amp_k = np.arange(spec.n_amp)

    def surv_of(v, ip, ia, k):
        alive = k < truth[v, ip, ia]
        if islands is not None and (v, ip, ia, k) in islands:
            alive = False
        return N if alive else 7

    # predicted survival: alive below truth + shift; boundaries from its step fit
    def fake_predict(net, knob_sets, spec_):
        out = np.zeros((n_lat, spec.n_pz, spec.n_angle, spec.n_amp), dtype=np.float32)
        kb = np.clip(truth + pred_shift, 0, spec.n_amp)
        out[...] = np.where(amp_k[None, None, None, :] < kb[..., None], 1.0, 0.0)
        return surrogate.step_fit_boundary(out, 1.0)

    tracked = []

    def fake_track_ics(lat, per, n_turns, tpb=0, return_stats=False):
        off = spec.offsets()
        res = []
        for v, c in enumerate(per):
            s = np.zeros(c.shape[1], dtype=np.int32)
            for j in range(c.shape[1]):
                # recover (ip, ia, k) from the coordinates
                d = np.abs(off[..., 0] - c[0, j]) + np.abs(off[..., 1] - c[2, j]) + np.abs(off[..., 2] - c[5, j])
                ip, ia, k = np.unravel_index(np.argmin(d), d.shape)
                s[j] = surv_of(v, ip, ia, k)
            tracked.append(c.shape[1])
            res.append(s)
        return (res, dict(pet=0.0)) if return_stats else res

    monkeypatch.setattr(surrogate, "predict_boundaries", fake_predict)
    monkeypatch.setattr(surrogate, "upload_variants", lambda ctx, k, s: _FakeLat())
    from bmx import tracker
    monkeypatch.setattr(tracker, "track_ics", fake_track_ics)
    runs, area, timing, n_tracked = surrogate.surrogate_scan(None, None, np.zeros((n_lat, 32)), spec, refine=refine,
                                                            **kw)
    if kw:
        return runs, n_tracked, spec, timing
    return runs, n_tracked, spec


@pytest.mark.parametrize("refine", [0, 6])
@pytest.mark.parametrize("shift", [0, 2, -3, 15, -15])
def test_band_search_recovers_true_boundary(monkeypatch, shift, refine):
    rng = np.random.default_rng(abs(shift) + 1)
    truth = rng.integers(5, 35, size=(4, 3, 6))
    runs, n_tracked, spec = _run(monkeypatch, truth, shift, refine=refine)
    assert np.array_equal(runs, truth)
    full = truth.size * spec.n_amp
    if abs(shift) <= 2:
        assert n_tracked < 0.35 * full, (n_tracked, full)


def test_boundary_at_grid_edges(monkeypatch):
    truth = np.zeros((2, 3, 6), dtype=np.int64)
    truth[0] = 40  # survives the whole grid
    truth[1] = 0   # lost at the smallest amplitude
    runs, _, _ = _run(monkeypatch, truth, 0)
    assert np.array_equal(runs, truth)


def test_island_below_band_is_missed_by_design(monkeypatch):
    """An isolated loss far below the predicted band is not tracked (the price
    of screening); the DA-area error measured against plain tracking is what
    bounds it."""
    truth = np.full((1, 3, 6), 30)
    runs, _, _ = _run(monkeypatch, truth, 0, islands={(0, 1, 2, 5)})
    assert runs[0, 1, 2] == 30


def test_replicated_knob_lattices_equal_built_ones():
    """upload_variants' numpy replication gives, byte for byte, the element
    records knob_lattice builds for each knob set."""
    import ctypes as C
    from bmx import abi, lattice

    captured = {}

    class _Ctx:
        def lattice(self, arr, species, variant_offsets=None):
            captured["arr"] = bytes(memoryview(arr).cast("B"))
            captured["offs"] = variant_offsets
            return _FakeLat()

    spec = surrogate.ScanSpec(n_cells=12)
    rng = np.random.default_rng(5)
    knobs = rng.uniform(-0.2, 0.2, (3, surrogate.N_KNOBS))
    surrogate.upload_variants(_Ctx(), knobs, spec)
    size = C.sizeof(abi.EleRec)
    for v in range(3):
        want = bytes(memoryview(lattice.to_array(surrogate.knob_lattice(knobs[v], spec))).cast("B"))
        a, b = captured["offs"][v] * size, captured["offs"][v + 1] * size
        assert captured["arr"][a:b] == want


def test_step_fit_boundary():
    p = np.array([[1, 1, 1, 0.98, 1, 0.3, 0.2, 0.1],   # step at 5
                  [1, 1, 1, 1, 1, 1, 1, 1],            # survives the whole ray
                  [0.2, 0.1, 0.1, 0, 0, 0, 0, 0],      # lost from the start
                  [1, 1, 0.2, 1, 1, 0.1, 0.1, 0.1]])   # one mispredicted point does not move it
    assert list(surrogate.step_fit_boundary(p, 1.0)) == [5, 8, 0, 5]


@pytest.mark.parametrize("shift", [0, 1, -1, 3, -2])
def test_bracket_mode_resolves_near_predictions(monkeypatch, shift):
    """Bracket mode, window kb-1..kb, probes 4 up / 3 down, two rounds: a
    prediction within reach of the second round's probes is resolved exactly,
    with a handful of tracked ICs per ray."""
    rng = np.random.default_rng(abs(shift) + 11)
    truth = rng.integers(5, 35, size=(4, 3, 6))
    runs, n_tracked, spec, timing = _run(monkeypatch, truth, shift, window=(1, 0), probe_up=4, probe_down=3,
                                         max_rounds=2)
    assert np.array_equal(runs, truth)
    assert timing["n_open"] == 0
    assert n_tracked <= truth.size * (2 + 4)
    assert len(timing["rounds"]) == (1 if shift == 0 else 2)


def test_bracket_mode_bounds_rounds_and_estimates_open_rays(monkeypatch):
    """A prediction far off leaves rays open after max_rounds: no third round
    runs, and each open ray's boundary lies inside the bracket the tracked
    ICs prove (above the top survivor, at or below the first loss)."""
    truth = np.full((2, 3, 6), 25)
    runs, n_tracked, spec, timing = _run(monkeypatch, truth, -15, window=(1, 0), probe_up=4, probe_down=3,
                                         max_rounds=2)
    assert len(timing["rounds"]) == 2
    assert timing["n_open"] == truth.size
    # tracked: kb-1, kb (both alive) then 4 above: the top survivor is kb + 4 = 14, no loss seen
    assert np.all(runs >= 15) and np.all(runs <= spec.n_amp)


def test_neighbour_estimate():
    """An open ray takes the 1/distance-weighted mean of the resolved rays
    within 3 angles (not across the ends of the 0..pi fan), clipped into its
    bracket; resolved rays keep their boundary; an open ray with no resolved
    neighbour keeps the prediction."""
    bnd = np.array([[20, 22, 40, 24, 26, 28]])
    a_top = np.array([[19, 21, 10, 23, 25, 27]])  # ray 2 open: survivor at 10, first loss at 40
    kb = np.full_like(bnd, 5)
    est = surrogate.neighbour_estimate(bnd, a_top, kb)
    # distances from angle 2: angles 1, 3 at 1; 0, 4 at 2; 5 at 3
    num = 22 + 24 + (20 + 26) / 2 + 28 / 3
    den = 1 + 1 + 1 + 1 / 3
    assert est[0, 2] == int(np.rint(num / den))
    # the end ray 0 sees angles 1..3 only
    a_e, b_e = a_top.copy(), bnd.copy()
    a_e[0, 0], b_e[0, 0] = 5, 40
    assert surrogate.neighbour_estimate(b_e, a_e, kb)[0, 0] == int(np.rint((22 + 24 / 3) / (1 + 1 / 3)))
    # no resolved neighbour: the prediction, clipped into the bracket
    a_n = np.array([[5, 5, 5, 5, 5, 5]])
    assert np.array_equal(surrogate.neighbour_estimate(bnd, a_n, kb), np.full_like(bnd, 6))
    assert np.array_equal(np.delete(est[0], 2), np.delete(bnd[0], 2))
    # clipped into the bracket
    a_top2 = a_top.copy()
    a_top2[0, 2] = 30
    assert surrogate.neighbour_estimate(bnd, a_top2, kb)[0, 2] == 31


def test_bracket_mode_mixed_misses_across_lattices(monkeypatch):
    """Predictions off by different amounts ray by ray, some above and some
    below the truth, on several lattices: the second round mixes probes
    above, probes below and bracket fills of different lattices in one
    launch, and every IC must be tracked on its own lattice."""
    rng = np.random.default_rng(5)
    truth = rng.integers(8, 32, size=(5, 3, 6))
    shift = rng.choice([-4, -3, -2, -1, 0, 1, 2, 3], size=truth.shape)  # within reach of 4 up / 3 down
    runs, n_tracked, spec, timing = _run(monkeypatch, truth, shift, window=(1, 0), probe_up=4, probe_down=3,
                                         max_rounds=2)
    assert timing["n_open"] == 0
    assert np.array_equal(runs, truth)


@pytest.mark.parametrize("shift", [0, 1, -1, 2, 8, -8])
def test_band_mode_single_round(monkeypatch, shift):
    """Band mode (one FP64 round of the +-10 % band, max_rounds = 1): where the
    true boundary lies in the band the ray is exact; a boundary above the
    band reads as the band's top + 1, one below it as the band's bottom (the
    network stands for the untracked amplitudes), never a neighbour
    estimate."""
    rng = np.random.default_rng(abs(shift) + 21)
    truth = rng.integers(12, 35, size=(4, 3, 6))
    runs, n_tracked, spec, timing = _run(monkeypatch, truth, shift, window=None, max_rounds=1)
    assert len(timing["rounds"]) == 1
    kb = np.clip(truth + shift, 0, spec.n_amp)
    amps = spec.amp_max_x * (np.arange(spec.n_amp) + 1) / spec.n_amp
    a_b = spec.amp_max_x * np.maximum(kb, 1) / spec.n_amp
    lo = np.clip(np.minimum(np.searchsorted(amps, a_b * 0.9, side="left"), kb - 1), 0, spec.n_amp - 1)
    hi = np.clip(np.maximum(np.searchsorted(amps, a_b * 1.1, side="right") - 1, kb), 0, spec.n_amp - 1)
    want = np.where(truth > hi, np.minimum(hi + 1, spec.n_amp), np.where(truth < lo, lo, truth))
    assert np.array_equal(runs, want)
    assert n_tracked == int(np.sum(hi - lo + 1))
    if abs(shift) <= 1:
        assert np.array_equal(runs, truth)


def test_replicated_error_seeded_knob_lattices_equal_built_ones():
    """The same byte-for-byte replication on the error-seeded rf ring."""
    import ctypes as C
    from bmx import abi, lattice

    captured = {}

    class _Ctx:
        def lattice(self, arr, species, variant_offsets=None):
            captured["arr"] = bytes(memoryview(arr).cast("B"))
            captured["offs"] = variant_offsets
            return _FakeLat()

    spec = surrogate.ScanSpec(n_cells=16, n_rf=8, error_seed=77)
    knobs = np.random.default_rng(6).uniform(-0.2, 0.2, (3, surrogate.N_KNOBS))
    surrogate.upload_variants(_Ctx(), knobs, spec)
    size = C.sizeof(abi.EleRec)
    for v in range(3):
        want = bytes(memoryview(lattice.to_array(surrogate.knob_lattice(knobs[v], spec))).cast("B"))
        a, b = captured["offs"][v] * size, captured["offs"][v + 1] * size
        assert captured["arr"][a:b] == want
