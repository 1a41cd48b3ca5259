"""Host logic of the surrogate-guided sextupole optimisation
(bmx.optimize), with the network and the FP64 tracker replaced by CPU
stand-ins: a synthetic DA-area landscape with a known optimum.  The loop
must screen n_cand settings per generation, track the top fraction, retrain
on schedule, improve on its start and report consistent counts; the plain
search must stop at its wall-time budget.  (The GPU run is in
test_gpu_surrogate_da.py.)"""
import numpy as np

from bmx import optimize, surrogate

OPT = np.linspace(-0.1, 0.1, surrogate.N_KNOBS)


def _true_area(k):
    return np.exp(-np.sum((np.asarray(k) - OPT) ** 2, axis=-1) / 0.05) * 1e-5


def _stubs(monkeypatch, noise=0.05):
    spec = surrogate.ScanSpec(n_turns=64, n_angle=4, n_amp=8, n_pz=3)
    calls = dict(train=0, fp64=0, screened=0)
    rng = np.random.default_rng(3)

    def fake_training_data(ctx, k, spec_, n_ic, seed=0, chunk=256, gamma=None):
        n = len(k) * n_ic
        return np.zeros((n, surrogate.N_KNOBS + 3), np.float32), np.zeros(n, np.float32), dict(total=0.0, track=0.0)

    class FakeNet:
        def __init__(self, *a, **kw):
            pass

    def fake_train(net, X, Y, **kw):
        calls["train"] += 1
        return [0.0]

    def fake_predicted(net, k, spec_):
        calls["screened"] += len(k)
        a = _true_area(k)
        return a * (1 + noise * rng.standard_normal(len(a)))

    def fake_eval(ctx, k, spec_):
        calls["fp64"] += len(k)
        surv = np.full((len(k), spec.n_pz, spec.n_angle, spec.n_amp), spec.n_turns, dtype=np.int32)
        return _true_area(k), surv, 0.001

    monkeypatch.setattr(surrogate, "training_data", fake_training_data)
    monkeypatch.setattr(surrogate, "Net", FakeNet)
    monkeypatch.setattr(surrogate, "train", fake_train)
    monkeypatch.setattr(optimize, "predicted_areas", fake_predicted)
    monkeypatch.setattr(optimize, "fp64_evaluate", fake_eval)
    return spec, calls


def test_surrogate_search_counts_and_improves(monkeypatch):
    spec, calls = _stubs(monkeypatch)
    r = optimize.surrogate_search(None, spec, n_gen=7, n_cand=512, top_frac=0.05, retrain_every=3, n_init=8,
                                  init_ics=16)
    n_top = int(round(0.05 * 512))
    assert r["n_tracked_per_gen"] == n_top and calls["fp64"] == 7 * n_top and calls["screened"] == 7 * 512
    assert calls["train"] == 1 + 2  # the initial training, then after generations 3 and 6
    best = r["best_area_mm2_per_gen"]
    assert all(b2 >= b1 for b1, b2 in zip(best, best[1:]))  # the archive's best never drops
    assert best[-1] > best[0]
    assert r["n_fp64"] == 7 * n_top


def test_plain_search_stops_at_its_budget(monkeypatch):
    spec, calls = _stubs(monkeypatch)
    import time
    t0 = time.perf_counter()
    p = optimize.plain_search(None, spec, wall_s=0.05, pop=20)
    assert time.perf_counter() - t0 < 5.0
    assert p["n_fp64"] == 20 * p["n_gen"] and calls["fp64"] == p["n_fp64"]
    assert p["wall_s"] >= 0.05
