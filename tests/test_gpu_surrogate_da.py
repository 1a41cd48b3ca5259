"""End-to-end surrogate DA scan on the GPU: survival data from the batched FP64
grid tracker, the MLP trained by libbmx's own forward/backward/Adam kernels,
and the screened scan (prediction, FP64 re-tracking of the boundary band,
extension fallback) on lattices the network has not seen, judged against
plain full-grid FP64 tracking of the same lattices.

What the trained network gains over the method without it: with the same
band search, an untrained network (random initial weights) needs many more
FP64-tracked initial conditions, because its predicted boundaries are far off
and the fallback has to walk to the true ones.
"""
import numpy as np
import pytest

from bmx import surrogate

pytestmark = pytest.mark.gpu


def test_surrogate_da_scan_end_to_end(ctx):
    spec = surrogate.ScanSpec(n_turns=64, n_angle=8, n_amp=24, n_pz=3, amp_max_x=0.02, amp_ratio_y=0.5,
                              pz_min=-0.002, pz_max=0.002, n_cells=30)
    rng = np.random.default_rng(2026)
    k_train = rng.uniform(-0.2, 0.2, (96, surrogate.N_KNOBS))
    k_test = rng.uniform(-0.2, 0.2, (12, surrogate.N_KNOBS))
    surv_tr, area_tr, t_tr = surrogate.plain_scan(ctx, k_train, spec)
    print(f"training scan: 96 lattices in {t_tr['total']:.2f} s; DA area {area_tr.min():.3e} .. {area_tr.max():.3e}",
          flush=True)
    assert area_tr.max() > 1.2 * area_tr.min(), "the knobs should move the DA"
    X = surrogate.features(k_train, spec)
    Y = surrogate.target(surv_tr.reshape(-1), spec.n_turns)
    net = surrogate.MLP(ctx, [35, 256, 256, 256, 256, 1], seed=3)
    losses = surrogate.train(net, X, Y, epochs=40, batch=2048, lr=2e-3)
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.3 * losses[0]

    surv_te, area_plain, t_plain = surrogate.plain_scan(ctx, k_test, spec)
    runs, area_s, t_s, n_tr = surrogate.surrogate_scan(ctx, net, k_test, spec)
    err = np.abs(area_s - area_plain) / area_plain
    full = len(k_test) * spec.n_ic
    print(f"surrogate: tracked {n_tr} of {full} ICs; DA-area error mean {err.mean():.4f} max {err.max():.4f}",
          flush=True)
    # the screened boundary equals the plain one wherever survival is monotonic along the ray
    plain_runs = np.cumprod(surv_te >= spec.n_turns, axis=-1).sum(-1)
    agree = np.mean(runs == plain_runs)
    print(f"rays with identical boundary: {agree:.3f}", flush=True)
    assert err.mean() <= 0.02
    assert n_tr < 0.6 * full

    untrained = surrogate.MLP(ctx, [35, 256, 256, 256, 256, 1], seed=4)
    _, area_u, _, n_u = surrogate.surrogate_scan(ctx, untrained, k_test, spec)
    print(f"untrained network: tracked {n_u} ICs", flush=True)
    assert n_tr < n_u

    # bracket mode on two streams: two FP64 rounds, the open rays from their neighbours.  Every
    # boundary it reports lies in the bracket its tracked ICs prove, the resolved rays agree with
    # the plain scan wherever survival is monotonic, and it tracks fewer ICs than the exact mode.
    runs_b, area_b, t_b, n_b = surrogate.surrogate_scan(ctx, net, k_test, spec, workers=2, window=(1, 0),
                                                        probe_up=4, probe_down=3, max_rounds=2)
    err_b = np.abs(area_b - area_plain) / area_plain
    print(f"bracket mode: tracked {n_b} ICs in {len(t_b['rounds'])} rounds, {t_b['n_open']} rays open; "
          f"DA-area error mean {err_b.mean():.4f} max {err_b.max():.4f}; identical {np.mean(runs_b == plain_runs):.3f}",
          flush=True)
    assert len(t_b["rounds"]) <= 2 * 2  # two rounds per stream
    assert n_b < n_tr
    assert t_b["n_open"] <= 0.25 * runs_b.size
    assert np.mean(runs_b == plain_runs) >= 0.7
    assert err_b.mean() <= 0.02


def test_surrogate_guided_optimisation_end_to_end(ctx):
    """The surrogate-guided sextupole search at reduced scale (3 generations
    of 256 candidates on a 16-cell error-seeded rf ring, 64 turns, top 5 %
    tracked in FP64, retraining after generation 2), then the plain search
    with its wall time: counts, a best area that never drops, and a network
    whose predicted areas of the tracked candidates are finite."""
    from bmx import optimize
    spec = surrogate.ScanSpec(n_turns=64, n_angle=16, n_amp=24, n_pz=3, amp_max_x=0.012, amp_ratio_y=1.0,
                              pz_min=-0.01, pz_max=0.01, n_cells=16, n_rf=8, error_seed=5)
    r = optimize.compare_search(ctx, spec, n_gen=3, n_cand=256, top_frac=0.05, retrain_every=2, n_init=32,
                                init_ics=256, init_epochs=2, retrain_epochs=1, batch=2048)
    s, p = r["surrogate"], r["plain"]
    print("optimisation:", {k: s[k] for k in ("best_area_mm2_per_gen", "final_network_area_err_mean", "n_fp64")},
          {k: p[k] for k in ("best_area_mm2_per_gen", "n_fp64")}, flush=True)
    assert s["n_tracked_per_gen"] == 13 and s["n_fp64"] == 3 * 13
    b = s["best_area_mm2_per_gen"]
    assert all(y >= x for x, y in zip(b, b[1:])) and b[-1] > 0
    assert np.isfinite(s["final_network_area_err_mean"])
    assert p["n_fp64"] == 13 * p["n_gen"] and p["n_gen"] >= 1
