"""Surrogate-guided sextupole optimisation: a search over the 32 sextupole-
family knobs for the largest on-momentum dynamic aperture, in which the DA
surrogate screens every candidate of a generation and only the most
promising few are tracked in FP64 -- and the same search with FP64 tracking
alone, for comparison at equal wall time.

No counterpart in the reference (its DA is evaluated point by point,
bmad/modules/dynamic_aperture_mod.f90:27-207).  The loop:

* generation 0 draws n_cand knob vectors from the uniform prior U(+-20 %);
  later generations mutate the best FP64-evaluated settings found so far
  (Gaussian steps, clipped to the prior's range) and draw a share afresh;
* the network predicts every grid ray of every candidate on the on-momentum
  slice (surrogate.predict_boundaries) and ranks the candidates by the area
  of the predicted boundary;
* the top fraction is tracked in FP64 on the full grid (every pz slice) for
  the spec's turns; those survival grids join the training set;
* every retrain_every generations the network is trained further on all
  tracking data gathered so far.

The plain search runs the same mutation scheme with FP64 tracking of every
candidate it draws (pop per generation) until it has used the surrogate
loop's wall time.
"""
import sys
import time

import numpy as np

from . import surrogate


def _say(msg):
    print(f"optimisation: {msg}", file=sys.stderr, flush=True)


def _mid_spec(spec):
    """The on-momentum slice of spec's grid, with the same network inputs
    (geometry() normalises pz by max|pz| of the full grid; pz = 0 stays 0)."""
    s = surrogate.ScanSpec(n_turns=spec.n_turns, n_angle=spec.n_angle, n_amp=spec.n_amp, n_pz=1,
                           amp_max_x=spec.amp_max_x, amp_ratio_y=spec.amp_ratio_y, pz_min=0.0, pz_max=0.0,
                           n_cells=spec.n_cells, n_rf=spec.n_rf, error_seed=spec.error_seed)
    return s


def predicted_areas(net, knob_sets, spec):
    """On-momentum DA area predicted by the network for every knob set [m^2]."""
    mid = _mid_spec(spec)
    kb = surrogate.predict_boundaries(net, knob_sets, mid)  # [n, 1, n_angle]
    return np.array([surrogate.area_from_runs(kb[v, 0], spec) for v in range(len(knob_sets))])


def fp64_evaluate(ctx, knob_sets, spec):
    """Full-grid FP64 tracking of the knob sets: (on-momentum areas [m^2],
    survival [n, n_pz, n_angle, n_amp], seconds)."""
    t0 = time.perf_counter()
    surv, area, _ = surrogate.plain_scan(ctx, knob_sets, spec)
    return area, surv, time.perf_counter() - t0


def _mutate(rng, parents, n, sigma, fresh):
    """n children: a share `fresh` from the uniform prior, the rest Gaussian
    steps of width sigma from parents drawn uniformly, clipped to +-20 %."""
    R = surrogate.KNOB_RANGE
    n_fresh = int(round(fresh * n))
    out = np.empty((n, surrogate.N_KNOBS))
    out[:n_fresh] = rng.uniform(-R, R, (n_fresh, surrogate.N_KNOBS))
    pick = parents[rng.integers(0, len(parents), n - n_fresh)]
    out[n_fresh:] = np.clip(pick + sigma * rng.standard_normal(pick.shape), -R, R)
    return out


def _grid_samples(rng, knob_sets, surv, spec, n_ic):
    """Training rows from full FP64 survival grids surv [n, n_pz, n_angle,
    n_amp]: n_ic random grid ICs of each lattice (features as
    surrogate.features, targets as surrogate.target)."""
    n = len(knob_sets)
    geo = surrogate.geometry(spec)
    X = np.empty((n, n_ic, surrogate.N_KNOBS + 3), dtype=np.float32)
    Y = np.empty((n, n_ic), dtype=np.float32)
    flat = surv.reshape(n, -1)
    for v in range(n):
        idx = rng.choice(spec.n_ic, size=min(n_ic, spec.n_ic), replace=False)
        X[v, :, :surrogate.N_KNOBS] = knob_sets[v] / surrogate.KNOB_RANGE
        X[v, :, surrogate.N_KNOBS:] = geo[idx]
        Y[v] = surrogate.target(flat[v, idx], spec.n_turns)
    return X.reshape(-1, surrogate.N_KNOBS + 3), Y.reshape(-1)


class _Archive:
    """Every FP64-evaluated knob set with its area, the best first on demand."""

    def __init__(self):
        self.k, self.a = [], []

    def add(self, k, a):
        self.k.append(np.asarray(k))
        self.a.append(np.asarray(a))

    def best(self, m):
        K, A = np.concatenate(self.k), np.concatenate(self.a)
        order = np.argsort(-A)[:m]
        return K[order], A[order]

    @property
    def n(self):
        return sum(len(a) for a in self.a)


def surrogate_search(ctx, spec, n_gen=20, n_cand=4096, top_frac=0.05, retrain_every=5, n_init=512, init_ics=8192,
                     init_epochs=6, retrain_epochs=2, n_parents=32, sigma=0.03, fresh=0.25, seed=20261206,
                     batch=16384):
    """The surrogate-guided search (module docstring).  Returns a dict: the
    best FP64 area per generation, the FP64-evaluated and screened counts,
    timings (one-time: initial training data and training; loop: screening,
    FP64 tracking, retraining), lattices/hour and the network's DA-area
    error on each generation's re-tracked candidates."""
    rng = np.random.default_rng(seed)
    R, NK = surrogate.KNOB_RANGE, surrogate.N_KNOBS
    n_top = max(1, int(round(top_frac * n_cand)))
    t_start = time.perf_counter()
    # one-time: the initial network from n_init random settings x init_ics random grid ICs
    _say(f"initial training data: {n_init} lattices x {init_ics} ICs, {spec.n_turns} turns")
    k0 = rng.uniform(-R, R, (n_init, NK))
    X0, Y0, t_gen = surrogate.training_data(ctx, k0, spec, init_ics, seed=seed + 1, chunk=256)
    net = surrogate.Net(ctx, [NK + 3, 256, 256, 256, 256, 1], seed=seed % 1000)
    t0 = time.perf_counter()
    surrogate.train(net, X0, Y0, epochs=init_epochs, batch=batch)
    t_init_train = time.perf_counter() - t0
    t_setup = time.perf_counter() - t_start
    _say(f"initial network in {t_setup:.1f} s")
    data_X, data_Y = [X0], [Y0]
    arch = _Archive()
    best_per_gen, err_per_gen = [], []
    t_screen = t_fp64 = t_retrain = 0.0
    n_screened = 0
    t_loop0 = time.perf_counter()
    for g in range(n_gen):
        if g == 0:
            cand = rng.uniform(-R, R, (n_cand, NK))
        else:
            parents, _ = arch.best(n_parents)
            cand = _mutate(rng, parents, n_cand, sigma, fresh)
        t0 = time.perf_counter()
        pred = predicted_areas(net, cand, spec)
        t_screen += time.perf_counter() - t0
        n_screened += n_cand
        top = np.argsort(-pred)[:n_top]
        area, surv, dt = fp64_evaluate(ctx, cand[top], spec)
        t_fp64 += dt
        arch.add(cand[top], area)
        nz = area > 0
        err = np.abs(pred[top][nz] - area[nz]) / area[nz] if nz.any() else np.zeros(1)
        err_per_gen.append((float(err.mean()), float(err.max()), float(np.mean(err <= 0.02))))
        # the re-tracked grids join the training set, init_ics random grid ICs per lattice
        Xg, Yg = _grid_samples(rng, cand[top], surv, spec, init_ics)
        data_X.append(Xg)
        data_Y.append(Yg)
        _, best = arch.best(1)
        best_per_gen.append(float(best[0]))
        _say(f"generation {g}: best FP64 area {best[0] * 1e6:.3f} mm^2, screened {n_cand}, tracked {n_top}, "
             f"network area error mean {err.mean():.3f}")
        if (g + 1) % retrain_every == 0 and g + 1 < n_gen:
            t0 = time.perf_counter()
            surrogate.train(net, np.concatenate(data_X), np.concatenate(data_Y), epochs=retrain_epochs, batch=batch,
                            lr=1e-3, lr_final=1e-4, seed=g)
            t_retrain += time.perf_counter() - t0
    t_loop = time.perf_counter() - t_loop0
    kbest, abest = arch.best(1)
    return {
        "n_gen": n_gen, "n_cand": n_cand, "n_tracked_per_gen": n_top, "retrain_every": retrain_every,
        "n_init": n_init, "init_ics": init_ics, "turns": spec.n_turns,
        "best_area_mm2_per_gen": [a * 1e6 for a in best_per_gen], "best_area_mm2": float(abest[0] * 1e6),
        "best_knobs": kbest[0].tolist(),
        "network_area_err_per_gen": err_per_gen,
        "final_network_area_err_mean": err_per_gen[-1][0], "final_network_area_err_max": err_per_gen[-1][1],
        "final_network_frac_within_2pct": err_per_gen[-1][2],
        "setup_s": t_setup, "init_data_s": t_gen["total"], "init_train_s": t_init_train,
        "loop_s": t_loop, "screen_s": t_screen, "fp64_s": t_fp64, "retrain_s": t_retrain,
        "lattices_screened_per_hour": n_screened / t_loop * 3600,
        "lattices_fp64_per_hour": arch.n / t_fp64 * 3600 if t_fp64 > 0 else None,
        "n_fp64": arch.n,
    }


def plain_search(ctx, spec, wall_s, pop=205, n_parents=32, sigma=0.03, fresh=0.25, seed=20261206):
    """The same search with FP64 tracking of every candidate: generations of
    pop settings (generation 0 from the uniform prior, then mutations of the
    best found) until wall_s seconds of wall time are used.  Returns the
    best area per generation, the lattices evaluated and lattices/hour."""
    rng = np.random.default_rng(seed)
    R, NK = surrogate.KNOB_RANGE, surrogate.N_KNOBS
    arch = _Archive()
    best_per_gen = []
    t0 = time.perf_counter()
    g = 0
    while True:
        cand = rng.uniform(-R, R, (pop, NK)) if g == 0 else _mutate(rng, arch.best(n_parents)[0], pop, sigma, fresh)
        area, _, _ = fp64_evaluate(ctx, cand, spec)
        arch.add(cand, area)
        best_per_gen.append(float(arch.best(1)[1][0]))
        g += 1
        el = time.perf_counter() - t0
        _say(f"plain generation {g - 1}: best FP64 area {best_per_gen[-1] * 1e6:.3f} mm^2 ({el:.0f} / {wall_s:.0f} s)")
        if el >= wall_s:
            break
    el = time.perf_counter() - t0
    return {"n_gen": g, "pop": pop, "best_area_mm2_per_gen": [a * 1e6 for a in best_per_gen],
            "best_area_mm2": best_per_gen[-1] * 1e6, "n_fp64": arch.n, "wall_s": el,
            "lattices_per_hour": arch.n / el * 3600}


def compare_search(ctx, spec, n_gen=20, **kw):
    """Surrogate-guided search, then the plain search given the surrogate
    loop's wall time (its one-time initial training included).  Returns
    both records and the comparison: screened lattices/hour against plain
    FP64 lattices/hour, and the best FP64 areas each found."""
    s = surrogate_search(ctx, spec, n_gen=n_gen, **kw)
    budget = s["setup_s"] + s["loop_s"]
    p = plain_search(ctx, spec, budget, pop=s["n_tracked_per_gen"])
    return {"surrogate": s, "plain": p,
            "lattices_per_hour_gain": s["lattices_screened_per_hour"] / p["lattices_per_hour"],
            "best_area_ratio": s["best_area_mm2"] / p["best_area_mm2"] if p["best_area_mm2"] > 0 else None}
