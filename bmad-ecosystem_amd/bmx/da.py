"""Dynamic-aperture drivers over the batched tracker.

`reference_orbits` forms the per-(variant, pz) reference orbit the way
dynamic_aperture_scan does (bmad/modules/dynamic_aperture_mod.f90:130-140):
4-D closed orbit at each starting pz without rf, or the 6-D closed orbit plus
pz times the dispersion vector with rf.  `dynamic_aperture_scan` is the
per-ray bisection of dynamic_aperture_point (:284-398) run for every ray of
every variant and pz at once: each bisection round is one bmx_track_ics
launch.  `grid_da` is the polar-grid scan of bmx_track_grid with the same
reference orbits.
"""
import math

import numpy as np

from . import tracker


def reference_orbits(lat, pz_list, rf_on, lat_no_rf=None, eta_delta=1e-6):
    """[n_var, n_pz, 6] reference orbits.

    rf_on: 6-D closed orbit of `lat` plus pz * (eta_x, eta_px, eta_y, eta_py, 0, 1),
    with the dispersion taken from 4-D closed orbits at pz = +-eta_delta of
    `lat_no_rf` (the same variants with the cavities off).
    rf off: 4-D closed orbit at each pz.  Returns (orbits, status[n_var, n_pz])."""
    nv = lat.n_variants
    pz_list = list(pz_list)
    out = np.zeros((nv, len(pz_list), 6))
    status = np.zeros((nv, len(pz_list)), dtype=np.int32)
    if rf_on:
        co6, st6 = tracker.closed_orbit(lat, 6)
        lr = lat_no_rf if lat_no_rf is not None else lat
        g = np.zeros((nv, 6))
        g[:, 5] = eta_delta
        cop, stp = tracker.closed_orbit(lr, 4, g)
        g[:, 5] = -eta_delta
        com, stm = tracker.closed_orbit(lr, 4, g)
        eta = np.zeros((nv, 6))
        eta[:, :4] = (cop[:, :4] - com[:, :4]) / (2 * eta_delta)
        eta[:, 5] = 1.0
        for j, pz in enumerate(pz_list):
            out[:, j] = co6 + pz * eta
            status[:, j] = np.maximum(st6, np.maximum(stp, stm))
    else:
        for j, pz in enumerate(pz_list):
            g = np.zeros((nv, 6))
            g[:, 5] = pz
            co, st = tracker.closed_orbit(lat, 4, g)
            out[:, j] = co
            status[:, j] = st
    return out, status


def angle_list(n_angle, min_angle=0.0, max_angle=math.pi):
    """dynamic_aperture_mod.f90:88-95."""
    if n_angle == 1:
        return np.array([(min_angle + max_angle) / 2])
    d = (max_angle - min_angle) / (n_angle - 1)
    return np.array([i * d + min_angle for i in range(n_angle)])


class _Rays:
    """Bisection state of dynamic_aperture_point for many rays."""

    def __init__(self, theta, x_init, y_init):
        n = len(theta)
        self.x0 = np.zeros(n)
        self.y0 = np.zeros(n)
        self.x1 = np.cos(theta) * x_init
        self.y1 = np.sin(theta) * y_init
        self.x2 = np.zeros(n)
        self.y2 = np.zeros(n)
        self.bracketed = np.zeros(n, dtype=bool)
        self.turn_lost = np.zeros(n, dtype=np.int64)
        self.done = np.zeros(n, dtype=bool)
        self.gave_up = np.zeros(n, dtype=bool)


def _bisect(track, ref, rays, n_turn, rel_accuracy, abs_accuracy, max_rounds=200):
    """Runs the bisection (dynamic_aperture_mod.f90:325-387) for every ray.
    track(ics[6, n_active], ray_index[n_active]) -> survival turns."""
    for _ in range(max_rounds):
        act = np.flatnonzero(~rays.done)
        if len(act) == 0:
            break
        ics = ref[act].T.copy()
        ics[0] += rays.x1[act]
        ics[2] += rays.y1[act]
        surv = track(ics, act)
        lost = surv < n_turn
        # change search interval end
        rays.x2[act[lost]] = rays.x1[act[lost]]
        rays.y2[act[lost]] = rays.y1[act[lost]]
        rays.turn_lost[act[lost]] = surv[lost] + 1
        rays.bracketed[act[lost]] = True
        ok = act[~lost]
        rays.x0[ok] = rays.x1[ok]
        rays.y0[ok] = rays.y1[ok]
        # new starting point
        b = act[rays.bracketed[act]]
        u = act[~rays.bracketed[act]]
        rays.x1[b] = (rays.x0[b] + rays.x2[b]) / 2
        rays.y1[b] = (rays.y0[b] + rays.y2[b]) / 2
        rays.x1[u] = 2 * rays.x1[u]
        rays.y1[u] = 2 * rays.y1[u]
        far = act[(rays.x1[act] > 1.0) | (rays.y1[act] > 1.0)]
        rays.gave_up[far] = True
        rays.done[far] = True
        # accuracy
        r = 0.5 * np.sqrt((rays.x0[b] + rays.x2[b]) ** 2 + (rays.y0[b] + rays.y2[b]) ** 2)
        acc = np.sqrt((rays.x2[b] - rays.x0[b]) ** 2 + (rays.y2[b] - rays.y0[b]) ** 2) <= \
            2 * (abs_accuracy + r * rel_accuracy)
        rays.done[b[acc]] = True
    return rays


def dynamic_aperture_scan(lat, ref_orbits, n_turn=100, n_angle=9, min_angle=0.0, max_angle=math.pi, x_init=1e-3,
                          y_init=1e-3, rel_accuracy=1e-2, abs_accuracy=1e-5, turns_per_block=0, track=None):
    """dynamic_aperture_scan (dynamic_aperture_mod.f90:27-207) for every variant
    of `lat` and every pz slice of ref_orbits [n_var, n_pz, 6] at once.

    As the reference: the theta = 0 and pi/2 rays run first from x_init /
    y_init; their results become the scale of every other ray.  Returns a dict
    of [n_var, n_pz, n_angle] arrays x, y (aperture point offsets from the
    reference orbit), i_turn (turn of loss of the bracketing point) and
    `gave_up` (no aperture below 1 m), plus the angle list."""
    ref_orbits = np.asarray(ref_orbits, dtype=np.float64)
    nv, npz = ref_orbits.shape[:2]
    if track is None:
        def track(ics, idx_var):
            per = [ics[:, idx_var == v] for v in range(nv)]
            s = tracker.track_ics(lat, per, n_turn, turns_per_block)
            out = np.zeros(ics.shape[1], dtype=np.int64)
            for v in range(nv):
                out[idx_var == v] = s[v]
            return out

    angles = angle_list(n_angle, min_angle, max_angle)
    n_grp = nv * npz
    ref_g = ref_orbits.reshape(n_grp, 6)
    var_g = np.repeat(np.arange(nv), npz)

    def run(theta_per_ray, grp_per_ray, xi, yi):
        rays = _Rays(theta_per_ray, xi, yi)
        ref = ref_g[grp_per_ray]
        _bisect(lambda ics, act: track(ics, var_g[grp_per_ray[act]]), ref, rays, n_turn, rel_accuracy,
                abs_accuracy)
        return rays

    # scale factors: theta = 0 and pi/2 first (dynamic_aperture_mod.f90:163-176)
    grp = np.arange(n_grp)
    xr = run(np.zeros(n_grp), grp, np.full(n_grp, x_init), np.full(n_grp, y_init))
    yr = run(np.full(n_grp, math.pi / 2), grp, np.full(n_grp, x_init), np.full(n_grp, y_init))
    xs, ys = xr.x1.copy(), yr.y1.copy()
    # every other angle with the per-(variant, pz) scales
    res = {k: np.zeros((n_grp, n_angle)) for k in ("x", "y")}
    res["i_turn"] = np.zeros((n_grp, n_angle), dtype=np.int64)
    res["gave_up"] = np.zeros((n_grp, n_angle), dtype=bool)
    other = [i for i, a in enumerate(angles) if abs(a) >= 1e-6 and abs(a - math.pi / 2) >= 1e-6]
    if other:
        th = np.tile(angles[other], n_grp)
        g_of = np.repeat(grp, len(other))
        rr = run(th, g_of, xs[g_of], ys[g_of])
        for k, ia in enumerate(other):
            sel = slice(k, None, len(other))
            res["x"][:, ia] = rr.x1[sel]
            res["y"][:, ia] = rr.y1[sel]
            res["i_turn"][:, ia] = rr.turn_lost[sel]
            res["gave_up"][:, ia] = rr.gave_up[sel]
    for ia, a in enumerate(angles):
        src = xr if abs(a) < 1e-6 else (yr if abs(a - math.pi / 2) < 1e-6 else None)
        if src is not None:
            res["x"][:, ia] = src.x1
            res["y"][:, ia] = src.y1
            res["i_turn"][:, ia] = src.turn_lost
            res["gave_up"][:, ia] = src.gave_up
    out = {k: v.reshape(nv, npz, n_angle) for k, v in res.items()}
    out["angles"] = angles
    return out


def polygon_area(x, y):
    """Area of the polygon through the origin and the points (x_i, y_i) in
    order: 1/2 sum (x_i y_{i+1} - x_{i+1} y_i)."""
    x = np.asarray(x)
    y = np.asarray(y)
    return 0.5 * float(np.sum(x[:-1] * y[1:] - x[1:] * y[:-1]))


def grid_boundary(survival, n_turns, amp_max_x, amp_ratio_y, angles):
    """Per-angle boundary of a polar grid slice survival[n_angle, n_amp]: the
    offset of the last IC of the contiguous surviving run from the smallest
    amplitude (the definition bmx_track_grid's DA area uses)."""
    n_angle, n_amp = survival.shape
    alive = survival >= n_turns
    run = np.cumprod(alive, axis=1).sum(axis=1)  # length of the surviving prefix
    amp = amp_max_x * run / n_amp
    return amp * np.cos(angles), amp * amp_ratio_y * np.sin(angles), run
