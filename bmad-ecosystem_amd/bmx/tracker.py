"""Python front end of libbmx: the track_beam / track_bunch / track_many call
shape of Bmad (bmad/multiparticle/beam_mod.f90:40-170, bmad/code/track_many.f90)
over the C ABI.

Particles live in a `Bunch` of device tensors (structure of arrays, FP64 x,
px, y, py, z, pz, beta; int32 state, ix_lost, turn_lost).  PyTorch is used
only for device memory; every computation runs in libbmx's HIP kernels.
"""
import ctypes as C
import weakref

import numpy as np

from . import abi, lattice as latmod

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


class BmxError(RuntimeError):
    pass


def _check(ctx, rc, what):
    if rc != 0:
        lib = abi.load_library()
        msg = lib.bmx_last_error(ctx).decode() if ctx else ""
        raise BmxError(f"{what} failed (rc={rc}): {msg}")


class Context:
    """A libbmx context on one HIP device (bmx_ctx_create)."""

    def __init__(self, device=0):
        self.lib = abi.load_library()
        abi.check_layout(self.lib)
        h = C.c_void_p()
        rc = self.lib.bmx_ctx_create(C.byref(h), int(device), 0)
        if rc != 0:
            raise BmxError(f"bmx_ctx_create(device={device}) failed with rc={rc} (no HIP device?)")
        self.h = h
        self.device = device
        self._nets = weakref.WeakSet()  # library-held networks (surrogate.Net), closed before the context

    def close(self):
        if self.h:
            for net in list(self._nets):
                net.close()
            self.lib.bmx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def schedule(self):
        """The context's turn-block schedule (bmx_ctx_get_schedule) as a dict."""
        sc = abi.Schedule()
        self.lib.bmx_ctx_get_schedule(self.h, C.byref(sc))
        return dict(turns_per_block=sc.turns_per_block, graph_depth=sc.graph_depth, probe_pet_per_s=sc.probe_pet_per_s)

    def set_schedule(self, turns_per_block=32, graph_depth=1):
        sc = abi.Schedule(int(turns_per_block), int(graph_depth), 0.0)
        _check(self.h, self.lib.bmx_ctx_set_schedule(self.h, C.byref(sc)), "bmx_ctx_set_schedule")

    def lattice(self, eles, species, variant_offsets=None, aperture_limit_on=True, max_aperture_limit=1e3,
                tracked_species=None, absolute_time_tracking=False, rad_maps=None, radiation=None):
        return Lattice(self, eles, species, variant_offsets, aperture_limit_on, max_aperture_limit, tracked_species,
                       absolute_time_tracking, rad_maps, radiation)


def check_lattice(eles, species, aperture_limit_on=True, max_aperture_limit=1e3, tracked_species=None, direction=1,
                  absolute_time_tracking=False):
    """(ok, n_ops, reason): whether the device path accepts this element list
    for tracking in `direction`, decided on the host with no device
    (bmx_lattice_check)."""
    lib = abi.load_library()
    arr = eles if isinstance(eles, C.Array) else latmod.to_array(eles)
    ref = abi.Ref(species, tracked_species or 0)
    cfg = abi.Config(1 if aperture_limit_on else 0, int(direction), max_aperture_limit,
                     1 if absolute_time_tracking else 0, 0)
    n_ops = C.c_int64(0)
    msg = C.create_string_buffer(512)
    rc = lib.bmx_lattice_check(arr, len(arr), C.byref(ref), C.byref(cfg), C.byref(n_ops), msg, len(msg))
    return rc == 0, n_ops.value, msg.value.decode()


class Lattice:
    """A compiled, device-resident element program (bmx_lattice_upload).
    tracked_species: the bunch species when it is the antiparticle of the
    reference species (its kicks and voltages change sign)."""

    def __init__(self, ctx, eles, species, variant_offsets=None, aperture_limit_on=True, max_aperture_limit=1e3,
                 tracked_species=None, absolute_time_tracking=False, rad_maps=None, radiation=None):
        """rad_maps: None, or the radiation maps of every element as a
        float64 array [n_ele, 2, 84] (per element rm0 then rm1, each
        ref_orb[6], damp_dmat[36] row-major, xfer_damp_vec[6], stoc_mat[36]
        row-major); radiation: dict of the bmx_radiation switches
        (damping_on, fluctuations_on, zero_average, synch_rad_scale, seed)
        (bmx_lattice_upload_rad)."""
        self.ctx = ctx
        self.species = species
        self.tracked_species = tracked_species or species
        arr = eles if isinstance(eles, C.Array) else latmod.to_array(eles)
        self.n_ele_total = len(arr)
        self.p0c = arr[0].p0c
        self.e_tot = arr[0].e_tot
        ref = abi.Ref(species, tracked_species or 0)
        cfg = abi.Config(1 if aperture_limit_on else 0, 0, max_aperture_limit, 1 if absolute_time_tracking else 0, 0)
        self.absolute_time_tracking = bool(absolute_time_tracking)
        if variant_offsets is None:
            offs_p = None
            n_var = 1
            self.n_ele = [len(arr)]
        else:
            offs = (C.c_int32 * len(variant_offsets))(*variant_offsets)
            offs_p = offs
            n_var = len(variant_offsets) - 1
            self.n_ele = [variant_offsets[i + 1] - variant_offsets[i] for i in range(n_var)]
        self.n_variants = n_var
        h = C.c_void_p()
        if rad_maps is None and radiation is None:
            rc = ctx.lib.bmx_lattice_upload(ctx.h, arr, len(arr), offs_p, n_var, C.byref(ref), C.byref(cfg),
                                            C.byref(h))
            _check(ctx.h, rc, "bmx_lattice_upload")
        else:
            r = dict(damping_on=True, fluctuations_on=False, zero_average=False, synch_rad_scale=1.0, seed=0)
            r.update(radiation or {})
            rad = abi.Radiation(int(bool(r["damping_on"])), int(bool(r["fluctuations_on"])),
                                int(bool(r["zero_average"])), 0, float(r["synch_rad_scale"]), int(r["seed"]))
            maps = np.ascontiguousarray(np.asarray(rad_maps, dtype=np.float64))
            if maps.shape != (len(arr), 2, 84):
                raise ValueError(f"rad_maps must be [n_ele, 2, 84], got {maps.shape}")
            self._rad_maps = maps  # kept alive for the call
            rc = ctx.lib.bmx_lattice_upload_rad(ctx.h, arr, len(arr), offs_p, n_var,
                                                maps.ctypes.data_as(C.POINTER(abi.RadMap)), C.byref(rad),
                                                C.byref(ref), C.byref(cfg), C.byref(h))
            _check(ctx.h, rc, "bmx_lattice_upload_rad")
        self.h = h

    def set_rad_stream(self, seed, next_call=0):
        """Seed of the fluctuation stream and the number of the next call that
        draws from it (bmx_lattice_set_rad_stream)."""
        _check(self.ctx.h, self.ctx.lib.bmx_lattice_set_rad_stream(self.h, int(seed), int(next_call)),
               "bmx_lattice_set_rad_stream")

    @classmethod
    def scaled(cls, ctx, eles, species, attr, sel, factors, aperture_limit_on=True, max_aperture_limit=1e3):
        """Variants of one element list with attribute attr (abi.ATTR_K1/K2/K3)
        of elements sel scaled by factors [n_variants, len(sel)]
        (bmx_lattice_upload_scaled)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.species = self.tracked_species = species
        arr = eles if isinstance(eles, C.Array) else latmod.to_array(eles)
        f = np.ascontiguousarray(np.asarray(factors, dtype=np.float64))
        s = np.ascontiguousarray(np.asarray(sel, dtype=np.int32))
        n_var = f.shape[0]
        if f.shape[1] != len(s):
            raise ValueError("factors must be [n_variants, len(sel)]")
        self.n_ele_total = len(arr) * n_var
        self.p0c = arr[0].p0c
        self.e_tot = arr[0].e_tot
        self.n_ele = [len(arr)] * n_var
        self.n_variants = n_var
        self.absolute_time_tracking = False
        ref = abi.Ref(species, 0)
        cfg = abi.Config(1 if aperture_limit_on else 0, 0, max_aperture_limit, 0, 0)
        h = C.c_void_p()
        rc = ctx.lib.bmx_lattice_upload_scaled(ctx.h, arr, len(arr), n_var, int(attr),
                                               s.ctypes.data_as(C.POINTER(C.c_int32)), len(s),
                                               f.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ref), C.byref(cfg),
                                               C.byref(h))
        _check(ctx.h, rc, "bmx_lattice_upload_scaled")
        self.h = h
        return self

    def cells(self):
        """(ops per repeating cell, runs of it) tracked by the lattice's
        runtime-compiled cell kernels; (0, 0) when it runs on the op
        interpreter alone (bmx_lattice_cells)."""
        n, k = C.c_int32(0), C.c_int64(0)
        _check(self.ctx.h, self.ctx.lib.bmx_lattice_cells(self.h, C.byref(n), C.byref(k)), "bmx_lattice_cells")
        return n.value, k.value

    def close(self):
        if self.h:
            self.ctx.lib.bmx_lattice_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Bunch:
    """Structure-of-arrays particle store in HBM."""

    NAMES = ("x", "px", "y", "py", "z", "pz")

    def __init__(self, coords, p0c, mc2, device="cuda:0", beta=None, e_tot=None, t=None):
        """t: particle times [s] (coord_struct%t), carried and updated by
        track_bunch when given (absolute-time rf reads them)."""
        c = torch.as_tensor(np.asarray(coords, dtype=np.float64).reshape(6, -1))
        self.n = c.shape[1]
        self.device = torch.device(device)
        self.cols = [c[i].contiguous().to(self.device) for i in range(6)]
        if beta is None:
            pz = c[5]
            pc = p0c * (1.0 + pz)
            b = pc / torch.sqrt(pc * pc + mc2 * mc2)
            if e_tot is not None:  # init_coord: beta = p0c / e_tot when pz == 0 (init_coord.f90:204-208)
                b = torch.where(pz == 0, torch.full_like(b, p0c / e_tot), b)
            beta = b
        self.beta = torch.as_tensor(beta, dtype=torch.float64).contiguous().to(self.device)
        self.state = torch.ones(self.n, dtype=torch.int32, device=self.device)
        self.ix_lost = torch.full((self.n,), -1, dtype=torch.int32, device=self.device)
        self.turn_lost = torch.full((self.n,), -1, dtype=torch.int32, device=self.device)
        self.t = None if t is None else torch.as_tensor(np.asarray(t, dtype=np.float64).reshape(-1)).to(self.device)

    def particles_struct(self):
        p = abi.Particles()
        for name, t in zip(self.NAMES, self.cols):
            setattr(p, name, t.data_ptr())
        p.beta = self.beta.data_ptr()
        p.state = self.state.data_ptr()
        p.ix_lost = self.ix_lost.data_ptr()
        p.turn_lost = self.turn_lost.data_ptr()
        p.t = None if self.t is None else self.t.data_ptr()
        p.n = self.n
        p.on_device = 1
        return p

    def numpy(self):
        out = {k: t.cpu().numpy() for k, t in zip(self.NAMES, self.cols)}
        out.update(beta=self.beta.cpu().numpy(), state=self.state.cpu().numpy(), ix_lost=self.ix_lost.cpu().numpy(),
                   turn_lost=self.turn_lost.cpu().numpy())
        if self.t is not None:
            out["t"] = self.t.cpu().numpy()
        return out


def track_bunch(lat, bunch, n_turns=1, ix_start=0, ix_end=0, direction=1, sync=True):
    """n_turns passes of track_bunch(lat, bunch, ele_start = ix_start, ele_end = ix_end,
    direction).  direction = -1 tracks backward, elements ix_start down to
    ix_end + 1 (the lattice's backward program, compiled on first use).
    Returns bmx_stats as a dict."""
    if sync:
        torch.cuda.synchronize(bunch.device)
    p = bunch.particles_struct()
    st = abi.Stats()
    rc = lat.ctx.lib.bmx_track_bunch(lat.ctx.h, lat.h, int(ix_start), int(ix_end), int(n_turns), int(direction),
                                     C.byref(p), C.byref(st))
    _check(lat.ctx.h, rc, "bmx_track_bunch")
    return dict(n_lost=st.n_lost, kernel_ms=st.kernel_ms)


def bunch_moments(ctx, bunch):
    """(n_live, sums[6], second moments[6, 6]) over the live particles of a
    device Bunch (bmx_bunch_moments)."""
    torch.cuda.synchronize(bunch.device)
    p = bunch.particles_struct()
    m = (C.c_double * 28)()
    rc = ctx.lib.bmx_bunch_moments(ctx.h, C.byref(p), m)
    _check(ctx.h, rc, "bmx_bunch_moments")
    a = np.array(m[:])
    sq = np.zeros((6, 6))
    k = 7
    for r in range(6):
        for c in range(r, 6):
            sq[r, c] = sq[c, r] = a[k]
            k += 1
    return int(round(a[0])), a[1:7], sq


def track_bunch_host(lat, coords, n_turns=1, beta=None, direction=1):
    """Same as track_bunch over full turns on host numpy arrays (the library
    stages them through HBM)."""
    c = np.array(coords, dtype=np.float64, copy=True).reshape(6, -1)
    cols = [np.ascontiguousarray(c[i]) for i in range(6)]
    p, extra = _host_particles(lat, cols, beta)
    st = abi.Stats()
    rc = lat.ctx.lib.bmx_track_bunch(lat.ctx.h, lat.h, 0, 0, int(n_turns), int(direction), C.byref(p), C.byref(st))
    _check(lat.ctx.h, rc, "bmx_track_bunch")
    out = {k: v for k, v in zip(Bunch.NAMES, cols)}
    out.update(extra)
    out.update(n_lost=st.n_lost, kernel_ms=st.kernel_ms)
    return out


def _host_particles(lat, cols, beta=None):
    n = len(cols[0])
    if beta is None:
        pc = lat.p0c * (1.0 + cols[5])
        m = abi.MASS[lat.species]
        beta = np.where(cols[5] == 0, lat.p0c / lat.e_tot, pc / np.sqrt(pc * pc + m * m))
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    state = np.ones(n, dtype=np.int32)
    ix_lost = np.full(n, -1, dtype=np.int32)
    turn_lost = np.full(n, -1, dtype=np.int32)
    p = abi.Particles()
    for name, a in zip(Bunch.NAMES, cols):
        setattr(p, name, a.ctypes.data)
    p.beta = beta.ctypes.data
    p.state = state.ctypes.data
    p.ix_lost = ix_lost.ctypes.data
    p.turn_lost = turn_lost.ctypes.data
    p.n = n
    p.on_device = 0
    return p, dict(beta=beta, state=state, ix_lost=ix_lost, turn_lost=turn_lost)


def track_multi(lat, coords_per_variant, n_turns=1):
    """Track one host coordinate set [6, n_v] per variant of `lat` in one launch
    (bmx_track_multi).  Returns a list of per-variant result dicts (x..pz,
    beta, state, ix_lost, turn_lost)."""
    if len(coords_per_variant) != lat.n_variants:
        raise ValueError("one coordinate array per variant is required")
    blocks = [np.asarray(c, dtype=np.float64).reshape(6, -1) for c in coords_per_variant]
    offs = np.zeros(lat.n_variants + 1, dtype=np.int64)
    offs[1:] = np.cumsum([b.shape[1] for b in blocks])
    allc = np.concatenate(blocks, axis=1) if offs[-1] else np.zeros((6, 0))
    cols = [np.ascontiguousarray(allc[i]) for i in range(6)]
    p, extra = _host_particles(lat, cols)
    st = abi.Stats()
    rc = lat.ctx.lib.bmx_track_multi(lat.ctx.h, lat.h, offs.ctypes.data_as(C.POINTER(C.c_int64)), int(n_turns),
                                     C.byref(p), C.byref(st))
    _check(lat.ctx.h, rc, "bmx_track_multi")
    out = []
    for v in range(lat.n_variants):
        a, b = offs[v], offs[v + 1]
        d = {k: col[a:b].copy() for k, col in zip(Bunch.NAMES, cols)}
        d.update({k: val[a:b].copy() for k, val in extra.items()})
        out.append(d)
    return out


def closed_orbit(lat, n_dim=4, guess=None):
    """Closed orbit of every variant (bmx_closed_orbit).  guess: [n_var, 6]
    (pz and, for n_dim = 4, z are held).  Returns (orbit [n_var, 6], status)."""
    orb = np.zeros((lat.n_variants, 6)) if guess is None else np.array(guess, dtype=np.float64).reshape(lat.n_variants, 6)
    orb = np.ascontiguousarray(orb)
    status = np.zeros(lat.n_variants, dtype=np.int32)
    st = abi.Stats()
    rc = lat.ctx.lib.bmx_closed_orbit(lat.ctx.h, lat.h, int(n_dim), orb.ctypes.data_as(C.POINTER(C.c_double)),
                                      status.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st))
    _check(lat.ctx.h, rc, "bmx_closed_orbit")
    return orb, status


def one_turn_matrix(lat, orbit=None, d_orb=None):
    """One-turn 6x6 matrix of every variant by tracking about orbit [n_var, 6]
    (zeros when None), central differences with steps d_orb[6] (default
    1e-5 each) (bmx_one_turn_matrix).  Returns (mat6 [n_var, 6, 6], status)."""
    nv = lat.n_variants
    orb = np.ascontiguousarray(np.zeros((nv, 6)) if orbit is None else np.asarray(orbit, np.float64).reshape(nv, 6))
    d = None if d_orb is None else np.ascontiguousarray(np.asarray(d_orb, np.float64).reshape(6))
    m = np.zeros((nv, 36))
    status = np.zeros(nv, dtype=np.int32)
    st = abi.Stats()
    dp = None if d is None else d.ctypes.data_as(C.POINTER(C.c_double))
    rc = lat.ctx.lib.bmx_one_turn_matrix(lat.ctx.h, lat.h, orb.ctypes.data_as(C.POINTER(C.c_double)), dp,
                                         m.ctypes.data_as(C.POINTER(C.c_double)),
                                         status.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st))
    _check(lat.ctx.h, rc, "bmx_one_turn_matrix")
    return m.reshape(nv, 6, 6), status


def transfer_matrix(lat, orbits, ix_start=0, ix_end=0, n_turns=1):
    """Exact transfer matrices (bmx_transfer_matrix): orbits [n, 6] (or one
    orbit [6]) tracked forward through elements ix_start+1 .. ix_end (whole
    turns when equal) for n_turns passes.  Returns (out [n, 6], mat [n, 6, 6]
    with mat[i, r, q] = d out_r / d in_q, state [n])."""
    v = np.ascontiguousarray(np.asarray(orbits, dtype=np.float64).reshape(-1, 6))
    n = v.shape[0]
    out = np.zeros_like(v)
    mat = np.zeros((n, 36))
    state = np.zeros(n, dtype=np.int32)
    dp = C.POINTER(C.c_double)
    rc = lat.ctx.lib.bmx_transfer_matrix(lat.ctx.h, lat.h, int(ix_start), int(ix_end), int(n_turns), n,
                                         v.ctypes.data_as(dp), out.ctypes.data_as(dp), mat.ctypes.data_as(dp),
                                         state.ctypes.data_as(C.POINTER(C.c_int32)))
    _check(lat.ctx.h, rc, "bmx_transfer_matrix")
    return out, mat.reshape(n, 6, 6), state


def track_ics(lat, ics_per_variant, n_turns, turns_per_block=0, return_stats=False):
    """Survival turns of explicit initial conditions, one [6, n_v] host array
    per variant (bmx_track_ics).  Returns a list of int32 arrays."""
    blocks = [np.asarray(c, dtype=np.float64).reshape(6, -1) for c in ics_per_variant]
    if len(blocks) != lat.n_variants:
        raise ValueError("one initial-condition array per variant is required")
    offs = np.zeros(lat.n_variants + 1, dtype=np.int64)
    offs[1:] = np.cumsum([b.shape[1] for b in blocks])
    allc = np.ascontiguousarray(np.concatenate(blocks, axis=1))
    surv = np.zeros(int(offs[-1]), dtype=np.int32)
    st = abi.Stats()
    rc = lat.ctx.lib.bmx_track_ics(lat.ctx.h, lat.h, allc.ctypes.data_as(C.POINTER(C.c_double)),
                                   offs.ctypes.data_as(C.POINTER(C.c_int64)), int(n_turns), int(turns_per_block),
                                   surv.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st))
    _check(lat.ctx.h, rc, "bmx_track_ics")
    out = [surv[offs[v]:offs[v + 1]].copy() for v in range(lat.n_variants)]
    if return_stats:
        return out, dict(pet=st.pet, n_live=st.n_live, n_lost=st.n_lost, kernel_ms=st.kernel_ms)
    return out


def tune_grid(lat, n_turns, n_angle, n_amp, n_pz, amp_max_x, amp_ratio_y=1.0, pz_min=0.0, pz_max=0.0,
              angle_min=0.0, angle_max=np.pi, ref_orbit=None):
    """Autotune the context's turn-block schedule for this scan (bmx_tune_grid);
    returns the chosen schedule as a dict."""
    g = abi.Grid(n_angle, n_amp, n_pz, 0, amp_max_x, amp_ratio_y, pz_min, pz_max, angle_min, angle_max)
    ref_p = None
    if ref_orbit is not None:
        ref = np.ascontiguousarray(np.asarray(ref_orbit, dtype=np.float64).reshape(lat.n_variants, n_pz, 6))
        ref_p = ref.ctypes.data_as(C.POINTER(C.c_double))
    sc = abi.Schedule()
    rc = lat.ctx.lib.bmx_tune_grid(lat.ctx.h, lat.h, C.byref(g), ref_p, int(n_turns), C.byref(sc))
    _check(lat.ctx.h, rc, "bmx_tune_grid")
    return dict(turns_per_block=sc.turns_per_block, graph_depth=sc.graph_depth, probe_pet_per_s=sc.probe_pet_per_s)


def track_grid(lat, n_turns, n_angle, n_amp, n_pz, amp_max_x, amp_ratio_y=1.0, pz_min=0.0, pz_max=0.0,
               angle_min=0.0, angle_max=np.pi, turns_per_block=0, ref_orbit=None, return_stats=False):
    """Batched DA grid over every variant of `lat` (bmx_track_grid).
    ref_orbit: None or [n_var, n_pz, 6] reference orbits the offsets are added to.
    Returns (survival[n_var, n_pz, n_angle, n_amp], da_area[n_var], kernel_ms)
    (plus the stats dict when return_stats)."""
    g = abi.Grid(n_angle, n_amp, n_pz, int(turns_per_block), amp_max_x, amp_ratio_y, pz_min, pz_max, angle_min,
                 angle_max)
    n_ic = n_angle * n_amp * n_pz
    surv = np.zeros(lat.n_variants * n_ic, dtype=np.int32)
    area = np.zeros(lat.n_variants, dtype=np.float64)
    ref_p = None
    if ref_orbit is not None:
        ref = np.ascontiguousarray(np.asarray(ref_orbit, dtype=np.float64).reshape(lat.n_variants, n_pz, 6))
        ref_p = ref.ctypes.data_as(C.POINTER(C.c_double))
    st = abi.Stats()
    rc = lat.ctx.lib.bmx_track_grid(lat.ctx.h, lat.h, C.byref(g), ref_p, int(n_turns),
                                    surv.ctypes.data_as(C.POINTER(C.c_int32)),
                                    area.ctypes.data_as(C.POINTER(C.c_double)), C.byref(st))
    _check(lat.ctx.h, rc, "bmx_track_grid")
    out = (surv.reshape(lat.n_variants, n_pz, n_angle, n_amp), area, st.kernel_ms)
    if return_stats:
        return out + (dict(pet=st.pet, n_live=st.n_live, n_lost=st.n_lost, kernel_ms=st.kernel_ms),)
    return out
