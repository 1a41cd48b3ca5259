"""ctypes front end of the CPU oracle (oracle/bmx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (libbmx and the bmx
package) never imports this module.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "bmad-ecosystem_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from bmx import abi  # noqa: E402  (struct layout of include/bmx.h)

LIB = os.path.join(HERE, "liboracle.so")
LIB_COUNT = os.path.join(HERE, "liboracle_count.so")

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def _setup(lib):
    lib.oracle_track1.restype = C.c_int
    lib.oracle_track1.argtypes = [C.POINTER(abi.EleRec), C.c_int, C.c_int, C.c_int, _DP, _DP, _DP, _DP, _IP, C.c_double]
    lib.oracle_set_com.restype = None
    lib.oracle_set_com.argtypes = [C.c_int, C.c_int, C.c_double]
    lib.oracle_set_tracked_species.restype = None
    lib.oracle_set_tracked_species.argtypes = [C.c_int]
    lib.oracle_set_direction.restype = None
    lib.oracle_set_direction.argtypes = [C.c_int]
    lib.oracle_set_times.restype = None
    lib.oracle_set_times.argtypes = [_DP]
    lib.oracle_hwang_edge.restype = None
    lib.oracle_hwang_edge.argtypes = [C.POINTER(abi.EleRec), C.c_int, C.c_int, _DP, C.c_int, C.c_int]
    lib.oracle_check_aperture.restype = C.c_int
    lib.oracle_check_aperture.argtypes = [C.POINTER(abi.EleRec), _DP, C.c_int, _DP]
    lib.oracle_multipole_ele_to_ab.restype = C.c_int
    lib.oracle_multipole_ele_to_ab.argtypes = [C.POINTER(abi.EleRec), C.c_int, C.c_int, C.c_int, _DP, _DP, _DP]
    for fn in (lib.oracle_track_bunch, lib.oracle_track_particle_outer):
        fn.restype = C.c_int
        fn.argtypes = [C.POINTER(abi.EleRec), C.c_int, C.c_int, C.c_int, C.c_int64] + [_DP] * 7 + [_IP] * 3 + [C.c_int]
    return lib


LIB_FMA = os.path.join(HERE, "liboracle_fma.so")
_lib = None
_lib_count = None
_lib_fma = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _setup(C.CDLL(LIB))
    return _lib


def lib_fma():
    global _lib_fma
    if _lib_fma is None:
        if not os.path.exists(LIB_FMA):
            build()
        _lib_fma = _setup(C.CDLL(LIB_FMA))
    return _lib_fma


def lib_count():
    global _lib_count
    if _lib_count is None:
        if not os.path.exists(LIB_COUNT):
            build()
        _lib_count = _setup(C.CDLL(LIB_COUNT))
        _lib_count.oracle_flops_read_reset.restype = C.c_longlong
    return _lib_count


def set_com(aperture_limit_on=1, absolute_time_tracking=0, max_aperture_limit=1e3, counting=False):
    (lib_count() if counting else lib()).oracle_set_com(aperture_limit_on, absolute_time_tracking, max_aperture_limit)


def track1(ele, ref_species, species, direction, vec):
    """One particle through one element; returns (vec6, (t0, t1), (beta0, beta1), state)."""
    vin = (C.c_double * 6)(*vec)
    vout = (C.c_double * 6)()
    t = (C.c_double * 2)()
    b = (C.c_double * 2)()
    st = C.c_int()
    rc = lib().oracle_track1(C.byref(ele), ref_species, species, direction, vin, vout, t, b, C.byref(st), 0.0)
    if rc != 0:
        raise RuntimeError(f"oracle_track1 returned {rc}")
    return list(vout), (t[0], t[1]), (b[0], b[1]), st.value


def hwang_edge(ele, ref_species, first_edge, vec, species, direction=1):
    v = (C.c_double * 6)(*vec)
    lib().oracle_hwang_edge(C.byref(ele), ref_species, 1 if first_edge else 2, v, species, direction)
    return list(v)


def check_aperture(ele, vec, species):
    v = (C.c_double * 6)(*vec)
    uf = C.c_double()
    st = lib().oracle_check_aperture(C.byref(ele), v, species, C.byref(uf))
    return st, uf.value


def beta_of(pz, p0c, mc2):
    pc = p0c * (1.0 + pz)
    return pc / np.sqrt(pc * pc + mc2 * mc2)


def track_bunch(eles, species, n_turns, coords, beta=None, n_threads=1, particle_outer=False, counting=False,
                fma=False, tracked_species=0, direction=1, t=None, absolute_time=False):
    """Track SoA coords (dict or (6, n) array) through the element array for n_turns.
    Returns dict with x..pz, beta, state, ix_lost, turn_lost (copies).
    fma=True uses the FMA-contracted build (a second FP evaluation order).
    tracked_species: the particles' species when not the reference species.
    direction = -1: every turn runs the elements last to first, backward.
    t: particle times carried through (returned as out["t"]); absolute_time:
    rf phases from them (bmad_com%absolute_time_tracking)."""
    L = lib_count() if counting else (lib_fma() if fma else lib())
    arr = eles if isinstance(eles, C.Array) else abi_array(eles)
    c = np.array(coords, dtype=np.float64, copy=True).reshape(6, -1)
    n = c.shape[1]
    cols = [np.ascontiguousarray(c[i]) for i in range(6)]
    if beta is None:
        beta = beta_of(cols[5], arr[0].p0c_start, abi.MASS[species])
        beta = np.where(cols[5] == 0, arr[0].p0c_start / arr[0].e_tot_start, beta)
    beta = np.ascontiguousarray(np.array(beta, dtype=np.float64))
    state = np.ones(n, dtype=np.int32)
    ix_lost = np.full(n, -1, dtype=np.int32)
    turn_lost = np.full(n, -1, dtype=np.int32)
    fn = L.oracle_track_particle_outer if particle_outer else L.oracle_track_bunch
    ptrs = [x.ctypes.data_as(_DP) for x in cols + [beta]]
    if direction != 1 and particle_outer:
        raise ValueError("the particle-outer loop tracks forward only")
    if (t is not None or absolute_time) and particle_outer:
        raise ValueError("the particle-outer loop carries no particle time")
    tt = None
    if t is not None:
        tt = np.ascontiguousarray(np.array(t, dtype=np.float64, copy=True).reshape(n))
    elif absolute_time:
        raise ValueError("absolute-time tracking needs the particle times t")
    L.oracle_set_tracked_species(tracked_species or 0)
    L.oracle_set_direction(direction)
    L.oracle_set_times(tt.ctypes.data_as(_DP) if tt is not None else None)
    if absolute_time:
        L.oracle_set_com(1, 1, 1e3)
    try:
        rc = fn(arr, len(arr), species, n_turns, n, *ptrs, state.ctypes.data_as(_IP), ix_lost.ctypes.data_as(_IP),
                turn_lost.ctypes.data_as(_IP), n_threads)
    finally:
        L.oracle_set_tracked_species(0)
        L.oracle_set_direction(1)
        L.oracle_set_times(None)
        if absolute_time:
            L.oracle_set_com(1, 0, 1e3)
    if rc != 0:
        raise RuntimeError(f"oracle tracking returned {rc}")
    names = ["x", "px", "y", "py", "z", "pz"]
    out = {k: v for k, v in zip(names, cols)}
    out.update(beta=beta, state=state, ix_lost=ix_lost, turn_lost=turn_lost)
    if tt is not None:
        out["t"] = tt
    return out


def abi_array(eles):
    arr = (abi.EleRec * len(eles))()
    for i, e in enumerate(eles):
        arr[i] = e
    return arr


def flops_read_reset():
    return lib_count().oracle_flops_read_reset()


LIB_DUAL = os.path.join(HERE, "liboracle_dual.so")
_lib_dual = None


def jacobian(eles, species, n_turns, vec):
    """Exact (dual-number) Jacobian of the oracle's n_turns map at vec.
    Returns (out6, J 6x6); raises if the particle is lost."""
    L = _dual()
    arr = eles if isinstance(eles, C.Array) else abi_array(eles)
    v = (C.c_double * 6)(*vec)
    out = (C.c_double * 6)()
    J = (C.c_double * 36)()
    beta0 = arr[0].p0c_start / arr[0].e_tot_start
    rc = L.oracle_jacobian(arr, len(arr), species, n_turns, v, beta0, out, J)
    if rc != 0:
        raise RuntimeError(f"particle lost (state {rc}) while forming the Jacobian")
    return np.array(out[:]), np.array(J[:]).reshape(6, 6)


def symp_error(M):
    """mat_symp_error (sim_utils/matrix/mat_symp_error.f90:31-71): max |M^T S M - S|."""
    S = np.zeros((6, 6))
    for k in range(3):
        S[2 * k, 2 * k + 1] = 1
        S[2 * k + 1, 2 * k] = -1
    return float(np.max(np.abs(M.T @ S @ M - S)))


# --------------------------------------------------------------------------
# radiation (track1_radiation and radiation_map_setup restated in bmx_oracle.c)
# --------------------------------------------------------------------------
def _dual():
    global _lib_dual
    if _lib_dual is None:
        if not os.path.exists(LIB_DUAL):
            build()
        _lib_dual = C.CDLL(LIB_DUAL)
    L = _lib_dual
    if not getattr(L, "_rad_ready", False):
        L.oracle_jacobian.restype = C.c_int
        L.oracle_jacobian.argtypes = [C.POINTER(abi.EleRec), C.c_int, C.c_int, C.c_int, _DP, C.c_double, _DP, _DP]
        L.oracle_rad_map_setup.restype = C.c_int
        L.oracle_rad_map_setup.argtypes = [C.POINTER(abi.EleRec), C.c_int, _DP, C.c_double, C.c_double, C.c_double, _DP]
        L.oracle_make_mat6.restype = C.c_int
        L.oracle_make_mat6.argtypes = [C.POINTER(abi.EleRec), C.c_int, _DP, _DP, _DP]
        L.oracle_ele_slice.restype = C.c_int
        L.oracle_ele_slice.argtypes = [C.POINTER(abi.EleRec), C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(abi.EleRec)]
        L._rad_ready = True
    return L


def make_mat6(ele, species, vec):
    """(orbit at the exit, 6x6 Jacobian) of one element at vec (make_mat6)."""
    v = (C.c_double * 6)(*vec)
    out = (C.c_double * 6)()
    M = (C.c_double * 36)()
    if _dual().oracle_make_mat6(C.byref(ele), species, v, out, M) != 0:
        raise RuntimeError("particle lost in make_mat6")
    return np.array(out[:]), np.array(M[:]).reshape(6, 6)


def ele_slice(ele, l_slice, offset, inc_up, inc_down):
    out = abi.EleRec()
    if _dual().oracle_ele_slice(C.byref(ele), l_slice, offset, int(inc_up), int(inc_down), C.byref(out)) != 0:
        raise RuntimeError("element not sliceable by the oracle")
    return out


def rad_map_setup(ele, species, orb_in, total_length, g2_integral=0.0, g3_integral=0.0):
    """radiation_map_setup of one element about the reference orbit orb_in:
    float64 [2, 84] (rm0, rm1 in the bmx_rad_map layout)."""
    v = (C.c_double * 6)(*orb_in)
    rm = np.zeros(168)
    rc = _dual().oracle_rad_map_setup(C.byref(ele), species, v, total_length, g2_integral, g3_integral,
                                      rm.ctypes.data_as(_DP))
    if rc != 0:
        raise RuntimeError(f"oracle_rad_map_setup returned {rc}")
    return rm.reshape(2, 84)


def rad_maps(eles, species, orbits, g2_integral=0.0, g3_integral=0.0):
    """Radiation maps of every element about the reference orbit at each
    element's entrance (orbits [n_ele, 6]): [n_ele, 2, 84]."""
    arr = eles if isinstance(eles, C.Array) else abi_array(eles)
    total = float(sum(e.l for e in arr))
    return np.stack([rad_map_setup(arr[i], species, orbits[i], total, g2_integral, g3_integral)
                     for i in range(len(arr))])


def _rad_lib(L):
    if not getattr(L, "_rad_set", False):
        L.oracle_set_radiation.restype = None
        L.oracle_set_radiation.argtypes = [_DP, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_uint64]
        L.oracle_philox4x32_10.restype = None
        L.oracle_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32]
        L.oracle_rad_gauss6.restype = None
        L.oracle_rad_gauss6.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                        _DP]
        L._rad_set = True
    return L


def philox4x32_10(ctr, key):
    L = _rad_lib(lib())
    c = (C.c_uint32 * 4)(*ctr)
    L.oracle_philox4x32_10(c, key[0], key[1])
    return list(c)


def rad_gauss6(pid, ie, edge, turn, call, seed):
    L = _rad_lib(lib())
    g = (C.c_double * 6)()
    L.oracle_rad_gauss6(pid, ie, edge, turn, call, seed & 0xffffffff, seed >> 32, g)
    return np.array(g[:])


def track_bunch_rad(eles, species, n_turns, coords, maps, damping_on=True, fluctuations_on=False,
                    zero_average=False, synch_rad_scale=1.0, seed=0, call=0, fma=False, **kw):
    """track_bunch with track1_radiation at every element (maps [n_ele, 2, 84])."""
    L = _rad_lib(lib_fma() if fma else lib())
    m = np.ascontiguousarray(np.asarray(maps, dtype=np.float64))
    L.oracle_set_radiation(m.ctypes.data_as(_DP), int(damping_on), int(fluctuations_on), int(zero_average),
                           float(synch_rad_scale), int(seed), int(call))
    try:
        return track_bunch(eles, species, n_turns, coords, fma=fma, **kw)
    finally:
        L.oracle_set_radiation(None, 0, 0, 0, 1.0, 0, 0)
