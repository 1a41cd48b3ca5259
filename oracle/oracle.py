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
    global _lib_dual
    if _lib_dual is None:
        if not os.path.exists(LIB_DUAL):
            build()
        _lib_dual = C.CDLL(LIB_DUAL)
        _lib_dual.oracle_jacobian.restype = C.c_int
        _lib_dual.oracle_jacobian.argtypes = [C.POINTER(abi.EleRec), C.c_int, C.c_int, C.c_int, _DP, C.c_double, _DP, _DP]
    arr = eles if isinstance(eles, C.Array) else abi_array(eles)
    v = (C.c_double * 6)(*vec)
    out = (C.c_double * 6)()
    J = (C.c_double * 36)()
    beta0 = arr[0].p0c_start / arr[0].e_tot_start
    rc = _lib_dual.oracle_jacobian(arr, len(arr), species, n_turns, v, beta0, out, J)
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
