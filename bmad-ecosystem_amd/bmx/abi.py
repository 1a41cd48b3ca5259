"""ctypes mirror of include/bmx.h (the libbmx C ABI).

The layout here must match the header byte for byte; `check_layout` compares
it with the sizes the compiled library reports through `bmx_sizeof`.
"""
import ctypes as C
import os

N_POLE = 22
NOALLOC_MAX_PARTICLES = 65536  # BMX_NOALLOC_MAX_PARTICLES

# particle states (bmad/modules/bmad_struct.f90:520-527)
NOT_SET, PRE_BORN, ALIVE, LOST, LOST_NEG_X, LOST_POS_X, LOST_NEG_Y, LOST_POS_Y, LOST_Z, LOST_PZ = (
    -1, 0, 1, 2, 3, 4, 5, 6, 7, 8)

# element keys
(KEY_MARKER, KEY_DRIFT, KEY_QUADRUPOLE, KEY_SBEND, KEY_SEXTUPOLE, KEY_OCTUPOLE, KEY_THICK_MULTIPOLE,
 KEY_MULTIPOLE, KEY_AB_MULTIPOLE, KEY_RFCAVITY, KEY_HKICKER, KEY_VKICKER, KEY_KICKER, KEY_INSTRUMENT,
 KEY_MONITOR, KEY_ECOLLIMATOR, KEY_RCOLLIMATOR, KEY_PIPE) = range(18)

KEY_NAMES = {
    KEY_MARKER: "marker", KEY_DRIFT: "drift", KEY_QUADRUPOLE: "quadrupole", KEY_SBEND: "sbend",
    KEY_SEXTUPOLE: "sextupole", KEY_OCTUPOLE: "octupole", KEY_THICK_MULTIPOLE: "thick_multipole",
    KEY_MULTIPOLE: "multipole", KEY_AB_MULTIPOLE: "ab_multipole", KEY_RFCAVITY: "rfcavity",
    KEY_HKICKER: "hkicker", KEY_VKICKER: "vkicker", KEY_KICKER: "kicker", KEY_INSTRUMENT: "instrument",
    KEY_MONITOR: "monitor", KEY_ECOLLIMATOR: "ecollimator", KEY_RCOLLIMATOR: "rcollimator", KEY_PIPE: "pipe",
}

# fringe types (bmad/modules/bmad_struct.f90:27,167-168)
FRINGE_NONE, FRINGE_SOFT_EDGE_ONLY, FRINGE_HARD_EDGE_ONLY, FRINGE_FULL, FRINGE_SAD_FULL, FRINGE_LINEAR_EDGE, \
    FRINGE_BASIC_BEND = 1, 2, 3, 4, 5, 6, 7
ENTRANCE_END, EXIT_END, BOTH_ENDS, NO_END = 1, 2, 3, 4
APERTURE_RECTANGULAR, APERTURE_ELLIPTICAL = 2, 3
ELECTRON, POSITRON, PROTON, ANTIPROTON, MUON, ANTIMUON = 1, 2, 3, 4, 5, 6
ATTR_K1, ATTR_K2, ATTR_K3 = 1, 2, 3

# physical constants (sim_utils/interfaces/physical_constants.f90:18-34)
C_LIGHT = 2.99792458e8
M_ELECTRON = 0.51099895069e6
M_PROTON = 0.93827208943e9
M_MUON = 105.6583755e6

MASS = {ELECTRON: M_ELECTRON, POSITRON: M_ELECTRON, PROTON: M_PROTON, ANTIPROTON: M_PROTON,
        MUON: M_MUON, ANTIMUON: M_MUON}
CHARGE = {ELECTRON: -1, POSITRON: 1, PROTON: 1, ANTIPROTON: -1, MUON: -1, ANTIMUON: 1}


class EleRec(C.Structure):
    _fields_ = [
        ("key", C.c_int32), ("is_on", C.c_int32), ("orientation", C.c_int32), ("num_steps", C.c_int32),
        ("fringe_type", C.c_int32), ("fringe_at", C.c_int32), ("aperture_type", C.c_int32),
        ("aperture_at", C.c_int32), ("field_master", C.c_int32), ("scale_multipoles", C.c_int32),
        ("has_mag_poles", C.c_int32), ("has_elec_poles", C.c_int32), ("coupler_at", C.c_int32),
        ("exact_multipoles", C.c_int32), ("reserved", C.c_int32 * 2),
    ] + [(n, C.c_double) for n in (
        "l k1 k2 k3 g dg e1 e2 fint fintx hgap hgapx tilt ref_tilt x_offset y_offset z_offset x_pitch y_pitch "
        "hkick vkick kick voltage rf_frequency phi0 phi0_multipass coupler_strength coupler_angle coupler_phase "
        "x1_limit x2_limit y1_limit y2_limit p0c e_tot p0c_start e_tot_start delta_ref_time ref_time "
        "r0_mag r0_elec px_aperture_width2 px_aperture_center py_aperture_width2 py_aperture_center "
        "z_aperture_width2 z_aperture_center pz_aperture_width2 pz_aperture_center").split()] + [
        ("a", C.c_double * N_POLE), ("b", C.c_double * N_POLE),
        ("a_elec", C.c_double * N_POLE), ("b_elec", C.c_double * N_POLE),
    ]


class Ref(C.Structure):
    _fields_ = [("species", C.c_int32), ("tracked_species", C.c_int32)]


class Config(C.Structure):
    _fields_ = [("aperture_limit_on", C.c_int32), ("direction", C.c_int32), ("max_aperture_limit", C.c_double),
                ("absolute_time_tracking", C.c_int32), ("reserved", C.c_int32)]


class Particles(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("x", "px", "y", "py", "z", "pz", "beta", "state", "ix_lost", "turn_lost",
                                          "t")] + [("n", C.c_int64), ("on_device", C.c_int32), ("reserved", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("n_live", C.c_int64), ("n_lost", C.c_int64), ("kernel_ms", C.c_double), ("pet", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("n_angle", C.c_int32), ("n_amp", C.c_int32), ("n_pz", C.c_int32), ("turns_per_block", C.c_int32),
                ("amp_max_x", C.c_double), ("amp_ratio_y", C.c_double), ("pz_min", C.c_double),
                ("pz_max", C.c_double), ("angle_min", C.c_double), ("angle_max", C.c_double)]


class Schedule(C.Structure):
    _fields_ = [("turns_per_block", C.c_int32), ("graph_depth", C.c_int32), ("probe_pet_per_s", C.c_double)]


class RadMap(C.Structure):
    """bmx_rad_map: one half-element radiation map (rad_map_struct), row-major matrices."""
    _fields_ = [("ref_orb", C.c_double * 6), ("damp_dmat", C.c_double * 36), ("xfer_damp_vec", C.c_double * 6),
                ("stoc_mat", C.c_double * 36)]


class Radiation(C.Structure):
    _fields_ = [("damping_on", C.c_int32), ("fluctuations_on", C.c_int32), ("zero_average", C.c_int32),
                ("reserved", C.c_int32), ("synch_rad_scale", C.c_double), ("seed", C.c_uint64)]


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libbmx.so")

# every entry point declared in include/bmx.h
EXPORTS = [
    "bmx_abi_version", "bmx_sizeof", "bmx_ctx_create", "bmx_ctx_destroy", "bmx_last_error",
    "bmx_lattice_upload", "bmx_lattice_upload_rad", "bmx_lattice_set_rad_stream", "bmx_lattice_upload_scaled",
    "bmx_lattice_destroy", "bmx_lattice_check", "bmx_lattice_cells", "bmx_track_bunch",
    "bmx_track_bunch_noalloc", "bmx_track_multi", "bmx_track_grid",
    "bmx_track_ics", "bmx_bunch_moments", "bmx_closed_orbit", "bmx_one_turn_matrix", "bmx_transfer_matrix", "bmx_ctx_set_schedule", "bmx_ctx_get_schedule",
    "bmx_tune_grid",
    "bmx_gemm_bf16", "bmx_colsum_f32", "bmx_adam_step", "bmx_f32_to_bf16", "bmx_mse_grad", "bmx_da_features", "bmx_step_fit", "bmx_mlp_forward", "bmx_da_screen", "bmx_ctx_sync",
    "bmx_surrogate_create", "bmx_surrogate_destroy", "bmx_surrogate_set_params", "bmx_surrogate_get_params",
    "bmx_surrogate_train", "bmx_surrogate_infer", "bmx_surrogate_screen",
]

_lib = None


def load_library(path=None):
    """Load libbmx.so (built by __graft_entry__.build()).  Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libbmx.so not found at {p}: build it with __graft_entry__.build()")
    lib = C.CDLL(p)
    lib.bmx_abi_version.restype = C.c_int32
    lib.bmx_sizeof.restype = C.c_int32
    lib.bmx_sizeof.argtypes = [C.c_int32]
    lib.bmx_ctx_create.restype = C.c_int32
    lib.bmx_ctx_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_uint32]
    lib.bmx_ctx_destroy.restype = None
    lib.bmx_ctx_destroy.argtypes = [C.c_void_p]
    lib.bmx_last_error.restype = C.c_char_p
    lib.bmx_last_error.argtypes = [C.c_void_p]
    lib.bmx_lattice_upload.restype = C.c_int32
    lib.bmx_lattice_upload.argtypes = [C.c_void_p, C.POINTER(EleRec), C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                       C.POINTER(Ref), C.POINTER(Config), C.POINTER(C.c_void_p)]
    lib.bmx_lattice_upload_rad.restype = C.c_int32
    lib.bmx_lattice_upload_rad.argtypes = [C.c_void_p, C.POINTER(EleRec), C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                           C.POINTER(RadMap), C.POINTER(Radiation), C.POINTER(Ref), C.POINTER(Config),
                                           C.POINTER(C.c_void_p)]
    lib.bmx_lattice_set_rad_stream.restype = C.c_int32
    lib.bmx_lattice_set_rad_stream.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
    lib.bmx_lattice_upload_scaled.restype = C.c_int32
    lib.bmx_lattice_upload_scaled.argtypes = [C.c_void_p, C.POINTER(EleRec), C.c_int32, C.c_int32, C.c_int32,
                                              C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_double),
                                              C.POINTER(Ref), C.POINTER(Config), C.POINTER(C.c_void_p)]
    lib.bmx_lattice_destroy.restype = None
    lib.bmx_lattice_destroy.argtypes = [C.c_void_p]
    lib.bmx_lattice_check.restype = C.c_int32
    lib.bmx_lattice_check.argtypes = [C.POINTER(EleRec), C.c_int32, C.POINTER(Ref), C.POINTER(Config),
                                      C.POINTER(C.c_int64), C.c_char_p, C.c_int32]
    lib.bmx_lattice_cells.restype = C.c_int32
    lib.bmx_lattice_cells.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
    lib.bmx_track_bunch.restype = C.c_int32
    lib.bmx_track_bunch.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    C.POINTER(Particles), C.POINTER(Stats)]
    lib.bmx_track_bunch_noalloc.restype = C.c_int32
    lib.bmx_track_bunch_noalloc.argtypes = lib.bmx_track_bunch.argtypes
    lib.bmx_track_multi.restype = C.c_int32
    lib.bmx_track_multi.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_int32, C.POINTER(Particles),
                                    C.POINTER(Stats)]
    lib.bmx_track_grid.restype = C.c_int32
    lib.bmx_track_grid.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Grid), C.POINTER(C.c_double), C.c_int32,
                                   C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(Stats)]
    lib.bmx_track_ics.restype = C.c_int32
    lib.bmx_track_ics.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32,
                                  C.c_int32, C.POINTER(C.c_int32), C.POINTER(Stats)]
    lib.bmx_bunch_moments.restype = C.c_int32
    lib.bmx_bunch_moments.argtypes = [C.c_void_p, C.POINTER(Particles), C.POINTER(C.c_double)]
    lib.bmx_ctx_set_schedule.restype = C.c_int32
    lib.bmx_ctx_set_schedule.argtypes = [C.c_void_p, C.POINTER(Schedule)]
    lib.bmx_ctx_get_schedule.restype = C.c_int32
    lib.bmx_ctx_get_schedule.argtypes = [C.c_void_p, C.POINTER(Schedule)]
    lib.bmx_tune_grid.restype = C.c_int32
    lib.bmx_tune_grid.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Grid), C.POINTER(C.c_double), C.c_int32,
                                  C.POINTER(Schedule)]
    lib.bmx_closed_orbit.restype = C.c_int32
    lib.bmx_closed_orbit.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                     C.POINTER(Stats)]
    lib.bmx_one_turn_matrix.restype = C.c_int32
    lib.bmx_one_turn_matrix.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(Stats)]
    lib.bmx_mlp_forward.restype = C.c_int32
    lib.bmx_mlp_forward.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_int64, C.c_void_p]
    lib.bmx_da_screen.restype = C.c_int32
    lib.bmx_da_screen.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_int32,
                                  C.c_int32, C.c_void_p, C.c_float, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
                                  C.c_float, C.c_void_p]
    lib.bmx_surrogate_create.restype = C.c_int32
    lib.bmx_surrogate_create.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                         C.POINTER(C.c_int64), C.POINTER(C.c_void_p)]
    lib.bmx_surrogate_destroy.restype = None
    lib.bmx_surrogate_destroy.argtypes = [C.c_void_p]
    lib.bmx_surrogate_set_params.restype = C.c_int32
    lib.bmx_surrogate_set_params.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    lib.bmx_surrogate_get_params.restype = C.c_int32
    lib.bmx_surrogate_get_params.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    lib.bmx_surrogate_train.restype = C.c_int32
    lib.bmx_surrogate_train.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_float,
                                        C.c_float, C.c_uint64, C.c_int32, C.POINTER(C.c_float)]
    lib.bmx_surrogate_infer.restype = C.c_int32
    lib.bmx_surrogate_infer.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    lib.bmx_surrogate_screen.restype = C.c_int32
    lib.bmx_surrogate_screen.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_float, C.c_int32, C.c_int32,
                                         C.c_void_p, C.c_int32, C.c_float, C.c_void_p]
    lib.bmx_transfer_matrix.restype = C.c_int32
    lib.bmx_transfer_matrix.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_int32)]
    lib.bmx_gemm_bf16.restype = C.c_int32
    lib.bmx_gemm_bf16.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                  C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                  C.c_int32, C.c_void_p]
    lib.bmx_colsum_f32.restype = C.c_int32
    lib.bmx_colsum_f32.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]
    lib.bmx_adam_step.restype = C.c_int32
    lib.bmx_adam_step.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32]
    lib.bmx_f32_to_bf16.restype = C.c_int32
    lib.bmx_f32_to_bf16.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    lib.bmx_mse_grad.restype = C.c_int32
    lib.bmx_mse_grad.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.bmx_da_features.restype = C.c_int32
    lib.bmx_da_features.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_float,
                                    C.c_void_p, C.c_void_p]
    lib.bmx_step_fit.restype = C.c_int32
    lib.bmx_step_fit.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_float, C.c_void_p]
    lib.bmx_ctx_sync.restype = C.c_int32
    lib.bmx_ctx_sync.argtypes = [C.c_void_p]
    if path is None:
        _lib = lib
    return lib


def check_layout(lib):
    """Compare ctypes struct sizes with the compiled library's."""
    sizes = {0: EleRec, 1: Particles, 2: Stats, 3: Grid, 4: Config, 5: Schedule, 6: RadMap, 7: Radiation}
    bad = {k: (C.sizeof(v), lib.bmx_sizeof(k)) for k, v in sizes.items() if C.sizeof(v) != lib.bmx_sizeof(k)}
    if bad:
        raise RuntimeError(f"ABI layout mismatch (ctypes, library): {bad}")
    return True
