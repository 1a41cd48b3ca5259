"""The Fortran side of the drop-in (bmad-ecosystem_amd/fortran): a program
built with the image's flang from bmx_c_binding.f90 and bmx_shim_core.f90
reads elements the way Bmad keeps them (ele%key, ele%value(1:75) by Bmad's
attribute indices, the ele_struct flags, the pole arrays) and a bunch as
bunch%particle(:) records, flattens the elements (bmx_flatten_element),
packs the bunch and -- on a GPU -- tracks it through bmx_c_binding and
libbmx, and checks the track_many_hook contract one particle at a time.

CPU: the flattened records equal, byte for byte, the bmx_ele_rec the
elements were written from.  GPU: the Fortran-driven tracking equals the
oracle's; the track_many contract (track_state = moving_forward or the
1-based element of loss) matches the oracle's loss element."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bmx import abi, beam, lattice
import oracle
from bmad_layout import bmad_record

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORT = os.path.join(ROOT, "bmad-ecosystem_amd", "fortran")
LIBDIR = os.path.join(ROOT, "bmad-ecosystem_amd", "bmx")
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"

DRIVER = """program drive
use, intrinsic :: iso_c_binding
use bmx_c_binding
use bmx_shim_core
implicit none
integer(c_int32_t) :: hdr(9)
real(c_double) :: value(75), ref_time, a(0:21), b(0:21), ae(0:21), be(0:21)
type(bmx_ele_rec), allocatable, target :: recs(:)
type(bmx_shim_bunch), target :: bun
type(bmx_stats) :: st
type(bmx_ref) :: ref
type(bmx_config) :: cfg
type(c_ptr) :: ctx, lat
real(c_double), allocatable :: vec(:,:), beta(:), t(:)
integer, allocatable :: state(:)
real(c_double) :: v1(6), b1
integer :: n_ele, n, i, mode, n_turns, s1, ts, n_one
integer(c_int32_t) :: rc
logical :: ok
open (10, file = 'args.txt')
read (10, *) mode, n_ele, n, n_turns, n_one
close (10)
allocate(recs(n_ele))
open (11, file = 'elements.bin', access = 'stream', form = 'unformatted')
do i = 1, n_ele
  read (11) hdr, value, ref_time, a, b, ae, be
  call bmx_flatten_element(int(hdr(1)), value, ref_time, hdr(2) /= 0, int(hdr(3)), hdr(4) /= 0, hdr(5) /= 0, &
                           int(hdr(6)), int(hdr(7)), hdr(8) /= 0, a, b, hdr(9) /= 0, ae, be, recs(i), ok)
  if (.not. ok) stop 3
enddo
close (11)
open (12, file = 'recs.bin', access = 'stream', form = 'unformatted', status = 'replace')
write (12) recs
close (12)
if (mode == 0) stop
allocate(vec(6, n), beta(n), t(n), state(n))
open (13, file = 'particles.bin', access = 'stream', form = 'unformatted')
read (13) vec, beta
close (13)
t = 0
state = bmx_alive
rc = bmx_ctx_create(ctx, 0, 0)
if (rc /= 0) stop 4
ref%species = bmx_species_from_bmad(-1)
rc = bmx_lattice_upload(ctx, recs, int(n_ele, c_int32_t), c_null_ptr, 1, ref, cfg, lat)
if (rc /= 0) then
  print *, bmx_error_string(ctx)
  stop 5
endif
call bmx_shim_pack(vec, state, beta, t, bun)
rc = bmx_shim_track(ctx, lat, 0, 0, n_turns, 1, bun, st, .false.)
if (rc /= 0) stop 6
call bmx_shim_unpack(bun, vec, state, beta, t)
open (14, file = 'out.bin', access = 'stream', form = 'unformatted', status = 'replace')
write (14) vec, beta, state, bun%ix_lost, bun%turn_lost
! the track_many_hook contract, one particle and one turn at a time
open (15, file = 'particles.bin', access = 'stream', form = 'unformatted')
read (15) vec, beta
close (15)
do i = 1, n_one
  v1 = vec(:, i)
  b1 = beta(i)
  s1 = bmx_alive
  rc = bmx_shim_track_one(ctx, lat, 0, 1, v1, b1, s1, ts)
  if (rc /= 0) stop 7
  write (14) v1, s1, ts
enddo
close (14)
call bmx_lattice_destroy(lat)
call bmx_ctx_destroy(ctx)
end program
"""


def _elements():
    """The C0 ring of 4 cells, then one of each further element kind the
    flattener has a branch for."""
    eles = lattice.fodo_ring_c0(n_cells=4)
    extra = [
        lattice.ele("rfcavity", l=0.5, voltage=2e5, rf_frequency=5e8, phi0=0.05, num_steps=5, coupler_strength=0.05,
                    coupler_angle=0.2, coupler_at=abi.BOTH_ENDS, x_limit=0.01, y_limit=0.01,
                    aperture_type=abi.APERTURE_ELLIPTICAL),
        lattice.ele("sbend", l=0.8, angle=0.01, e1=0.01, e2=-0.02, fint=0.4, hgap=0.01, k1=0.05, ref_tilt=0.02,
                    num_steps=2, x_offset=1e-4, y_pitch=2e-5, x_limit=0.01, y_limit=0.01,
                    aperture_type=abi.APERTURE_ELLIPTICAL),
        lattice.ele("hkicker", l=0.2, kick=1e-5, tilt=0.1, num_steps=1),
        lattice.ele("ecollimator", l=0.3, x_limit=0.01, y_limit=0.008, px_aperture_width2=1e-2,
                    pz_aperture_width2=5e-2),
        lattice.ele("sextupole", l=0.2, k2=3.0, num_steps=2, tilt=0.05, y_offset=2e-5, hkick=1e-6,
                    x_limit=0.01, y_limit=0.01, aperture_type=abi.APERTURE_ELLIPTICAL),
        lattice.ele("quadrupole", l=0.3, k1=0.2, num_steps=3, b={2: 0.5}, a_elec={1: 1e3}, x_limit=0.01,
                    y_limit=0.01, aperture_type=abi.APERTURE_ELLIPTICAL),
        lattice.ele("octupole", l=0.2, k3=20.0, num_steps=2),
        lattice.ele("multipole", knl={2: 0.3, 3: 5.0}, tn={2: 0.1}),
        lattice.ele("marker"),
    ]
    return lattice.set_reference(eles + extra, abi.ELECTRON, e_tot=1e9)


def _run(tmp_path, eles, mode, coords=None, n_turns=1, n_one=0):
    if not os.path.exists(FLANG):
        pytest.skip("no Fortran compiler in this image")
    for name in ("bmx_c_binding", "bmx_shim_core"):
        subprocess.check_call([FLANG, "-c", os.path.join(FORT, name + ".f90"), "-o", name + ".o"], cwd=tmp_path)
    (tmp_path / "drive.f90").write_text(DRIVER)
    subprocess.check_call([FLANG, "drive.f90", "bmx_c_binding.o", "bmx_shim_core.o", f"-L{LIBDIR}", "-lbmx",
                           f"-Wl,-rpath,{LIBDIR}", "-o", "drive"], cwd=tmp_path)
    with open(tmp_path / "elements.bin", "wb") as f:
        for e in eles:
            f.write(bmad_record(e))
    n = 0 if coords is None else coords.shape[1]
    (tmp_path / "args.txt").write_text(f"{mode} {len(eles)} {n} {n_turns} {n_one}\n")
    if coords is not None:
        aos = np.ascontiguousarray(coords.T)  # vec(6, n) column-major = particle records one after another
        beta = oracle.beta_of(coords[5], eles[0].p0c, abi.M_ELECTRON)
        beta = np.where(coords[5] == 0, eles[0].p0c / eles[0].e_tot, beta)
        with open(tmp_path / "particles.bin", "wb") as f:
            f.write(aos.astype("<f8").tobytes() + beta.astype("<f8").tobytes())
    print(f"fortran shim: {len(eles)} elements, {n} particles, mode {mode}", flush=True)
    subprocess.check_call([str(tmp_path / "drive")], cwd=tmp_path)
    return tmp_path


def test_flattener_round_trip(tmp_path):
    """Every element, written in Bmad's layout, flattens back to its record."""
    eles = _elements()
    _run(tmp_path, eles, 0)
    got = (tmp_path / "recs.bin").read_bytes()
    want = b"".join(bytes(e) for e in lattice.to_array(eles))
    size = len(bytes(eles[0]))
    assert len(got) == len(want)
    bad = [i for i in range(len(eles)) if got[i * size:(i + 1) * size] != want[i * size:(i + 1) * size]]
    assert not bad, [abi.KEY_NAMES[eles[i].key] for i in bad]


@pytest.mark.gpu
def test_fortran_driven_tracking_matches_oracle(tmp_path):
    """A bunch handed over as Fortran particle records, flattened lattice,
    bmx_c_binding's bmx_track_bunch through bmx_shim_track, 2 turns (the
    extra elements make the ring lose about 85 % of the bunch over them):
    states and loss elements equal to the oracle's, coordinates within
    max(1e-12, 4 x the oracle's own plain-vs-FMA spread); then the
    track_many contract for 64 particles, one turn each."""
    from helpers import coord_errors
    eles = _elements()
    coords, _ = beam.c0_bunch(eles[:7], n=2048, seed=31)
    coords *= 8.0  # about half the bunch is lost on the apertures in the first turn
    _run(tmp_path, eles, 1, coords, n_turns=2, n_one=64)
    raw = (tmp_path / "out.bin").read_bytes()
    n = coords.shape[1]
    o = 0
    vec = np.frombuffer(raw, "<f8", 6 * n, o).reshape(n, 6).T
    o += 48 * n
    o += 8 * n  # beta
    state = np.frombuffer(raw, "<i4", n, o)
    o += 4 * n
    ix_lost = np.frombuffer(raw, "<i4", n, o)
    o += 8 * n  # ix_lost, turn_lost
    ref = oracle.track_bunch(eles, abi.ELECTRON, 2, coords)
    fma = oracle.track_bunch(eles, abi.ELECTRON, 2, coords, fma=True)
    assert np.array_equal(state, ref["state"])
    lost = ref["state"] != abi.ALIVE
    assert lost.any() and (~lost).any()
    assert np.array_equal(ix_lost[lost], ref["ix_lost"][lost])
    g = {k: vec[i] for i, k in enumerate(("x", "px", "y", "py", "z", "pz"))}
    path = 2 * sum(e.l for e in eles)
    errs = coord_errors(g, ref, ~lost, path)
    tol = max(1e-12, 4 * max(coord_errors(fma, ref, ~lost, path).values()))
    print("fortran shim parity:", {k: f"{v:.1e}" for k, v in errs.items()}, f"bar {tol:.1e}", flush=True)
    assert max(errs.values()) <= tol
    one = oracle.track_bunch(eles, abi.ELECTRON, 1, coords[:, :64])
    rec = np.dtype([("v", "<f8", 6), ("s", "<i4"), ("ts", "<i4")])
    r = np.frombuffer(raw[o:], rec, 64)
    assert np.array_equal(r["s"], one["state"])
    want_ts = np.where(one["state"] == abi.ALIVE, -9, one["ix_lost"] + 1)
    assert np.array_equal(r["ts"], want_ts)
    alive = one["state"] == abi.ALIVE
    g1 = {k: r["v"][:, i] for i, k in enumerate(("x", "px", "y", "py", "z", "pz"))}
    assert max(coord_errors(g1, one, alive, sum(e.l for e in eles)).values()) <= 1e-12
