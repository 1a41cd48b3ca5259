"""bmx_track_bunch_noalloc (tracking with no device allocation: the loss
counter and the staging of host-resident arrays live in the library's static
per-device scratch) against bmx_track_bunch, through the C ABI: identical
coordinates, beta, states, loss elements, loss turns and loss counts, on
calls in a row of decreasing size (so that the scratch keeps the previous
call's contents), forward with a wrapping element range, backward, on
device-resident arrays larger than the staging area, and the size limit
refused without touching the device."""
import ctypes as C

import numpy as np
import pytest

from bmx import abi, beam, lattice, tracker

pytestmark = pytest.mark.gpu

FIELDS = ("x", "px", "y", "py", "z", "pz", "beta", "state", "ix_lost", "turn_lost")


def _host_call(lat, fn, coords, n_turns, direction=1, ix_start=0, ix_end=0):
    c = np.array(coords, dtype=np.float64, copy=True).reshape(6, -1)
    cols = [np.ascontiguousarray(c[i]) for i in range(6)]
    p, extra = tracker._host_particles(lat, cols)
    st = abi.Stats()
    rc = fn(lat.ctx.h, lat.h, int(ix_start), int(ix_end), int(n_turns), int(direction), C.byref(p), C.byref(st))
    out = dict(zip(tracker.Bunch.NAMES, cols))
    out.update(extra)
    return rc, out, st.n_lost


def _assert_same(a, b, what):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


@pytest.fixture(scope="module")
def ring(ctx):
    eles = lattice.fodo_ring_c0(n_cells=4)
    coords, _ = beam.c0_bunch(eles[:7], n=3000, seed=11)
    lat = ctx.lattice(eles, abi.ELECTRON)
    yield lat, coords
    lat.close()


def test_noalloc_matches_track_bunch_on_host_arrays(ring):
    lat, coords = ring
    lib = lat.ctx.lib
    big = coords * 20.0  # about a third of the particles are lost on the aperture
    print("noalloc: host arrays, 3000 particles x 28 elements x 3 turns", flush=True)
    rc0, ref, lost0 = _host_call(lat, lib.bmx_track_bunch, big, 3)
    rc1, got, lost1 = _host_call(lat, lib.bmx_track_bunch_noalloc, big, 3)
    assert rc0 == 0 and rc1 == 0, lib.bmx_last_error(lat.ctx.h)
    assert 0 < lost0 < 3000 and lost1 == lost0
    _assert_same(got, ref, "first call")
    # second call, smaller input, wrapping element range: the scratch still holds the first call's data
    small = big[:, :1000] * 0.9
    print("noalloc: host arrays, 1000 particles, elements 6..28 then 1..3, 2 turns", flush=True)
    rc0, ref, lost0 = _host_call(lat, lib.bmx_track_bunch, small, 2, ix_start=5, ix_end=3)
    rc1, got, lost1 = _host_call(lat, lib.bmx_track_bunch_noalloc, small, 2, ix_start=5, ix_end=3)
    assert rc0 == 0 and rc1 == 0, lib.bmx_last_error(lat.ctx.h)
    assert lost1 == lost0 and lost0 > 0
    _assert_same(got, ref, "second call")
    print("noalloc: host arrays, 500 particles, backward, 1 turn", flush=True)
    back = coords[:, :500] * 5.0
    rc0, ref, lost0 = _host_call(lat, lib.bmx_track_bunch, back, 1, direction=-1)
    rc1, got, lost1 = _host_call(lat, lib.bmx_track_bunch_noalloc, back, 1, direction=-1)
    assert rc0 == 0 and rc1 == 0, lib.bmx_last_error(lat.ctx.h)
    assert lost1 == lost0
    _assert_same(got, ref, "backward call")


def test_noalloc_device_arrays_beyond_the_staging_area(ring):
    lat, coords = ring
    n = abi.NOALLOC_MAX_PARTICLES + 4464  # device arrays are tracked in place at any size
    reps = -(-n // coords.shape[1])
    c = np.tile(coords * 15.0, reps)[:, :n]
    b0 = tracker.Bunch(c, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    b1 = tracker.Bunch(c, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    lib = lat.ctx.lib
    print(f"noalloc: device arrays, {n} particles x 28 elements x 2 turns", flush=True)
    tracker.track_bunch(lat, b0, n_turns=2)
    import torch
    torch.cuda.synchronize()
    p = b1.particles_struct()
    st = abi.Stats()
    rc = lib.bmx_track_bunch_noalloc(lat.ctx.h, lat.h, 0, 0, 2, 1, C.byref(p), C.byref(st))
    assert rc == 0, lib.bmx_last_error(lat.ctx.h)
    a, b = b0.numpy(), b1.numpy()
    assert int(np.sum(a["state"] != abi.ALIVE)) == st.n_lost > 0
    _assert_same(b, a, "device arrays")


def test_noalloc_refuses_host_bunches_beyond_its_scratch(ring):
    lat, coords = ring
    lib = lat.ctx.lib
    n = abi.NOALLOC_MAX_PARTICLES + 1
    c = np.tile(coords, -(-n // coords.shape[1]))[:, :n]
    rc, _, _ = _host_call(lat, lib.bmx_track_bunch_noalloc, c, 1)
    assert rc < 0
    assert b"65536" in lib.bmx_last_error(lat.ctx.h)
