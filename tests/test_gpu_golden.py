"""The device path against the reference's own golden vectors, with no oracle
in between: every Bmad_Standard row of regression_tests/tracking_method_test
in the forward variant (direction +1, orientation +1, relative time) and in
the -Anti_O variant (orientation -1 and the antiparticle tracked, except for
sbend and rfcavity, which keep both: tracking_method_test.f90:158-199), and
the backward -Anti_D / -Anti_OD variants (direction = -1, the species flipped
where direction x orientation = -1), whose element the device path accepts,
tracked through bmx_track_bunch and compared with the values the reference
printed (11 significant digits).  The backward rows and the -Abs_Time rows
are tracked with absolute time, as the reference harness does
(tracking_method_test.f90:170-175), from init_coord's particle time at the
start edge (init_coord.f90:222).  Elements the device rejects are skipped
with the reason bmx_lattice_check gives."""
import json
import os

import numpy as np
import pytest

from bmx import abi, tracker
from helpers import SPECIES, ele_from_spec, printed_match

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TM = json.load(open(os.path.join(GOLD, "tracking_method_bmad_standard.json")))
ROWS = TM["rows"]


@pytest.mark.parametrize("row", ROWS, ids=[r["name"] + r["variant"] for r in ROWS])
def test_device_tracking_method_row(ctx, row):
    sp = SPECIES[TM["particle"]]
    spec = TM["elements"][row["name"]]
    e = ele_from_spec(spec, sp, e_tot=TM["e_tot"])
    direction, osign, abs_time = TM["variants"][row["variant"]]
    abs_time = abs_time or direction == -1
    if spec["key"] not in ("sbend", "rfcavity"):
        e.orientation = osign
    # the harness flips the species where direction x orientation = -1
    tracked = abi.POSITRON + abi.ELECTRON - sp if direction * e.orientation == -1 else None
    ok, _, msg = tracker.check_lattice([e], sp, tracked_species=tracked, direction=direction,
                                       absolute_time_tracking=abs_time)
    if not ok:
        pytest.skip(f"device path rejects the element: {msg}")
    n = 64  # one wave of copies of the golden start vector
    coords = np.repeat(np.asarray(TM["start"], dtype=np.float64)[:, None], n, axis=1)
    lat = ctx.lattice([e], sp, tracked_species=tracked, absolute_time_tracking=abs_time)
    # init_coord at the start edge: the entrance forward, the exit backward
    pc = lat.p0c * (1.0 + coords[5])
    beta = pc / np.sqrt(pc * pc + abi.MASS[sp] ** 2)
    t_edge = e.ref_time - (e.delta_ref_time if direction == 1 else 0.0)
    t0 = t_edge - coords[4] / (beta * abi.C_LIGHT)
    b = tracker.Bunch(coords, lat.p0c, abi.MASS[sp], e_tot=lat.e_tot, t=t0 if abs_time else None)
    ix = 0 if direction == 1 else 1
    tracker.track_bunch(lat, b, n_turns=1, direction=direction, ix_start=ix, ix_end=0)
    out = b.numpy()
    lat.close()
    assert np.all(out["state"] == abi.ALIVE)
    vec = [float(out[k][0]) for k in ("x", "px", "y", "py", "z", "pz")]
    for k in ("x", "px", "y", "py", "z", "pz"):
        assert np.all(out[k] == out[k][0]), "copies of one particle must track identically"
    bad = [(i, v, g) for i, (v, g) in enumerate(zip(vec, row["vec"])) if not printed_match(v, g, TM["printed_digits"])]
    print("golden:", row["name"] + row["variant"], " ".join(f"{v - g:+.1e}" for v, g in zip(vec, row["vec"])), flush=True)
    assert not bad, f"coordinates off the printed golden value: {bad}"
