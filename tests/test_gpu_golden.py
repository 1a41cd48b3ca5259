"""The device path against the reference's own golden vectors, with no oracle
in between: every Bmad_Standard row of regression_tests/tracking_method_test
in the forward variant (direction +1, orientation +1, relative time) and in
the -Anti_O variant (orientation -1 and the antiparticle tracked, except for
sbend and rfcavity, which keep both: tracking_method_test.f90:158-199), and
the backward -Anti_D / -Anti_OD variants (direction = -1, the species flipped
where direction x orientation = -1), whose element the device path accepts,
tracked through bmx_track_bunch and compared with the values the reference
printed (11 significant digits).  The backward rf rows are tracked with
absolute time there and are refused by the device path.  Elements the
device rejects are skipped with the reason bmx_lattice_check gives."""
import json
import os

import numpy as np
import pytest

from bmx import abi, tracker
from helpers import SPECIES, ele_from_spec, printed_match

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TM = json.load(open(os.path.join(GOLD, "tracking_method_bmad_standard.json")))
ROWS = [r for r in TM["rows"] if r["variant"] in ("", "-Anti_O", "-Anti_D", "-Anti_OD")]


@pytest.mark.parametrize("row", ROWS, ids=[r["name"] + r["variant"] for r in ROWS])
def test_device_tracking_method_row(ctx, row):
    sp = SPECIES[TM["particle"]]
    spec = TM["elements"][row["name"]]
    e = ele_from_spec(spec, sp, e_tot=TM["e_tot"])
    direction, osign, _ = TM["variants"][row["variant"]]
    if spec["key"] not in ("sbend", "rfcavity"):
        e.orientation = osign
    # the harness flips the species where direction x orientation = -1
    tracked = abi.POSITRON + abi.ELECTRON - sp if direction * e.orientation == -1 else None
    ok, _, msg = tracker.check_lattice([e], sp, tracked_species=tracked, direction=direction)
    if not ok:
        pytest.skip(f"device path rejects the element: {msg}")
    n = 64  # one wave of copies of the golden start vector
    coords = np.repeat(np.asarray(TM["start"], dtype=np.float64)[:, None], n, axis=1)
    lat = ctx.lattice([e], sp, tracked_species=tracked)
    b = tracker.Bunch(coords, lat.p0c, abi.MASS[sp], e_tot=lat.e_tot)
    tracker.track_bunch(lat, b, n_turns=1, direction=direction)
    out = b.numpy()
    lat.close()
    assert np.all(out["state"] == abi.ALIVE)
    vec = [float(out[k][0]) for k in ("x", "px", "y", "py", "z", "pz")]
    for k in ("x", "px", "y", "py", "z", "pz"):
        assert np.all(out[k] == out[k][0]), "copies of one particle must track identically"
    bad = [(i, v, g) for i, (v, g) in enumerate(zip(vec, row["vec"])) if not printed_match(v, g, TM["printed_digits"])]
    print("golden:", row["name"] + row["variant"], " ".join(f"{v - g:+.1e}" for v, g in zip(vec, row["vec"])), flush=True)
    assert not bad, f"coordinates off the printed golden value: {bad}"
