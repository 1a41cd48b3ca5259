"""Shared test helpers: element construction from golden-fixture specs and
parity metrics."""
import numpy as np

from bmx import abi, lattice

FRINGE = {"full": abi.FRINGE_FULL, "hard_edge_only": abi.FRINGE_HARD_EDGE_ONLY, "basic_bend": abi.FRINGE_BASIC_BEND,
          "none": abi.FRINGE_NONE}
SPECIES = {"positron": abi.POSITRON, "electron": abi.ELECTRON, "antimuon": abi.ANTIMUON, "proton": abi.PROTON}
NAMES = ("x", "px", "y", "py", "z", "pz")


ENDS = {"entrance_end": abi.ENTRANCE_END, "exit_end": abi.EXIT_END, "both_ends": abi.BOTH_ENDS, "no_end": abi.NO_END}


def ele_from_spec(spec, species, e_tot=None, p0c=None):
    spec = dict(spec)
    key = spec.pop("key")
    if "fringe_type" in spec and isinstance(spec["fringe_type"], str):
        spec["fringe_type"] = FRINGE[spec["fringe_type"]]
    for k in ("coupler_at", "fringe_at", "aperture_at"):
        if isinstance(spec.get(k), str):
            spec[k] = ENDS[spec[k]]
    e = lattice.ele(key, **spec)
    lattice.set_reference([e], species, e_tot=e_tot, p0c=p0c)
    return e


def ring_from_fixture(fx, name):
    """Element list of one lattice of a ring fixture (elements, lines, use,
    lattices[name]): the lines expanded in order, the file's voltage sign and
    sextupole scale applied, the reference energy set, and every cavity at
    harmon times the revolution frequency (Bmad's harmon bookkeeping)."""
    opt = fx["lattices"][name]

    def expand(n):
        if n in fx["lines"]:
            return [x for m in fx["lines"][n] for x in expand(m)]
        return [n]

    species = SPECIES[fx["particle"]]
    eles = []
    for n in expand(fx["use"]):
        spec = dict(fx["elements"][n])
        if spec["key"] == "sextupole":
            spec["k2"] = spec["k2"] * opt["sextupole_scale"]
        if spec["key"] == "rfcavity":
            spec["voltage"] = spec["voltage"] * opt["voltage_sign"]
        eles.append(lattice.ele(spec.pop("key"), **spec))
    lattice.set_reference(eles, species, e_tot=fx["e_tot"])
    t_rev = lattice.revolution_time(eles)
    for e in eles:
        if e.key == abi.KEY_RFCAVITY:
            e.rf_frequency = fx["harmon"] / t_rev
    return eles


def printed_match(val, golden, digits=11):
    """True if val rounds to the golden value printed with `digits` significant digits."""
    if golden == 0:
        return abs(val) < 5e-21
    e = np.floor(np.log10(abs(golden)))
    half_ulp = 0.5 * 10.0 ** (e - (digits - 1))
    return abs(val - golden) <= half_ulp * 1.0001


def coord_errors(gpu, ref, mask, path_length):
    """Per-coordinate parity metric: max|gpu - ref| / scale with scale = max|ref|
    for x, px, y, py, pz and scale = max(max|z|, 1e-3 * path_length) for z.
    z is the small difference of path lengths of order the distance tracked,
    so its attainable accuracy is set by that distance, not by |z|."""
    out = {}
    for k in NAMES:
        a, b = np.asarray(gpu[k])[mask], np.asarray(ref[k])[mask]
        if a.size == 0:
            out[k] = 0.0
            continue
        scale = np.max(np.abs(b))
        if k == "z":
            scale = max(scale, 1e-3 * path_length)
        out[k] = float(np.max(np.abs(a - b)) / (scale + 1e-300))
    return out


def rel_err(gpu, ref, mask=None):
    """max over coordinates of max|gpu - ref| / max|ref| on the masked particles."""
    worst = 0.0
    for k in NAMES:
        a, b = np.asarray(gpu[k]), np.asarray(ref[k])
        if mask is not None:
            a, b = a[mask], b[mask]
        if a.size == 0:
            continue
        scale = np.max(np.abs(b)) + 1e-300
        worst = max(worst, float(np.max(np.abs(a - b)) / scale))
    return worst
