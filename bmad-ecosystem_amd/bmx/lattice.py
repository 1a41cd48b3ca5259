"""Lattice records for libbmx: element constructors with Bmad's defaults,
reference-energy bookkeeping, and the benchmark ring lattices.

Element defaults follow bmad/parsing/set_ele_defaults.f90 (fringe_type none,
fringe_at both_ends; sbend fringe basic_bend; rfcavity fringe full, 10 steps;
ecollimator elliptical aperture) and the step rules of
bmad/code/attribute_bookkeeper.f90:296-404 (drift-like elements one step).
num_steps must be given explicitly where Bmad would ask PTC's check_bend
for it (sbend, quadrupole, sextupole, kickers).
"""
import math

import numpy as np

from . import abi

_KEYS = {v: k for k, v in abi.KEY_NAMES.items()}

# default tilt of a bare "tilt" keyword: pi / (2 (n+1)) for the element's order
DEFAULT_TILT = {"quadrupole": math.pi / 4, "sextupole": math.pi / 6, "octupole": math.pi / 8}


def ele(key, **kw):
    """Build one abi.EleRec.

    key: element type name ("drift", "quadrupole", "sbend", ...).
    Keyword attributes use Bmad names (l, k1, k2, k3, g, angle, dg, e1, e2,
    fint, fintx, hgap, hgapx, tilt, ref_tilt, x_offset, y_offset, z_offset,
    x_pitch, y_pitch, hkick, vkick, kick, voltage, rf_frequency, phi0,
    x_limit, y_limit, x1_limit, ..., num_steps, fringe_type, fringe_at,
    aperture_type, aperture_at, orientation, is_on, r0_mag, r0_elec).
    Multipoles: a={n: An}, b={n: Bn}, a_elec, b_elec; for "multipole"
    knl={n: KnL}, tn={n: Tn}.
    """
    e = abi.EleRec()
    if key not in _KEYS:
        raise ValueError(f"unknown element key {key!r}")
    k = _KEYS[key]
    e.key = k
    e.is_on = 1
    e.orientation = 1
    e.fringe_type = abi.FRINGE_NONE
    e.fringe_at = abi.BOTH_ENDS
    e.aperture_type = abi.APERTURE_RECTANGULAR
    e.aperture_at = abi.EXIT_END
    e.scale_multipoles = 1
    e.exact_multipoles = 1
    e.coupler_at = abi.EXIT_END
    if key == "sbend":
        e.fringe_type = abi.FRINGE_BASIC_BEND
    if key == "rfcavity":
        e.fringe_type = abi.FRINGE_FULL
    if key == "ecollimator":
        e.aperture_type = abi.APERTURE_ELLIPTICAL
    kw = dict(kw)
    l = float(kw.pop("l", 0.0))
    e.l = l
    if "angle" in kw:
        ang = float(kw.pop("angle"))
        kw["g"] = ang / l if l != 0 else 0.0
    if "tilt" in kw and kw["tilt"] is True:
        kw["tilt"] = DEFAULT_TILT[key]
    if "x_limit" in kw:
        v = float(kw.pop("x_limit"))
        kw.setdefault("x1_limit", v)
        kw.setdefault("x2_limit", v)
    if "y_limit" in kw:
        v = float(kw.pop("y_limit"))
        kw.setdefault("y1_limit", v)
        kw.setdefault("y2_limit", v)
    for name, arr in (("a", "a"), ("b", "b"), ("a_elec", "a_elec"), ("b_elec", "b_elec")):
        if name in kw:
            d = kw.pop(name)
            for n, v in d.items():
                getattr(e, arr)[int(n)] = float(v)
            if name in ("a", "b"):
                e.has_mag_poles = 1
            else:
                e.has_elec_poles = 1
    if key == "multipole":
        knl = kw.pop("knl", {})
        tn = kw.pop("tn", {})
        for n, v in knl.items():
            e.a[int(n)] = float(v)
        for n, v in tn.items():
            e.b[int(n)] = float(v)
        e.has_mag_poles = 1 if knl else 0
    int_fields = {"num_steps", "fringe_type", "fringe_at", "aperture_type", "aperture_at", "orientation", "is_on",
                  "field_master", "scale_multipoles", "coupler_at", "exact_multipoles"}
    for name, v in kw.items():
        if name in int_fields:
            setattr(e, name, int(v))
        else:
            if not hasattr(e, name):
                raise ValueError(f"unknown attribute {name!r}")
            setattr(e, name, float(v))
    if key in ("sbend",):
        if "fintx" not in kw:
            e.fintx = e.fint
        if "hgapx" not in kw:
            e.hgapx = e.hgap
    if e.num_steps == 0:
        if key in ("drift", "pipe", "ecollimator", "rcollimator", "instrument", "monitor", "marker",
                   "multipole", "ab_multipole") or l == 0:
            e.num_steps = 1
        elif key == "rfcavity":
            e.num_steps = 10
        elif key in ("octupole", "thick_multipole"):
            # ds_step = bmad_com%default_ds_step = 0.2 m (attribute_bookkeeper.f90:390-400)
            e.num_steps = max(1, _nint(abs(l) / 0.2))
        else:
            raise ValueError(f"{key}: set num_steps explicitly (Bmad would call PTC's check_bend)")
    return e


def _nint(x):
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def set_reference(eles, species, e_tot=None, p0c=None):
    """Fill p0c, e_tot, p0c_start, e_tot_start, delta_ref_time, ref_time
    along the element list (constant reference energy: no lcavity/e_gun)."""
    m = abi.MASS[species]
    if e_tot is None:
        e_tot = math.sqrt(p0c * p0c + m * m)
    if p0c is None:
        p0c = math.sqrt(e_tot * e_tot - m * m)
    beta0 = p0c / e_tot
    t = 0.0
    for e in eles:
        e.p0c = e.p0c_start = p0c
        e.e_tot = e.e_tot_start = e_tot
        e.delta_ref_time = e.l / (beta0 * abi.C_LIGHT)
        t += e.delta_ref_time
        e.ref_time = t
    return eles


def to_array(eles):
    arr = (abi.EleRec * len(eles))()
    for i, e in enumerate(eles):
        arr[i] = e
    return arr


# ---------------------------------------------------------------------------
# benchmark lattices
# ---------------------------------------------------------------------------

def fodo_ring_c0(n_cells=100, aperture=0.01, species=abi.ELECTRON, e_tot=1e9):
    """Ring of n_cells FODO cells QF SF D B D QD SD: quads L=0.5 m, k1=+-0.8 m^-2,
    thin sextupoles k2L=+-1.5 m^-2 next to each quad, drifts 1.0 m, sector bend
    L=2.0 m, angle=2 pi/200, hard-edge fringe.  Elliptical 1 cm aperture on
    every element at the exit end."""
    ap = dict(x_limit=aperture, y_limit=aperture, aperture_type=abi.APERTURE_ELLIPTICAL)
    eles = []
    for _ in range(n_cells):
        eles += [
            ele("quadrupole", l=0.5, k1=0.8, num_steps=1, **ap),
            ele("multipole", knl={2: 1.5}, **ap),
            ele("drift", l=1.0, **ap),
            ele("sbend", l=2.0, angle=2 * math.pi / 200, fringe_type=abi.FRINGE_HARD_EDGE_ONLY, num_steps=1, **ap),
            ele("drift", l=1.0, **ap),
            ele("quadrupole", l=0.5, k1=-0.8, num_steps=1, **ap),
            ele("multipole", knl={2: -1.5}, **ap),
        ]
    return set_reference(eles, species, e_tot=e_tot)


def ring_c1(n_cells=500, n_rf=8, seed=20261203, species=abi.ELECTRON, e_tot=6e9, x_ap=0.02, y_ap=0.01,
            k1_err=None, k2_err=None, offsets=None):
    """Large ring from the C0 cell pattern: cell QF SF(thick) OC D B D QD SD(thick),
    8 elements per cell.  Thick sextupoles L=0.2 m with k2 L = +-1.5 at 4 kick-drift
    slices, thin octupoles k3L ~ U(-50, 50) (seeded), bends closing the ring
    (angle 2 pi / n_cells), n_rf of the 1.0 m drifts replaced by 1 MV
    rfcavities at phi0 = 0 (stable above transition) whose frequency is the
    harmonic of the revolution frequency nearest 500 MHz (h = 4503 on the
    500-cell ring), so that the reference particle sees the same rf phase on
    every turn in absolute time as in relative time; elliptical aperture
    x_ap x y_ap.  Optional per-element relative errors (arrays indexed by
    quad / sextupole count) and transverse offsets."""
    rng = np.random.default_rng(seed)
    ap = dict(x_limit=x_ap, y_limit=y_ap, aperture_type=abi.APERTURE_ELLIPTICAL)
    k3l = rng.uniform(-50.0, 50.0, size=n_cells)
    rf_cells = set(int(round(i * n_cells / n_rf)) for i in range(n_rf)) if n_rf > 0 else set()
    eles = []
    iq = isx = 0
    for c in range(n_cells):
        def qerr():
            nonlocal iq
            f = 1.0 + (k1_err[iq] if k1_err is not None else 0.0)
            off = offsets[("q", iq)] if offsets is not None else (0.0, 0.0)
            iq += 1
            return f, off

        def serr():
            nonlocal isx
            f = 1.0 + (k2_err[isx] if k2_err is not None else 0.0)
            off = offsets[("s", isx)] if offsets is not None else (0.0, 0.0)
            isx += 1
            return f, off

        fq, oq = qerr()
        fs, os_ = serr()
        cell = [
            ele("quadrupole", l=0.5, k1=0.8 * fq, num_steps=1, x_offset=oq[0], y_offset=oq[1], **ap),
            ele("sextupole", l=0.2, k2=7.5 * fs, num_steps=4, x_offset=os_[0], y_offset=os_[1], **ap),
            ele("multipole", knl={3: k3l[c]}, **ap),
        ]
        if c in rf_cells:
            cell.append(ele("rfcavity", l=1.0, voltage=1e6, rf_frequency=500e6, phi0=0.0, num_steps=10, **ap))
        else:
            cell.append(ele("drift", l=1.0, **ap))
        fq, oq = qerr()
        fs, os_ = serr()
        cell += [
            ele("sbend", l=2.0, angle=2 * math.pi / n_cells, fringe_type=abi.FRINGE_HARD_EDGE_ONLY, num_steps=1, **ap),
            ele("drift", l=1.0, **ap),
            ele("quadrupole", l=0.5, k1=-0.8 * fq, num_steps=1, x_offset=oq[0], y_offset=oq[1], **ap),
            ele("sextupole", l=0.2, k2=-7.5 * fs, num_steps=4, x_offset=os_[0], y_offset=os_[1], **ap),
        ]
        eles += cell
    set_reference(eles, species, e_tot=e_tot)
    return set_rf_harmonic(eles, 500e6)


def revolution_time(eles):
    """Reference time of one pass through the element list [s]."""
    return eles[-1].ref_time - (eles[0].ref_time - eles[0].delta_ref_time)


def set_rf_harmonic(eles, f_target):
    """Set every rfcavity's frequency to the harmonic h / T of the
    revolution time T nearest f_target (h = nint(f_target T))."""
    t_rev = revolution_time(eles)
    h = max(1, _nint(f_target * t_rev))
    for e in eles:
        if e.key == abi.KEY_RFCAVITY:
            e.rf_frequency = h / t_rev
    return eles


def da_variants(n_var=256, seed=20261204, min_cells=63, max_cells=500, n_rf=8, offsets_scale=1.0):
    """Error-seeded C1-style variants of uneven length (63..500 cells = 504..4000
    elements): quad k1 x (1 + 1e-3 g), sextupole k2 x (1 + 5e-3 g), x/y offsets
    offsets_scale x 50 um g on quadrupoles and sextupoles.  Returns (list of
    element lists)."""
    rng = np.random.default_rng(seed)
    out = []
    for v in range(n_var):
        nc = int(rng.integers(min_cells, max_cells + 1))
        nq = ns = 2 * nc
        k1e = 1e-3 * rng.standard_normal(nq)
        k2e = 5e-3 * rng.standard_normal(ns)
        offs = {}
        for i in range(nq):
            offs[("q", i)] = tuple(offsets_scale * 50e-6 * rng.standard_normal(2))
        for i in range(ns):
            offs[("s", i)] = tuple(offsets_scale * 50e-6 * rng.standard_normal(2))
        out.append(ring_c1(n_cells=nc, n_rf=n_rf, seed=seed + 1000 + v, k1_err=k1e, k2_err=k2e, offsets=offs))
    return out


def EleRec_copy(e):
    """Independent copy of an element record."""
    return abi.EleRec.from_buffer_copy(e)
