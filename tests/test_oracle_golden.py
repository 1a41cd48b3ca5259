"""The CPU oracle against the reference's own golden vectors:
regression_tests/tracking_method_test (Bmad_Standard rows, printed to 11
significant digits), fringe_test (Hwang edge, printed to 7) and aperture_test
(loss states and unstable factors)."""
import json
import os

import numpy as np
import pytest

from bmx import abi
import oracle
from helpers import SPECIES, ele_from_spec, printed_match

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


TM = _load("tracking_method_bmad_standard.json")


@pytest.mark.parametrize("row", TM["rows"], ids=[r["name"] + r["variant"] for r in TM["rows"]])
def test_tracking_method_row(row):
    """tracking_method_test.f90:158-199 harness: init_coord at the element's
    start edge, species flipped to the antiparticle when direction x
    orientation = -1, absolute time for backward tracking and -Abs_Time rows."""
    spec = TM["elements"][row["name"]]
    ref_species = SPECIES[TM["particle"]]
    e = ele_from_spec(spec, ref_species, e_tot=TM["e_tot"])
    direction, osign, abs_time = TM["variants"][row["variant"]]
    if spec["key"] not in ("sbend", "rfcavity"):
        e.orientation = osign
    species = ref_species if direction * e.orientation == 1 else abi.POSITRON + abi.ELECTRON - ref_species
    oracle.set_com(1, 1 if (abs_time or direction == -1) else 0)
    try:
        vec, t, beta, state = oracle.track1(e, ref_species, species, direction, TM["start"])
    finally:
        oracle.set_com(1, 0)
    assert state == abi.ALIVE
    bad = [(i, v, g) for i, (v, g) in enumerate(zip(vec, row["vec"])) if not printed_match(v, g, TM["printed_digits"])]
    assert not bad, f"coordinates off the printed golden value: {bad}"


def test_tracking_method_row_count():
    # every element the device path supports appears in all four orientation/direction variants
    # 24 elements x the 4 orientation/direction variants + the -Abs_Time rows of the 3 rf cavities
    assert len(TM["rows"]) == 99


FR = _load("fringe_test.json")


@pytest.mark.parametrize("edge,key", [(True, "Up-Orb"), (False, "Dn-Orb")])
def test_fringe_hwang_orbit(edge, key):
    """fringe_test.f90: hwang_bend_edge_kick orbit change, electron 1 MeV."""
    sp = SPECIES[FR["particle"]]
    e = ele_from_spec(FR["element"], sp, e_tot=FR["e_tot"])
    out = oracle.hwang_edge(e, sp, edge, FR["start"], sp)
    d = np.array(out) - np.array(FR["start"])
    gold = FR["rows"][key]["vals"]
    for v, g in zip(d, gold):
        assert printed_match(v, g, 7), (list(d), gold)


@pytest.mark.parametrize("edge,prefix,symp_key", [(True, "Up", "Up-Symp-Err"), (False, "Dn", "Dn-Symp-Err")])
def test_fringe_hwang_jacobian(edge, prefix, symp_key):
    """The Jacobian of the Hwang edge map by central differences of the
    oracle against the Up-/Dn-Mat rows, and its symplectic error against the
    pinned mat_symp_error (sim_utils/matrix/mat_symp_error.f90:31-71)."""
    sp = SPECIES[FR["particle"]]
    e = ele_from_spec(FR["element"], sp, e_tot=FR["e_tot"])
    x0 = np.array(FR["start"])
    J = np.zeros((6, 6))
    for j in range(6):
        h = 1e-6 * max(abs(x0[j]), 1e-2)
        xp, xm = x0.copy(), x0.copy()
        xp[j] += h
        xm[j] -= h
        J[:, j] = (np.array(oracle.hwang_edge(e, sp, edge, xp, sp)) - np.array(oracle.hwang_edge(e, sp, edge, xm, sp))) / (2 * h)
    for i in range(6):
        gold = FR["rows"][f"{prefix}-Mat{i + 1}"]["vals"]
        row = J[i] - np.eye(6)[i]
        # Rows x, px, y, py are derivatives of the map.  Bmad fills the z row of
        # its edge matrix by symplectic completion (fringe_mod.f90:405-410), not by
        # differentiating dz, so that row is left to the symplectic-error check.
        if i >= 4:
            continue
        for j in range(6):
            assert abs(row[j] - gold[j]) <= max(1e-10, 6e-7 * abs(gold[j])), (i, j, row[j], gold[j])
    S = np.zeros((6, 6))
    for k in range(3):
        S[2 * k, 2 * k + 1] = 1
        S[2 * k + 1, 2 * k] = -1
    err = np.max(np.abs(J.T @ S @ J - S))
    gold = FR["rows"][symp_key]["vals"][0]
    assert abs(err - gold) <= 0.01 * gold


AP = _load("aperture_test.json")


def _f32(*v):
    """aperture_test.f90 assigns xw2, x0, yw2, y0, eps from default-real
    (single precision) literals, so their double values are the float32 ones."""
    return [float(np.float32(x)) for x in v]


def _aperture_grid(e, sp, kind, sign):
    """aperture_test.f90:86-180 3x3 grids just inside (-) / outside (+) the limits."""
    (eps,) = _f32(1e-5)
    states, ufs = [], []
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            if kind == "Pxy":
                xw2, x0, yw2, y0 = _f32(0.01, 0.01, 0.015, 0.0)
                d = (xw2 - eps) if sign == "-" else (xw2 + eps)
                dy = (yw2 - eps) if sign == "-" else (yw2 + eps)
                vec = [0.0, x0 + i * d, 0.0, y0 + j * dy, -0.6, 0.02]
            elif kind == "xy":
                xw2, x0, yw2, y0 = _f32(0.01, 0.01, 0.015, 0.0)
                d = (xw2 - eps) if sign == "-" else (xw2 + eps)
                dy = (yw2 - eps) if sign == "-" else (yw2 + eps)
                vec = [x0 + i * d, 0.1 if sign == "-" else 0.0, y0 + j * dy, 0.0, -0.6, 0.02]
            else:
                xw2, x0, yw2, y0 = _f32(0.2, -0.6, 0.03, 0.02)
                d = (xw2 - eps) if sign == "-" else (xw2 + eps)
                dy = (yw2 - eps) if sign == "-" else (yw2 + eps)
                vec = [0.0, 0.01, 0.0, 0.0, x0 + i * d, y0 + j * dy]
            st, uf = oracle.check_aperture(e, vec, sp)
            states.append(st)
            ufs.append(uf)
    return states, ufs


@pytest.mark.parametrize("ename", ["EC", "RC"])
@pytest.mark.parametrize("kind", ["Pxy", "xy", "zPz"])
@pytest.mark.parametrize("sign", ["-", "+"])
def test_aperture_states(ename, kind, sign):
    sp = SPECIES[AP["particle"]]
    e = ele_from_spec(AP["elements"][ename], sp, p0c=AP["p0c"])
    states, ufs = _aperture_grid(e, sp, kind, sign)
    gs = AP["rows"][f"{ename}_S_{kind}{sign}"]["vals"]
    gu = AP["rows"][f"{ename}_U_{kind}{sign}"]["vals"]
    assert states == [int(v) for v in gs]
    for u, g in zip(ufs, gu):
        assert printed_match(u, g, 3) or abs(u - g) <= 1e-8, (ufs, gu)


@pytest.mark.parametrize("row", [r for r in TM["rows"] if r["variant"] in ("-Anti_D", "-Anti_OD")],
                         ids=lambda r: r["name"] + r["variant"])
def test_backward_bunch_tracking_matches_the_golden_rows(row):
    """The oracle's backward multi-particle loop (track_bunch with direction
    = -1, the device's parity reference for backward tracking) reproduces the
    -Anti_D / -Anti_OD rows of tracking_method_test with absolute time, as
    the harness tracks them (tracking_method_test.f90:170-172)."""
    spec = TM["elements"][row["name"]]
    ref_species = SPECIES[TM["particle"]]
    e = ele_from_spec(spec, ref_species, e_tot=TM["e_tot"])
    if row["variant"] == "-Anti_OD" and spec["key"] not in ("sbend", "rfcavity"):
        e.orientation = -1
    species = ref_species if -1 * e.orientation == 1 else abi.POSITRON + abi.ELECTRON - ref_species
    coords = np.array(TM["start"], dtype=np.float64)[:, None]
    # init_coord at the exit end (init_coord.f90:222), absolute time as the harness sets it
    pc = e.p0c * (1.0 + coords[5])
    beta = pc / np.sqrt(pc * pc + abi.MASS[species] ** 2)
    t0 = e.ref_time - coords[4] / (beta * abi.C_LIGHT)
    out = oracle.track_bunch([e], ref_species, 1, coords, tracked_species=species, direction=-1, t=t0,
                             absolute_time=True)
    assert out["state"][0] == abi.ALIVE
    vec = [out[k][0] for k in ("x", "px", "y", "py", "z", "pz")]
    bad = [(i, v, g) for i, (v, g) in enumerate(zip(vec, row["vec"])) if not printed_match(v, g, TM["printed_digits"])]
    assert not bad, f"coordinates off the printed golden value: {bad}"


MP = _load("multipole_test.json")


@pytest.mark.parametrize("key", sorted(MP["rows"]))
def test_multipole_ele_to_ab_rows(key):
    """multipole_ele_to_ab against regression_tests/multipole_test (rows
    MULT-<use_ele_tilt><include_kicks>-QQ/BB/MM, printed to 5 digits, ABS 0
    there): a1, a3, a5, b2, b4, b6 of the magnetic poles (tilt rotation by
    (n+1) tilt, ref_tilt for the bend; k1 L, k2 L / 2 pulled in with the
    kicks; the bend's l (g + dg) scaling; the kt -> ab form of a multipole
    element) and the order of the electric poles."""
    row = MP["rows"][key]
    name = key.split("-")[0]
    sp = SPECIES[MP["particle"]]
    e = ele_from_spec(MP["elements"][name], sp, e_tot=MP["e_tot"])
    import ctypes as C
    a, b = (C.c_double * abi.N_POLE)(), (C.c_double * abi.N_POLE)()
    L = oracle.lib()
    L.oracle_multipole_ele_to_ab(C.byref(e), int(row["use_ele_tilt"]), 1, int(row["include_kicks"]), a, b, None)
    got = [a[1], a[3], a[5], b[2], b[4], b[6]]
    for v, g in zip(got, row["vals"]):
        assert printed_match(v, g, MP["printed_digits"]), (got, row["vals"])
    ix = L.oracle_multipole_ele_to_ab(C.byref(e), int(row["use_ele_tilt"]), 2, int(row["include_kicks"]), a, b,
                                      None)
    assert ix == row["ix_elec"]


M6 = _load("mat6_calc_method_test.json")
# The reference's matrix for a bend with electric multipoles approximates the
# energy change of the electric kick (bend_apply_multipole_kicks' make_matrix
# branch): its z and pz rows differ from the derivative of its own map by up
# to 2e-8 (dpz/dpy: -5.41e-7 printed against -5.6452e-7 for SBEND3 by central
# differences of the pinned map).  Those rows are checked against central
# differences of the map instead.
M6_APPROX_ROWS = {"SBEND3": (4, 5), "SBEND7": (4, 5), "RBEND6": (4, 5)}


def _m6_element(name):
    return ele_from_spec(M6["elements"][name], abi.POSITRON, e_tot=M6["e_tot"])


def m6_check(name, J, out, fd):
    """The golden rows of one element against a Jacobian J and orbit out at
    the start; fd(i) gives row i by central differences of the map."""
    rows = M6["rows"][name]
    x0 = np.array(M6["start"])
    skip = M6_APPROX_ROWS.get(name, ())
    for i in range(6):
        r = rows[f"MatrixRow{i + 1}"]
        want = np.array(r["vals"])
        if i in skip:
            got = fd(i)
            assert np.max(np.abs(J[i] - got)) <= 1e-9 * max(1.0, np.max(np.abs(got))), (name, i)
            assert np.max(np.abs(J[i] - want)) <= 1e-7, (name, i)  # the size of the reference's approximation
            continue
        assert np.max(np.abs(J[i] - want)) <= r["tol"], (name, i, J[i], want)
    if not skip:
        vec0 = out - J @ x0
        assert np.max(np.abs(vec0 - np.array(rows["Vector"]["vals"]))) <= rows["Vector"]["tol"], name
    gold = rows["Symp_Err"]["vals"][0]
    sym = oracle.symp_error(J)
    assert abs(sym - gold) <= max(1e-14, (1e-2 if skip else 1e-6) * gold), (name, sym, gold)


@pytest.mark.parametrize("name", sorted(M6["elements"]))
def test_mat6_bmad_standard_rows(name):
    """mat6_calc_method_test: the oracle's dual-number Jacobian at
    particle_start against every Bmad_Standard matrix row, vec0 and the
    symplectic error of the 23 elements of this path (positron, 1 MeV)."""
    e = _m6_element(name)
    x0 = np.array(M6["start"])
    out, J = oracle.jacobian([e], abi.POSITRON, 1, x0)

    def fd(i, h=1e-6):
        col = np.zeros(6)
        for j in range(6):
            vp, vm = x0.copy(), x0.copy()
            vp[j] += h
            vm[j] -= h
            a = oracle.track_bunch([e], abi.POSITRON, 1, vp.reshape(6, 1))
            b = oracle.track_bunch([e], abi.POSITRON, 1, vm.reshape(6, 1))
            k = ("x", "px", "y", "py", "z", "pz")[i]
            col[j] = (a[k][0] - b[k][0]) / (2 * h)
        return col

    m6_check(name, J, out, fd)


NL = _load("nonlin_small_ring.json")


def _small_ring():
    return [ele_from_spec(sp, abi.POSITRON, p0c=NL["p0c"]) for sp in NL["elements"]]


def check_coupled_twiss(T):
    """twiss_coupled of the small ring's one-turn matrix T against the
    coupled Twiss Bmad recorded for it (beginning[...] of small_ring.bmad)."""
    from bmx import beam
    tw = beam.twiss_coupled(T)
    g = {k: v["value"] for k, v in NL["twiss"].items()}
    got = dict(beta_a=tw["beta_a"], alpha_a=tw["alpha_a"], beta_b=tw["beta_b"], alpha_b=tw["alpha_b"],
               cmat_11=tw["cmat"][0, 0], cmat_12=tw["cmat"][0, 1], cmat_21=tw["cmat"][1, 0], cmat_22=tw["cmat"][1, 1],
               eta_x=tw["eta_x"], etap_x=tw["etap_x"], eta_y=tw["eta_y"], etap_y=tw["etap_y"])
    for k, v in got.items():
        assert abs(v - g[k]) <= 1e-12 * max(1.0, abs(g[k])), (k, v, g[k])
    return tw


def test_coupled_twiss_of_the_nonlin_test_ring():
    """nonlin_test/small_ring.bmad records the ring's coupled periodic Twiss
    (tilted quadrupole): the oracle's exact one-turn matrix, decoupled by
    twiss_coupled, gives beta, alpha of both modes, the coupling matrix and
    the lab dispersion to 1e-12 (the file prints 18 digits)."""
    _, T = oracle.jacobian(_small_ring(), abi.POSITRON, 1, np.zeros(6))
    tw = check_coupled_twiss(T)
    assert 0.9 < tw["gamma_c"] < 1.0


def test_coupled_twiss_reduces_to_uncoupled_and_rebuilds_the_map():
    """Invariants of the decomposition: without coupling it is twiss_from_mat2
    of each plane; for a coupled map V U conj(V) = T (V symplectic)."""
    from bmx import beam, lattice
    eles = lattice.fodo_ring_c0(n_cells=10)
    _, T = oracle.jacobian(eles, abi.ELECTRON, 1, np.zeros(6))
    tw = beam.twiss_coupled(T)
    bx, ax, _ = beam.twiss_from_mat2(T[0:2, 0:2])
    by, ay, _ = beam.twiss_from_mat2(T[2:4, 2:4])
    assert tw["gamma_c"] == 1.0 and np.all(tw["cmat"] == 0)
    assert np.allclose([tw["beta_a"], tw["alpha_a"], tw["beta_b"], tw["alpha_b"]], [bx, ax, by, ay], rtol=1e-14)
    _, T = oracle.jacobian(_small_ring(), abi.POSITRON, 1, np.zeros(6))
    tw = beam.twiss_coupled(T)
    g, Cm = tw["gamma_c"], tw["cmat"]
    V = np.block([[g * np.eye(2), Cm], [-beam._symp_conj2(Cm), g * np.eye(2)]])
    Vc = np.block([[g * np.eye(2), -Cm], [beam._symp_conj2(Cm), g * np.eye(2)]])
    assert np.allclose(V @ Vc, np.eye(4), atol=1e-14)
    U = Vc @ T[:4, :4] @ V
    assert np.max(np.abs(U[0:2, 2:4])) < 1e-14 and np.max(np.abs(U[2:4, 0:2])) < 1e-14


def test_chromatic_twiss_of_the_nonlin_test_ring():
    """dbeta/dpz and dalpha/dpz of both modes recorded in small_ring.bmad,
    by chrom_calc's central differences (delta 1e-4) of the coupled Twiss at
    the off-momentum closed orbits: to 1e-9 relative (the differences'
    rounding at delta 1e-4)."""
    from bmx import beam
    eles = _small_ring()
    res = beam.chromatic_twiss(lambda v: oracle.jacobian(eles, abi.POSITRON, 1, v))
    g = {k: v["value"] for k, v in NL["twiss"].items()}
    for k in ("dbeta_dpz_a", "dbeta_dpz_b", "dalpha_dpz_a", "dalpha_dpz_b"):
        assert abs(res[k] - g[k]) <= 1e-9 * abs(g[k]), (k, res[k], g[k])
