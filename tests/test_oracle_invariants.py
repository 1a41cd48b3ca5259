"""The oracle against closed forms and invariants that the mathematics fixes,
chosen so that a dropped term, a wrong sign or a transposed operand fails:

* exact drift (straight line, z from the velocity ratio);
* linear quadrupole body = matrix exponential of the Hill equations;
* exact sector-bend body = circle / exit-plane geometry;
* thin multipole kicks = the textbook complex polynomials;
* one-turn maps symplectic to 1e-12 (dual-number Jacobian);
* no energy change without rf; RF energy gain = V sin(phase) for a thin cavity.
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from bmx import abi, beam, lattice
import oracle

E_TOT = 1e9
M = abi.M_ELECTRON
P0C = math.sqrt(E_TOT ** 2 - M ** 2)
BETA0 = P0C / E_TOT
RNG = np.random.default_rng(1234)


def _beta(pz):
    pc = P0C * (1 + pz)
    return pc / math.sqrt(pc * pc + M * M)


def _track(e, vec):
    lattice.set_reference([e], abi.ELECTRON, e_tot=E_TOT)
    out, t, b, st = oracle.track1(e, abi.ELECTRON, abi.ELECTRON, 1, vec)
    assert st == abi.ALIVE
    return np.array(out)


def _samples(n, s):
    return [RNG.standard_normal(6) * np.array(s) for _ in range(n)]


@pytest.mark.parametrize("vec", _samples(8, [1e-3, 5e-3, 1e-3, 5e-3, 1e-3, 1e-2]))
def test_drift_closed_form(vec):
    L = 2.7
    out = _track(lattice.ele("drift", l=L), vec)
    x, px, y, py, z, pz = [np.longdouble(v) for v in vec]
    ps = np.sqrt((1 + pz) ** 2 - px ** 2 - py ** 2)
    bratio = np.longdouble(_beta(float(pz))) / np.longdouble(BETA0)
    exp = [x + L * px / ps, px, y + L * py / ps, py, z + L * (bratio - (1 + pz) / ps), pz]
    for o, e in zip(out, exp):
        assert abs(o - float(e)) <= 2e-15 * max(1.0, abs(float(e))) + 1e-17


@pytest.mark.parametrize("k1", [0.8, -0.8, 2.5])
def test_quadrupole_linear_part(k1):
    L = 0.5
    for vec in _samples(4, [1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-2]):
        out = _track(lattice.ele("quadrupole", l=L, k1=k1, num_steps=1), vec)
        rp = 1 + vec[5]
        Ax = np.array([[0, 1 / rp], [-k1, 0]])
        Ay = np.array([[0, 1 / rp], [k1, 0]])
        ex = expm(Ax * L) @ vec[0:2]
        ey = expm(Ay * L) @ vec[2:4]
        assert np.allclose(out[0:2], ex, rtol=1e-13, atol=1e-17)
        assert np.allclose(out[2:4], ey, rtol=1e-13, atol=1e-17)
        assert out[5] == vec[5]


@pytest.mark.parametrize("g,pz,x,px", [(0.05, 0.0, 1e-3, 2e-3), (0.05, 0.01, -2e-3, -1e-3), (0.3, -0.02, 5e-3, 1e-2)])
def test_sbend_body_geometry(g, pz, x, px):
    """Exit coordinates of a sector bend (no edges) from circle geometry: the
    particle circle of radius (1+pz)/g intersected with the exit plane at angle gL."""
    L = 1.2
    e = lattice.ele("sbend", l=L, g=g, fringe_type=abi.FRINGE_NONE, num_steps=1)
    out = _track(e, [x, px, 0.0, 0.0, 0.0, pz])
    rho = 1 / g
    rp = 1 + pz
    R = rp * rho
    # frame: bend centre at origin, entrance plane along +X, beam moving towards +Y at entrance
    P0 = np.array([rho + x, 0.0])
    phi = math.asin(px / rp)  # angle of motion from the reference direction (+Y), positive outward (+X)
    d0 = np.array([math.sin(phi), math.cos(phi)])
    # the bend turns the particle towards the origin: its circle's centre is R along -(d0 rotated by -90 deg)
    C = P0 - R * np.array([d0[1], -d0[0]])
    th = g * L
    u = np.array([math.cos(th), math.sin(th)])  # exit plane direction from origin
    # intersection of |P - C| = R with P = s u, s > 0
    b = -2 * np.dot(u, C)
    cc = np.dot(C, C) - R * R
    s = (-b + math.sqrt(b * b - 4 * cc)) / 2
    P1 = s * u
    x_out = s - rho
    # direction at exit: tangent of the particle circle, clockwise motion
    r1 = (P1 - C) / R
    d1 = np.array([r1[1], -r1[0]])
    if np.dot(d1, np.array([-math.sin(th), math.cos(th)])) < 0:
        d1 = -d1
    sin_out = np.dot(d1, u)  # component along the outward radial direction
    px_out = rp * sin_out
    r0 = (P0 - C) / R
    ang = math.acos(max(-1.0, min(1.0, float(np.dot(r0, r1)))))
    path = R * ang
    z_out = _beta(pz) / BETA0 * L - path
    assert abs(out[0] - x_out) < 1e-13, (out[0], x_out)
    assert abs(out[1] - px_out) < 1e-13, (out[1], px_out)
    assert abs(out[4] - z_out) < 1e-13, (out[4], z_out)


@pytest.mark.parametrize("n,knl", [(1, 0.7), (2, 1.5), (3, -40.0), (4, 300.0)])
def test_thin_multipole_polynomial(n, knl):
    """Normal KnL kick: dpx + i dpy = -KnL/n! (x - i y)^n conjugated -> textbook
    dpx = -KnL/n! Re[(x + i y)^n], dpy = KnL/n! Im[(x + i y)^n]."""
    x, y = 3e-3, -2e-3
    out = _track(lattice.ele("multipole", knl={n: knl}), [x, 1e-4, y, -2e-4, 0.0, 0.0])
    zz = complex(x, y) ** n
    f = knl / math.factorial(n)
    assert abs(out[1] - (1e-4 - f * zz.real)) < 1e-15
    assert abs(out[3] - (-2e-4 + f * zz.imag)) < 1e-15


def test_skew_multipole_is_rotated_normal():
    """A multipole with tilt T = pi/(2(n+1)) is the skew component: the kick equals
    the normal kick evaluated in coordinates rotated by T and rotated back."""
    n, knl, T = 2, 1.1, math.pi / 6
    x, y = 2e-3, 1e-3
    out = _track(lattice.ele("multipole", knl={n: knl}, tn={n: T}), [x, 0, y, 0, 0, 0])
    xr, yr = x * math.cos(T) + y * math.sin(T), -x * math.sin(T) + y * math.cos(T)
    zz = complex(xr, yr) ** n
    f = knl / math.factorial(n)
    kx, ky = -f * zz.real, f * zz.imag
    ex = kx * math.cos(T) - ky * math.sin(T)
    ey = kx * math.sin(T) + ky * math.cos(T)
    assert abs(out[1] - ex) < 1e-16 and abs(out[3] - ey) < 1e-16


@pytest.mark.parametrize("ring", ["c0", "c1"])
def test_one_turn_map_symplectic(ring):
    if ring == "c0":
        eles = lattice.fodo_ring_c0()
        coords, _ = beam.c0_bunch(eles[:7], n=4, seed=9)
    else:
        eles = lattice.ring_c1()
        tw = beam.periodic_twiss(lattice.ring_c1(n_rf=0)[:8])
        coords = beam.matched_gaussian(4, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed=9)
    for i in range(coords.shape[1]):
        out, J = oracle.jacobian(eles, abi.ELECTRON, 1, coords[:, i])
        assert oracle.symp_error(J) <= 1e-12
        r = oracle.track_bunch(eles, abi.ELECTRON, 1, coords[:, i:i + 1])
        assert np.array_equal(out, np.array([r[k][0] for k in ("x", "px", "y", "py", "z", "pz")]))


def test_no_energy_change_without_rf():
    eles = lattice.fodo_ring_c0(n_cells=10)
    coords, _ = beam.c0_bunch(eles[:7], n=64, seed=2)
    r = oracle.track_bunch(eles, abi.ELECTRON, 3, coords)
    assert np.array_equal(r["pz"], coords[5])


def test_thin_cavity_energy_gain():
    """Zero-length cavity: two half kicks at one phase, dE = V sin(2 pi (phi0 + z f / (beta c)))."""
    V, f, phi0 = 1e6, 500e6, 0.05
    e = lattice.ele("rfcavity", l=0.0, voltage=V, rf_frequency=f, phi0=phi0)
    z, pz = 2e-3, 1e-3
    out = _track(e, [0, 0, 0, 0, z, pz])
    b = _beta(pz)
    E0 = P0C * (1 + pz) / b
    dE = V * math.sin(2 * math.pi * (phi0 + z * f / (b * abi.C_LIGHT)))
    E1 = E0 + dE
    pz1 = math.sqrt(E1 * E1 - M * M) / P0C - 1
    assert abs(out[5] - pz1) < 1e-14


@pytest.mark.parametrize("ft", [abi.FRINGE_FULL, abi.FRINGE_SAD_FULL, abi.FRINGE_SOFT_EDGE_ONLY])
def test_oracle_refuses_unmodelled_bend_edges(ft):
    """exact_bend_edge_kick and sad_soft_bend_edge_kick are not restated:
    the oracle refuses such a bend instead of tracking past its edges; with
    fringe_at = no_end nothing is missing and it tracks."""
    e = lattice.set_reference([lattice.ele("sbend", l=1.0, angle=0.02, num_steps=1, fringe_type=ft)], abi.ELECTRON,
                              e_tot=1e9)
    with pytest.raises(RuntimeError):
        oracle.track_bunch(e, abi.ELECTRON, 1, np.zeros((6, 4)))
    e[0].fringe_at = abi.NO_END
    r = oracle.track_bunch(e, abi.ELECTRON, 1, np.zeros((6, 4)))
    assert np.all(r["state"] == abi.ALIVE)


@pytest.mark.parametrize("kind", ["hwang", "hard_edge", "linear_edge", "dg", "k2", "k1"])
def test_reversed_sbend_is_the_mirrored_bend_for_the_antiparticle(kind):
    """A sector bend turned around (orientation -1) and crossed by the
    antiparticle meets the physical exit face first and the same field
    relative to its charge: it equals the bend with its two faces swapped
    (e1 <-> e2, fint <-> fintx, hgap <-> hgapx) crossed by the particle.
    Exact for the edges (Hwang, hard edge, linear), dg and multipoles; for the
    k1 body every coordinate but z is exact, and z differs by the constant
    path-length term the reference signs with orientation x direction
    (track_a_bend.f90:655), a function of pz alone."""
    rng = np.random.default_rng(3)
    c = rng.standard_normal((6, 2000)) * np.array([2e-3, 2e-4, 2e-3, 2e-4, 2e-3, 2e-3])[:, None]
    base = dict(l=1.2, angle=0.05, e1=0.03, e2=-0.02, fint=0.6, hgap=0.02, fintx=0.4, hgapx=0.03, num_steps=1)
    extra = {"hwang": {}, "hard_edge": dict(fringe_type=abi.FRINGE_HARD_EDGE_ONLY),
             "linear_edge": dict(fringe_type=abi.FRINGE_LINEAR_EDGE), "dg": dict(dg=0.003),
             "k2": dict(k2=2.0, num_steps=4), "k1": dict(k1=0.3, num_steps=4)}[kind]
    kw = dict(base, **extra)
    mirrored = dict(kw, e1=kw["e2"], e2=kw["e1"], fint=kw["fintx"], fintx=kw["fint"], hgap=kw["hgapx"],
                    hgapx=kw["hgap"])
    rev = lattice.set_reference([lattice.ele("sbend", orientation=-1, **kw)], abi.ELECTRON, e_tot=1e9)
    mir = lattice.set_reference([lattice.ele("sbend", **mirrored)], abi.ELECTRON, e_tot=1e9)
    a = oracle.track_bunch(rev, abi.ELECTRON, 1, c, tracked_species=abi.POSITRON)
    b = oracle.track_bunch(mir, abi.ELECTRON, 1, c)
    assert np.array_equal(a["state"], b["state"])
    for k in ("x", "px", "y", "py", "pz") + (("z",) if kind != "k1" else ()):
        assert np.max(np.abs(a[k] - b[k])) <= 1e-15 * np.max(np.abs(b[k])), k
    if kind == "k1":
        dz = a["z"] - b["z"]
        assert np.max(np.abs(dz)) > 1e-8  # the orientation-signed term is there
        # the same pz with other transverse coordinates: the same z difference
        c2 = c.copy()
        c2[:5] = rng.standard_normal((5, c.shape[1])) * np.array([2e-3, 2e-4, 2e-3, 2e-4, 2e-3])[:, None]
        a2 = oracle.track_bunch(rev, abi.ELECTRON, 1, c2, tracked_species=abi.POSITRON)
        b2 = oracle.track_bunch(mir, abi.ELECTRON, 1, c2)
        dz2 = a2["z"] - b2["z"]
        assert np.max(np.abs(dz2 - dz)) <= 1e-12 * np.max(np.abs(dz))
