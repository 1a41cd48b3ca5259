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


# ---------------------------------------------------------------------------
# Branch switches: each side of every switch against the mathematics
# ---------------------------------------------------------------------------

def _special():
    import ctypes as C
    L = oracle.lib()
    L.oracle_special.restype = C.c_double
    L.oracle_special.argtypes = [C.c_int, C.c_double]
    L.oracle_low_energy_z.restype = C.c_double
    L.oracle_low_energy_z.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    return L


@pytest.mark.parametrize("which,cut,tol", [(0, 1e-8, 1e-16), (1, 1e-8, 1e-16), (2, 0.1, 1e-13)])
def test_special_function_cutoffs(which, cut, tol):
    """sinc (sinc.f90:15-48), cosc (cosc.f90) and sincc (sincc.f90) on both
    sides of their series cut-offs against sin x / x, (1 - cos x) / x^2 and
    (x - sin x) / x^3 in 40-digit arithmetic: a wrong series term would show
    on the series side, a wrong cut-off as a jump."""
    import mpmath as mp
    mp.mp.dps = 40
    exact = {0: lambda x: mp.sin(x) / x, 1: lambda x: (1 - mp.cos(x)) / x ** 2,
             2: lambda x: (x - mp.sin(x)) / x ** 3}[which]
    L = _special()
    for f in (0.5, 0.99, 0.9999999, 1.0000001, 1.01, 2.0):
        x = cut * f
        e = exact(mp.mpf(x))
        assert abs(float((L.oracle_special(which, x) - e) / e)) <= tol, (which, x)


def test_low_energy_z_correction_switch():
    """low_energy_z_correction.f90:44: the series branch below
    m (beta0 pz)^2 = 3e-7 E and the exact L (beta - beta0) / beta0 above it,
    for protons at 1 GeV (switch at pz = 1.63e-3), against 40-digit
    arithmetic: the series side is within its own truncation (1e-10 of the
    correction at the switch; a wrong coefficient would be ~1e-4), the exact
    side within its cancellation (1e-12)."""
    import mpmath as mp
    mp.mp.dps = 40
    L = _special()
    m, E = abi.M_PROTON, 1e9
    p0c = math.sqrt(E * E - m * m)
    b0 = p0c / E
    thr = math.sqrt(3e-7 * E / m) / b0
    for f, tol in ((0.5, 2e-11), (0.99, 1e-10), (0.9999, 1e-10), (1.0001, 1e-12), (1.01, 1e-12), (2.0, 1e-12)):
        pz = thr * f
        got = L.oracle_low_energy_z(abi.PROTON, p0c, E, pz, 1.0)
        pc = mp.mpf(p0c) * (1 + mp.mpf(pz))
        beta = pc / mp.sqrt(pc ** 2 + mp.mpf(m) ** 2)
        e = (beta - mp.mpf(b0)) / mp.mpf(b0)
        assert abs(float((got - e) / e)) <= tol, (f, float((got - e) / e))


def test_sbend_near_zero_orbit_switch():
    """track_a_bend.f90:136-137: orbits with |x g|, |px|, |py|, |pz| < 1e-9
    take the semi-linear map, larger ones the exact body.  At these
    amplitudes the map is linear to 1e-18, so output / amplitude is the same
    on both sides of the switch: transverse coordinates to the exact body's
    rounding at 1e-9 amplitudes (1e-8 relative), z, a difference of path
    lengths of order L = 1 m, to 1e-15 m (a wrong semi-linear term would
    differ at O(1))."""
    b = lattice.set_reference([lattice.ele("sbend", l=1.0, g=0.01, fringe_type=abi.FRINGE_NONE, num_steps=1)],
                              abi.ELECTRON, e_tot=1e9)
    v = np.array([0.3e-9, 1e-9, 0.5e-9, 0.7e-9, 0.2e-9, 0.8e-9])
    names = ("x", "px", "y", "py", "z", "pz")
    outs = {}
    for s in (0.5, 0.999, 1.001, 1.1):
        r = oracle.track_bunch(b, abi.ELECTRON, 1, (v * s)[:, None])
        outs[s] = {k: r[k][0] for k in names}
    for s in (0.5, 1.001, 1.1):
        for k in names:
            a, ref = outs[s][k] / s, outs[0.999][k] / 0.999
            if k == "z":
                assert abs(outs[s][k] - s * ref) <= 1e-15, (s, k)
            else:
                assert abs(a - ref) <= 1e-8 * abs(ref), (s, k)


def test_sbend_k1_map_small_argument_switch():
    """sbend_body_with_k1_map (track_a_bend.f90:546-559): below arg =
    sqrt|k1 + g g_tot| L / sqrt(1+pz) = 1e-6 the x plane takes s_x, c_x to
    second order in arg, above it sin / cos (sinh / cosh).  On-momentum
    particles (x_c = 0; off momentum x_c = (g(1+pz) - g_tot) / k_x diverges
    as k_x -> 0) give x, px, y, py continuous across the switch (the change
    from one side to the other is the local slope's, to 1e-15).  z is not:
    the small-argument branch sets z2 = +g L^2 / (2 (1+pz)) (:549) where both
    exact branches tend to -g L^2 / (2 (1+pz)), the sign of the path-length
    change of a particle with px > 0.  The oracle (and the device after it)
    follow the reference: the test pins the size of that jump, g L^2 px, and
    that z is flat on each side."""
    g, L = 0.01, 1.0
    c = np.array([[1e-3, -5e-4], [2e-4, 1e-4], [-1e-3, 2e-3], [1e-4, -3e-4], [1e-3, 0.0], [0.0, 0.0]])
    res = {}
    for f in (0.997, 0.999, 1.001, 1.003):
        e = lattice.set_reference([lattice.ele("sbend", l=L, g=g, k1=-g * g + 1e-12 * f,
                                               fringe_type=abi.FRINGE_NONE, num_steps=1)], abi.ELECTRON, e_tot=1e9)
        res[f] = oracle.track_bunch(e, abi.ELECTRON, 1, c)
    for k in ("x", "px", "y", "py", "pz"):
        pred = res[0.999][k] + (res[0.999][k] - res[0.997][k])  # the small-argument side, extrapolated
        post = res[1.001][k] - (res[1.003][k] - res[1.001][k])  # the exact side, extrapolated back
        scale = np.max(np.abs(res[0.999][k])) + 1e-300
        assert np.max(np.abs(pred - res[1.001][k])) <= 1e-15 * scale, k
        assert np.max(np.abs(post - res[0.999][k])) <= 1e-15 * scale, k
    assert np.max(np.abs(res[0.997]["z"] - res[0.999]["z"])) <= 1e-15
    # 1 - cos(arg) cancels at arg = 1e-6: z2 px is good to |z2 px| eps / (1 - c_x), 2e-10 m here
    assert np.max(np.abs(res[1.001]["z"] - res[1.003]["z"])) <= 1e-9
    jump = res[0.999]["z"] - res[1.001]["z"]  # to the exact side's eps / (1 - c_x) = 2.2e-4 relative
    assert np.allclose(jump, g * L * L * c[1] / (1 + c[5]), rtol=1e-3, atol=0), jump


# ---------------------------------------------------------------------------
# Linear bend edge and rf coupler
# ---------------------------------------------------------------------------

def test_linear_bend_edge_kick_closed_form():
    """linear_bend_edge_kick (fringe_mod.f90:151-212) is the thin linear
    edge map of the MAD physics guide: px += h tan(e) x, py -= h tan(e - psi) y,
    psi = 2 fint hgap h (1 + sin^2 e) / cos e, with e1/fint/hgap at the
    entrance and e2/fintx/hgapx at the exit.  So the exact one-element
    Jacobian (dual numbers) at the reference orbit of a bend with linear
    edges is E2 J_body E1 with J_body the same bend without edges.  With no
    fringe integral its linear part also equals that of the Hwang edge map,
    which fringe_test pins."""
    h = 0.03
    kw = dict(l=0.6, g=h, e1=0.05, e2=-0.04, fint=0.5, hgap=0.02, fintx=0.3, hgapx=0.03, num_steps=1)
    lin = lattice.set_reference([lattice.ele("sbend", fringe_type=abi.FRINGE_LINEAR_EDGE, **kw)], abi.ELECTRON,
                                e_tot=1e9)
    none = lattice.set_reference([lattice.ele("sbend", fringe_type=abi.FRINGE_NONE, **kw)], abi.ELECTRON, e_tot=1e9)
    zero = np.zeros(6)
    _, J_lin = oracle.jacobian(lin, abi.ELECTRON, 1, zero)
    _, J_none = oracle.jacobian(none, abi.ELECTRON, 1, zero)

    def edge(e, fint, hgap):
        psi = 2 * fint * hgap * h * (1 + math.sin(e) ** 2) / math.cos(e)
        E = np.eye(6)
        E[1, 0] = h * math.tan(e)
        E[3, 2] = -h * math.tan(e - psi)
        return E

    J_ref = edge(kw["e2"], kw["fintx"], kw["hgapx"]) @ J_none @ edge(kw["e1"], kw["fint"], kw["hgap"])
    assert np.max(np.abs(J_lin - J_ref)) <= 1e-14 * np.max(np.abs(J_ref))
    kw0 = dict(kw, fint=0.0, fintx=0.0)
    lin0 = lattice.set_reference([lattice.ele("sbend", fringe_type=abi.FRINGE_LINEAR_EDGE, **kw0)], abi.ELECTRON,
                                 e_tot=1e9)
    hw0 = lattice.set_reference([lattice.ele("sbend", fringe_type=abi.FRINGE_BASIC_BEND, **kw0)], abi.ELECTRON,
                                e_tot=1e9)
    _, J_l0 = oracle.jacobian(lin0, abi.ELECTRON, 1, zero)
    _, J_h0 = oracle.jacobian(hw0, abi.ELECTRON, 1, zero)
    assert np.max(np.abs(J_l0 - J_h0)) <= 1e-14 * np.max(np.abs(J_h0))


def test_rf_coupler_kick_angle_symmetry():
    """rf_coupler_kick (rf_coupler_kick.f90:41-76) kicks along the direction
    2 pi coupler_angle in the x-y plane and the cavity body is symmetric
    under rotations about the axis: turning coupler_angle by a quarter turn
    maps (x, px, y, py) to (-y, -py, x, px) after the cavity for particles
    that start on the axis (dp_x = dp cos 2 pi a, dp_y = dp sin 2 pi a),
    and leaves z, pz unchanged."""
    def cav(angle):
        return lattice.set_reference([lattice.ele("rfcavity", l=1.0, voltage=1e6, rf_frequency=500e6, phi0=0.1,
                                                  num_steps=5, coupler_strength=0.2, coupler_angle=angle,
                                                  coupler_phase=0.1, coupler_at=abi.BOTH_ENDS)],
                                     abi.ELECTRON, e_tot=1e9)
    rng = np.random.default_rng(9)
    c = np.zeros((6, 64))
    c[4] = rng.uniform(-3e-3, 3e-3, 64)
    c[5] = rng.uniform(-2e-3, 2e-3, 64)
    a = oracle.track_bunch(cav(0.1), abi.ELECTRON, 1, c)
    b = oracle.track_bunch(cav(0.35), abi.ELECTRON, 1, c)
    assert np.max(np.abs(a["px"])) > 1e-6  # the coupler kicks
    for ka, kb, sgn in (("x", "y", 1), ("px", "py", 1), ("y", "x", -1), ("py", "px", -1)):
        scale = np.max(np.abs(a[ka])) + np.max(np.abs(a[kb]))
        assert np.max(np.abs(sgn * b[kb] - a[ka])) <= 1e-14 * scale, ka
    for k in ("z", "pz"):
        assert np.max(np.abs(b[k] - a[k])) <= 1e-14 * np.max(np.abs(a[k])), k
