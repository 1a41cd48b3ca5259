"""Pins of the oracle's radiation code: track1_radiation
(bmad/modules/radiation_mod.f90:62-141) and radiation_map_setup with
rad1_damp_and_stoc_mats (radiation_mod.f90:160-232,
bmad/modules/rad_6d_mod.f90:406-718), restated in oracle/bmx_oracle.c.

The damping is pinned to the reference's own numbers by
regression_tests/closed_orbit_test: the 6-d closed orbit's z of two SLS-2
rings with radiation damping on sits where the cavities restore the energy
the maps take, and the test records it with the synchrotron tune at
ABS 5e-8.  No golden vector covers the fluctuations without spin and
Bmad's own random stream (regression_tests/long_term_tracking_test runs
ramping, spin and ran_gauss), so the maps are further pinned by what the
physics fixes:
the energy loss per turn U0/E = (2/3) r_e gamma^3 I2, Robinson's theorem
(the damping decrements of the three modes add up to 2 U0/E, i.e.
-ln det M = 4 U0/E for the one-turn matrix with damping), the partition
numbers J_y = 1 and J_x = 1 - I4/I2 of a flat ring, the equilibrium energy
spread C_q gamma^2 I3 / (J_z I2) and the vertical emittance of the photons'
opening angle (13/55) C_q / (J_y I2) * integral of beta_y |g|^3, with the
radiation integrals formed here from the ring's own optics.  The Gaussian
stream is Philox4x32-10, pinned by its published known-answer vectors.
"""
import json
import os
import sys

import numpy as np
import pytest
import scipy.linalg as sla

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "bmad-ecosystem_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

from bmx import abi, lattice  # noqa: E402
import oracle  # noqa: E402
from helpers import ring_from_fixture  # noqa: E402

R_E = 2.8179403227e-15
HBAR = 4.135667696e-15 / (2 * np.pi)
C_LIGHT = 2.99792458e8


def gamma_of(e):
    return np.sqrt(e.p0c ** 2 + abi.M_ELECTRON ** 2) / abi.M_ELECTRON


def c_q():
    return 55.0 / (32.0 * np.sqrt(3.0)) * HBAR * C_LIGHT / abi.M_ELECTRON


def no_apertures(eles):
    for e in eles:
        e.x1_limit = e.x2_limit = e.y1_limit = e.y2_limit = 0.0
    return eles


def one_turn(eles, maps):
    """One-turn matrix with damping (I + D1) M (I + D0) per element, the one
    without, and the turn's accumulated stochastic covariance."""
    sp = abi.ELECTRON
    I = np.eye(6)
    Md, M0, B = I.copy(), I.copy(), np.zeros((6, 6))
    Ms = []
    for i, e in enumerate(eles):
        M = oracle.make_mat6(e, sp, [0.0] * 6)[1]
        Ms.append(M)
        D0, D1 = maps[i, 0, 6:42].reshape(6, 6), maps[i, 1, 6:42].reshape(6, 6)
        S0, S1 = maps[i, 0, 48:].reshape(6, 6), maps[i, 1, 48:].reshape(6, 6)
        B = (I + D0) @ B @ (I + D0).T + S0 @ S0.T
        B = M @ B @ M.T
        B = (I + D1) @ B @ (I + D1).T + S1 @ S1.T
        Md = (I + D1) @ M @ (I + D0) @ Md
        M0 = M @ M0
    return Md, M0, B, Ms


def radiation_integrals(eles, M0, Ms, n_sub=32):
    """I2, I3, I4 and the opening-angle integral of beta_y |g|^3 over the
    sector bends (Simpson over n_sub slices), with the periodic dispersion
    and Twiss of the one-turn matrix M0 carried through the elements."""
    sp = abi.ELECTRON
    A, b = M0[:4, :4], M0[:4, 5]
    eta0 = np.linalg.solve(np.eye(4) - A, b)

    def twiss2(m):
        c = (m[0, 0] + m[1, 1]) / 2
        s = np.sign(m[0, 1]) * np.sqrt(1 - c * c)
        return m[0, 1] / s, (m[0, 0] - m[1, 1]) / (2 * s)

    by, ay = twiss2(M0[2:4, 2:4])
    Ty = np.array([[by, -ay], [-ay, (1 + ay * ay) / by]])
    vec = np.array([eta0[0], eta0[1], eta0[2], eta0[3], 0.0, 1.0])
    I2 = I3 = I4 = Iy = 0.0
    w = np.ones(n_sub + 1)
    w[1:-1:2], w[2:-1:2] = 4, 2
    for e, M in zip(eles, Ms):
        if e.key == abi.KEY_SBEND:
            hs = e.l / n_sub
            v4, vy = [], []
            for k in range(n_sub + 1):
                Msl = np.eye(6) if k == 0 else oracle.make_mat6(oracle.ele_slice(e, k * hs, 0.0, True, False), sp,
                                                                [0.0] * 6)[1]
                et = Msl @ vec
                ty = Msl[2:4, 2:4] @ Ty @ Msl[2:4, 2:4].T
                v4.append(et[0] * e.g ** 3)
                vy.append(ty[0, 0] * abs(e.g) ** 3)
            I4 += hs / 3 * np.dot(w, v4)
            Iy += hs / 3 * np.dot(w, vy)
            I2 += e.g ** 2 * e.l
            I3 += abs(e.g) ** 3 * e.l
        vec = M @ vec
        Ty = M[2:4, 2:4] @ Ty @ M[2:4, 2:4].T
    return I2, I3, I4, Iy


def mode_decrements(Md, u0e):
    """Damping decrement -ln|lambda| of each eigenvalue in units of U0/(2E)."""
    lam = np.linalg.eigvals(Md)
    return lam, -np.log(np.abs(lam)) / (u0e / 2)


# ---------------------------------------------------------------------------
# the Gaussian stream
# ---------------------------------------------------------------------------
def test_philox_known_answers():
    # Random123 known-answer vectors of philox4x32-10 (Salmon et al., SC'11)
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, out in kat:
        assert oracle.philox4x32_10(ctr, key) == out


def test_gaussian_deviates_are_standard_normal():
    g = np.array([oracle.rad_gauss6(pid, 3, pid % 2, 7, 0, 12345) for pid in range(20000)])
    assert g.shape == (20000, 6)
    # mean 0 and variance 1 per component to 4 standard errors; no correlation between components
    assert np.all(np.abs(g.mean(axis=0)) < 4 / np.sqrt(20000))
    assert np.all(np.abs(g.var(axis=0) - 1) < 4 * np.sqrt(2 / 20000))
    c = np.corrcoef(g.T)
    assert np.max(np.abs(c - np.eye(6))) < 4 / np.sqrt(20000)
    # every counter word changes the deviates
    base = oracle.rad_gauss6(1, 2, 0, 3, 4, 99)
    for args in ((2, 2, 0, 3, 4, 99), (1, 3, 0, 3, 4, 99), (1, 2, 1, 3, 4, 99), (1, 2, 0, 4, 4, 99),
                 (1, 2, 0, 3, 5, 99), (1, 2, 0, 3, 4, 98)):
        assert not np.allclose(oracle.rad_gauss6(*args), base)


# ---------------------------------------------------------------------------
# per-element maps
# ---------------------------------------------------------------------------
def test_bend_energy_loss_is_kd_g2_L():
    """On the design orbit a sector bend's two half maps take
    -(2/3) r_e gamma^3 g^2 L from pz (the energy loss U/E of the element)
    and leave the reference orbit as their reference."""
    eles = lattice.fodo_ring_c0(n_cells=4)
    b = [e for e in eles if e.key == abi.KEY_SBEND][0]
    kd = 2.0 / 3.0 * R_E * gamma_of(b) ** 3
    rm = oracle.rad_map_setup(b, abi.ELECTRON, [0.0] * 6, 28.0)
    assert rm[0, 47] + rm[1, 47] == pytest.approx(-kd * b.g ** 2 * b.l, rel=1e-12)
    assert np.all(rm[:, :6] == 0.0)
    for h in range(2):
        S = rm[h, 48:].reshape(6, 6)
        assert np.all(np.triu(S, 1) == 0.0) and np.all(np.diag(S) >= 0.0)  # Cholesky factor
    # the two halves mirror each other on the design orbit: the x and z
    # terms of the damping vector flip sign, pz and the y plane do not
    assert rm[0, 43] == pytest.approx(-rm[1, 43], rel=1e-9)
    assert rm[0, 46] == pytest.approx(-rm[1, 46], rel=1e-9)


def test_off_axis_quadrupole_energy_loss():
    """A quadrupole crossed off axis radiates: g = k1 x(s) (rel_p = 1), and
    its maps take kd * integral of (k1 x(s))^2 ds from pz, with x(s) = x0
    cos(sqrt(k1) s) + px0 sin(sqrt(k1) s) / sqrt(k1)."""
    q = lattice.fodo_ring_c0(n_cells=1)[0]
    k = np.sqrt(q.k1)
    x0, px0 = 1e-3, 2e-4
    rm = oracle.rad_map_setup(q, abi.ELECTRON, [x0, px0, 0.0, 0.0, 0.0, 0.0], 7.0)
    kd = 2.0 / 3.0 * R_E * gamma_of(q) ** 3
    s = np.linspace(0.0, q.l, 4001)
    x = x0 * np.cos(k * s) + px0 * np.sin(k * s) / k
    integral = np.trapezoid((q.k1 * x) ** 2, s)
    assert rm[0, 47] + rm[1, 47] == pytest.approx(-kd * integral, rel=2e-4)
    # on axis: no angle change, no map
    rm0 = oracle.rad_map_setup(q, abi.ELECTRON, [0.0] * 6, 7.0)
    assert np.all(rm0[:, 6:] == 0.0)


# ---------------------------------------------------------------------------
# ring-level pins
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c0_ring():
    eles = no_apertures(lattice.fodo_ring_c0(n_cells=100))
    maps = oracle.rad_maps(eles, abi.ELECTRON, np.zeros((len(eles), 6)))
    Md, M0, B, Ms = one_turn(eles, maps)
    return eles, maps, Md, M0, B, Ms


@pytest.fixture(scope="module")
def rf_ring():
    eles = no_apertures(lattice.ring_c1(n_cells=50, n_rf=8))
    maps = oracle.rad_maps(eles, abi.ELECTRON, np.zeros((len(eles), 6)))
    Md, M0, B, Ms = one_turn(eles, maps)
    return eles, maps, Md, M0, B, Ms


def test_ring_energy_loss_and_robinson(c0_ring, rf_ring):
    for (eles, maps, Md, M0, B, Ms), tol in ((c0_ring, 1e-7), (rf_ring, 5e-5)):
        kd = 2.0 / 3.0 * R_E * gamma_of(eles[0]) ** 3
        I2 = sum(e.g ** 2 * e.l for e in eles if e.key == abi.KEY_SBEND)
        u0e = kd * I2
        assert maps[:, :, 47].sum() == pytest.approx(-u0e, rel=1e-12)
        # Robinson: the decrements of all modes add up to 2 U0/E -> -ln det = 4 U0/E
        assert -np.log(np.linalg.det(Md)) == pytest.approx(4 * u0e, rel=tol)
        assert abs(np.linalg.det(M0) - 1.0) < 1e-12


def test_flat_ring_partition_numbers(c0_ring):
    """J_y = 1 and J_x = 1 - I4/I2 (sector bends with parallel faces, no
    gradient: I4 = integral of eta g^3); the energy mode of this ring with no
    rf is the real eigenvalue with decrement 2 J_z = 2 (2 + I4/I2)."""
    eles, maps, Md, M0, B, Ms = c0_ring
    kd = 2.0 / 3.0 * R_E * gamma_of(eles[0]) ** 3
    I2, I3, I4, Iy = radiation_integrals(eles, M0, Ms)
    u0e = kd * I2
    lam, dec = mode_decrements(Md, u0e)
    phase = np.abs(np.angle(lam))
    mu_x = np.arccos(np.trace(M0[0:2, 0:2]) / 2)
    mu_y = np.arccos(np.trace(M0[2:4, 2:4]) / 2)
    jx = dec[np.argmin(np.abs(phase - mu_x))]
    jy = dec[np.argmin(np.abs(phase - mu_y))]
    real = (np.abs(lam.imag) < 1e-12) & (dec > 1.0)
    assert jy == pytest.approx(1.0, abs=1e-6)
    assert jx == pytest.approx(1.0 - I4 / I2, abs=2e-4)
    assert abs(1.0 - I4 / I2 - 1.0) > 5e-3  # the test resolves the I4 term
    assert dec[real][0] == pytest.approx(2 * (2.0 + I4 / I2), abs=4e-4)


def test_equilibrium_energy_spread_and_opening_angle_emittance(rf_ring, c0_ring):
    """Equilibrium of the one-turn map with damping and the accumulated
    stochastic covariance (Sigma = M Sigma M^T + B): sigma_pz^2 =
    C_q gamma^2 I3 / (J_z I2) with the ring's measured J_z, and the vertical
    emittance of the opening angle (13/55) C_q / (J_y I2) int beta_y |g|^3."""
    for ring, check_pz in ((rf_ring, True), (c0_ring, False)):
        eles, maps, Md, M0, B, Ms = ring
        gam = gamma_of(eles[0])
        kd = 2.0 / 3.0 * R_E * gam ** 3
        I2, I3, I4, Iy = radiation_integrals(eles, M0, Ms)
        u0e = kd * I2
        lam, dec = mode_decrements(Md, u0e)
        mu_y = np.arccos(np.trace(M0[2:4, 2:4]) / 2)
        jy = dec[np.argmin(np.abs(np.abs(np.angle(lam)) - mu_y))]
        # the y plane is decoupled: its 2x2 block has its own equilibrium
        Sy = sla.solve_discrete_lyapunov(Md[2:4, 2:4], B[2:4, 2:4])
        eps_y = np.sqrt(np.linalg.det(Sy))
        assert eps_y == pytest.approx(13.0 / 55.0 * c_q() / (jy * I2) * Iy, rel=1e-4)
        if check_pz:
            Sig = sla.solve_discrete_lyapunov(Md, B)
            mu_s = min(np.abs(np.angle(lam))[np.abs(np.angle(lam)) > 1e-6],
                       key=lambda a: abs(a - np.arccos(np.trace(M0[4:6, 4:6]) / 2)))
            jz = dec[np.argmin(np.abs(np.abs(np.angle(lam)) - mu_s))]
            assert 1.9 < jz < 2.2
            assert np.sqrt(Sig[5, 5]) == pytest.approx(np.sqrt(c_q() * gam ** 2 * I3 / (jz * I2)), rel=1e-3)


def test_tracking_applies_the_maps_as_their_linear_composition(c0_ring):
    """One turn of oracle tracking with damping on, for particles near the
    design orbit, is the map the element maps compose to: (I + D1) M (I + D0)
    per element (radiation_zero_average, so the orbit stays at the particles'
    amplitudes, where the second-order terms of the tracking stay below 1e-12
    of them); and a particle on the design orbit loses the damping vectors'
    energy carried through the ring.  A flipped sign, or a kick applied in the
    element frame instead of the lab frame, fails it."""
    eles, maps, Md, M0, B, Ms = c0_ring
    I = np.eye(6)
    c = np.zeros(6)  # the turn's affine constant (the maps' reference orbits are zero here)
    for i, M in enumerate(Ms):
        D0, D1 = maps[i, 0, 6:42].reshape(6, 6), maps[i, 1, 6:42].reshape(6, 6)
        c = (I + D0) @ c + maps[i, 0, 42:48]
        c = M @ c
        c = (I + D1) @ c + maps[i, 1, 42:48]
    rng = np.random.default_rng(5)
    x0 = rng.normal(size=(6, 32)) * np.array([[1e-5], [1e-6], [1e-5], [1e-6], [1e-5], [1e-6]])

    def change(x):  # what radiation changes in one turn
        out = oracle.track_bunch_rad(eles, abi.ELECTRON, 1, x, maps, zero_average=True)
        plain = oracle.track_bunch(eles, abi.ELECTRON, 1, x)
        assert np.all(out["state"] == abi.ALIVE)
        return np.array([out[n] - plain[n] for n in ("x", "px", "y", "py", "z", "pz")])

    # its odd part in the amplitudes against the composed matrices: the
    # ring's second-order terms drop out, and at these amplitudes the change
    # stands well above the rounding of 1 + pz (2e-16) that blurs it nearer
    # the axis
    odd = (change(x0) - change(-x0)) / 2
    pred = (Md - M0) @ x0
    for k in range(6):
        scale = np.max(np.abs(pred[k]))
        assert scale > 1e-8 * np.max(np.abs(x0[k])), k  # damping moves every coordinate
        assert np.max(np.abs(odd[k] - pred[k])) < (5e-3 if k == 4 else 1e-3) * scale, k
    z = oracle.track_bunch_rad(eles, abi.ELECTRON, 1, np.zeros((6, 1)), maps)
    assert z["pz"][0] == pytest.approx(c[5], rel=1e-6)
    assert c[5] == pytest.approx(maps[:, :, 47].sum(), rel=1e-6)


def test_fluctuating_tracking_draws_the_philox_stream():
    """A drift-free check of the stochastic kick: one bend with only its
    stochastic matrices, damping off: each particle's change at the exit edge
    (entrance-edge kick carried through the bend) is S0, S1 applied to the
    Philox deviates of (particle, element, edge, turn, call)."""
    eles = no_apertures(lattice.fodo_ring_c0(n_cells=1))
    ib = [i for i, e in enumerate(eles) if e.key == abi.KEY_SBEND][0]
    maps = oracle.rad_maps(eles, abi.ELECTRON, np.zeros((len(eles), 6)))
    one = [eles[ib]]
    m1 = maps[ib:ib + 1]
    x0 = np.zeros((6, 8))
    seed, call = 0xDEADBEEF12345, 3
    out = oracle.track_bunch_rad(one, abi.ELECTRON, 1, x0, m1, damping_on=False, fluctuations_on=True, seed=seed,
                                 call=call)
    M = oracle.make_mat6(eles[ib], abi.ELECTRON, [0.0] * 6)[1]
    S0, S1 = m1[0, 0, 48:].reshape(6, 6), m1[0, 1, 48:].reshape(6, 6)
    for p in range(8):
        g0 = oracle.rad_gauss6(p, 0, 0, 0, call, seed)
        g1 = oracle.rad_gauss6(p, 0, 1, 0, call, seed)
        pred = M @ (S0 @ g0) + S1 @ g1
        got = np.array([out[n][p] for n in ("x", "px", "y", "py", "z", "pz")])
        # to the bend's second-order terms (z: the path length at second order in the pz kick, ~1e-14 m)
        # (a deviate drawn with another counter differs by order one)
        assert np.all(np.abs(got - pred) < 1e-5 * np.abs(pred) + 5e-14)


# ---------------------------------------------------------------------------
# closed_orbit_test: the SLS-2 rings ad05f_pos_RF.lat and ad05f_neg_RF.lat
# ---------------------------------------------------------------------------
with open(os.path.join(ROOT, "tests", "golden", "closed_orbit_ad05f.json")) as _f:
    CO = json.load(_f)
VEC = ("x", "px", "y", "py", "z", "pz")


def g_integrals(eles):
    """g_integrals_calc (bmad/low_level/g_integrals_calc.f90:44-51): sums of
    g^2 L and g^3 L over the bends, which set the Romberg tolerance."""
    g2 = sum(e.g ** 2 * e.l for e in eles if e.key == abi.KEY_SBEND)
    g3 = sum(e.g ** 3 * e.l for e in eles if e.key == abi.KEY_SBEND)
    return g2, g3


def oracle_closed_orbit_6d(arr, maps, wavelen):
    """The oracle's one-turn map with damping closed in all six coordinates:
    Newton steps with a finite-difference Jacobian from z = 0, z kept in
    [-wavelen / 2, wavelen / 2), then, as closed_orbit_calc does when the
    orbit found is longitudinally unstable, the most stable of the starts
    z + j wavelen / 8 and Newton again."""
    h = 1e-8

    def fold(orb):
        z = orb[4] % wavelen
        orb[4] = z - wavelen if z >= wavelen / 2 else z
        return orb

    def turn(orb):
        c = np.tile(np.asarray(orb, float)[:, None], (1, 7))
        for j in range(6):
            c[j, j + 1] += h
        out = oracle.track_bunch_rad(arr, abi.ELECTRON, 1, c, maps, damping_on=True)
        assert (out["state"] == abi.ALIVE).all()
        v = np.stack([out[k] for k in VEC])
        return v[:, 0], (v[:, 1:] - v[:, :1]) / h

    def newton(orb):
        for _ in range(30):
            t, m = turn(orb)
            dx = np.linalg.solve(np.eye(6) - m, t - orb)
            orb = fold(orb + dx)
            if np.abs(dx).max() < 1e-12:
                break
        return orb, m

    orb, m = newton(np.zeros(6))
    if np.abs(np.linalg.eigvals(m)).max() > 1.000001:
        starts = [orb + np.array([0, 0, 0, 0, j * wavelen / 8, 0]) for j in range(-3, 5) if j != 0]
        rho = [np.abs(np.linalg.eigvals(turn(s)[1])).max() for s in starts]
        orb, m = newton(starts[int(np.argmin(rho))])
    return orb, m


def z_tune(arr, maps, orb):
    """calc_z_tune (bmad/code/calc_z_tune.f90:36-59) of the matrix
    lat_make_mat6 builds about the closed orbit: each element's make_mat6
    (its body without radiation) at the orbit's entrance into it, the orbit
    carried on with radiation (track1_radiation moves the orbit, not
    ele%mat6).  Returns (tune / 2 pi, orbit after one turn)."""
    t1 = np.eye(6)
    v = np.array(orb, float).reshape(6, 1)
    beta = None
    for i in range(len(arr)):
        t1 = oracle.make_mat6(arr[i], abi.ELECTRON, v[:, 0])[1] @ t1
        out = oracle.track_bunch_rad(oracle.abi_array([arr[i]]), abi.ELECTRON, 1, v, maps[i:i + 1],
                                     damping_on=True, beta=beta)
        v = np.stack([out[k] for k in VEC])
        beta = out["beta"]
    a = t1
    denom = 2 * (a[4, 4] * a[5, 5] - a[4, 5] * a[5, 4])
    assert (a[4, 4] + a[5, 5]) / denom - 1 < -1e-9  # stable
    ev = np.linalg.eigvals(a)
    ang = np.abs(np.angle(ev))
    # the synchrotron mode is the pair of the smallest phase advance
    return np.sign(a[4, 5]) * ang.min() / (2 * np.pi), v[:, 0]


@pytest.mark.parametrize("name", sorted(CO["lattices"]))
def test_closed_orbit_z_and_synchrotron_tune_with_damping(name):
    """closed_orbit_test (regression_tests/closed_orbit_test/closed_orbit_test.f90:
    45-80, output.correct:4-13): the 6-d closed orbit's z of the 3120-element
    SLS-2 ring with radiation damping on, and its synchrotron tune.  With the
    cavities on, z is where the rf restores the energy the bends' maps take
    per turn (sin(2 pi z / lambda) = U0 / V), so it pins the maps' xfer_damp_vec
    and damping matrices (radiation_map_setup about the zero orbit, which the
    parser's lat_make_mat6 leaves in map_ref_orb_in), bends with gradients
    and edge angles, the thick sextupoles and the rf to 5e-8 m.  The tune
    pins the matrices along that orbit."""
    eles = ring_from_fixture(CO, name)
    arr = oracle.abi_array(eles)
    g2, g3 = g_integrals(eles)
    maps = oracle.rad_maps(arr, abi.ELECTRON, np.zeros((len(arr), 6)), g2, g3)
    f_rf = min(e.rf_frequency for e in eles if e.key == abi.KEY_RFCAVITY)
    wavelen = C_LIGHT * eles[0].p0c / eles[0].e_tot / f_rf
    orb, _ = oracle_closed_orbit_6d(arr, maps, wavelen)
    rows = CO["rows"]
    for tag in ("1-B", "2-A", "2-B"):
        r = rows[f"{name}-{tag} Z"]
        assert abs(orb[4] - r["value"]) <= r["tol"], (tag, orb[4], r["value"])
    tune, closing = z_tune(arr, maps, orb)
    assert np.abs(closing - orb).max() < 1e-12
    for tag in ("1-B", "2-B"):
        r = rows[f"{name}-{tag} Z_tune"]
        assert abs(tune - r["value"]) <= r["tol"], (tag, tune, r["value"])
