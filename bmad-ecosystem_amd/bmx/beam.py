"""Linear optics of a ring (transfer matrices, periodic Twiss) and seeded,
Twiss-matched Gaussian bunches -- the set-up that feeds the tracker.

The matched-beam recipe follows init_random_distribution
(bmad/multiparticle/beam_utils.f90:539-709): Gaussian r(6), then
x = sqrt(eps beta) r1, px = sqrt(eps / beta) (r2 - alpha r1) per plane, z and pz
independent.  The random stream is numpy's PCG64 with a recorded seed (the
reference's ran_gauss stream is not reproduced).
"""
import math

import numpy as np

from . import abi


def _drift(l):
    m = np.eye(4)
    m[0, 1] = l
    m[2, 3] = l
    return m


def _quad(l, k1):
    m = np.eye(4)
    for k, sl in ((k1, slice(0, 2)), (-k1, slice(2, 4))):
        if k > 0:
            s = math.sqrt(k)
            blk = [[math.cos(s * l), math.sin(s * l) / s], [-s * math.sin(s * l), math.cos(s * l)]]
        elif k < 0:
            s = math.sqrt(-k)
            blk = [[math.cosh(s * l), math.sinh(s * l) / s], [s * math.sinh(s * l), math.cosh(s * l)]]
        else:
            blk = [[1.0, l], [0.0, 1.0]]
        m[sl, sl] = blk
    return m


def _sbend(l, g, k1=0.0):
    m = np.eye(4)
    kx = k1 + g * g
    for k, sl in ((kx, slice(0, 2)), (-k1, slice(2, 4))):
        if k > 0:
            s = math.sqrt(k)
            blk = [[math.cos(s * l), math.sin(s * l) / s], [-s * math.sin(s * l), math.cos(s * l)]]
        elif k < 0:
            s = math.sqrt(-k)
            blk = [[math.cosh(s * l), math.sinh(s * l) / s], [s * math.sinh(s * l), math.cosh(s * l)]]
        else:
            blk = [[1.0, l], [0.0, 1.0]]
        m[sl, sl] = blk
    return m


def _edge(g, e):
    m = np.eye(4)
    m[1, 0] = g * math.tan(e)
    m[3, 2] = -g * math.tan(e)
    return m


def transfer_matrix_4d(eles):
    """4x4 (x, px, y, py) on-momentum linear map of an element list."""
    m = np.eye(4)
    for e in eles:
        k = e.key
        if k == abi.KEY_QUADRUPOLE:
            t = _quad(e.l, e.k1)
        elif k == abi.KEY_SBEND:
            t = _edge(e.g, e.e2) @ _sbend(e.l, e.g, e.k1) @ _edge(e.g, e.e1)
        elif k in (abi.KEY_MULTIPOLE,):
            t = np.eye(4)
            kl = e.a[1] * math.cos(-2 * e.b[1]) if e.has_mag_poles else 0.0
            t[1, 0] = -kl
            t[3, 2] = kl
        else:
            t = _drift(e.l)
        m = t @ m
    return m


def periodic_twiss(eles):
    """(beta_x, alpha_x, beta_y, alpha_y, nu_x, nu_y) at the lattice start from the
    one-period 4x4 map (uncoupled)."""
    m = transfer_matrix_4d(eles)
    out = []
    tunes = []
    for sl in (slice(0, 2), slice(2, 4)):
        b = m[sl, sl]
        cmu = 0.5 * (b[0, 0] + b[1, 1])
        if abs(cmu) >= 1:
            raise ValueError("lattice period is linearly unstable")
        smu = math.copysign(math.sqrt(1 - cmu * cmu), b[0, 1])
        beta = b[0, 1] / smu
        alpha = (b[0, 0] - b[1, 1]) / (2 * smu)
        out += [beta, alpha]
        mu = math.atan2(smu, cmu)
        tunes.append((mu if mu >= 0 else mu + 2 * math.pi) / (2 * math.pi))
    return tuple(out) + tuple(tunes)


def twiss_from_mat2(m):
    """(beta, alpha, phi) of one plane from its 2x2 one-turn block, as
    twiss_from_mat2 (bmad/low_level/twiss_from_mat2.f90:19-60): the block
    normalised by its determinant, cos phi = trace / 2, phi in [0, 2 pi).
    Raises ValueError for an unstable or non-symplectic block."""
    m = np.asarray(m, dtype=np.float64)
    det = m[0, 0] * m[1, 1] - m[0, 1] * m[1, 0]
    m = m / det
    t_cos = (m[0, 0] + m[1, 1]) / 2.0
    if abs(t_cos) >= 1.0:
        raise ValueError("unstable matrix")
    t_sin = math.copysign(1.0, m[0, 1]) * math.sqrt(1.0 - t_cos * t_cos)
    phi = math.atan2(t_sin, t_cos)
    if phi < 0:
        phi += 2 * math.pi
    alpha = (m[0, 0] - m[1, 1]) / (2.0 * t_sin)
    radical = -(1 + alpha * alpha) * m[0, 1] / m[1, 0]
    if radical <= 0:
        raise ValueError("non-symplectic matrix")
    return math.sqrt(radical), alpha, phi


def twiss_from_tracking(lat, orbit=None, d_orb=None, variant=0):
    """Periodic Twiss (beta_x, alpha_x, beta_y, alpha_y, nu_x, nu_y) at the
    lattice start from the one-turn matrix the device forms by tracking about
    orbit (tracker.one_turn_matrix), its x and y blocks taken as uncoupled, as
    twiss_from_tracking (bmad/code/twiss_from_tracking.f90) feeds
    twiss_at_start.  Also returns the matrix."""
    from . import tracker
    m, status = tracker.one_turn_matrix(lat, orbit, d_orb)
    if status[variant] != 0:
        raise ValueError("a probe particle was lost")
    M = m[variant]
    bx, ax, px = twiss_from_mat2(M[0:2, 0:2])
    by, ay, py = twiss_from_mat2(M[2:4, 2:4])
    return (bx, ax, by, ay, px / (2 * math.pi), py / (2 * math.pi)), M


def _symp_conj2(a):
    """Symplectic conjugate of a 2x2 block: [[d, -b], [-c, a]]."""
    return np.array([[a[1, 1], -a[0, 1]], [-a[1, 0], a[0, 0]]])


def twiss_coupled(m):
    """Normal-mode (coupled) periodic Twiss from a one-turn matrix m (6x6 or
    4x4 transverse block; 6x6 also gives the dispersion), as Bmad's
    twiss_from_mat6 (bmad/code/twiss_from_mat6.f90:86-139) forms it with
    mat_symp_decouple (bmad/low_level/mat_symp_decouple.f90:63-113), the
    Edwards-Teng / Sagan-Rubin decomposition T = V U V^-1:
      with T = [[M, m], [n, N]] and H = m + conj(n) (conj = symplectic
      conjugate), t = tr M - tr N, D = t^2 + 4 det H (D <= 0: stop band),
      gamma = sqrt(1/2 + 1/2 sqrt(t^2 / D)), C = -sign(t) H / (gamma sqrt D),
      V = [[gamma I, C], [-conj(C), gamma I]], U = conj(V) T V,
    the a and b modes' Twiss from U's diagonal blocks (twiss_from_mat2), the
    lab dispersion (I - T4)^-1 m[0:4, 5] and the normal-mode one conj(V)
    times it.  Returns a dict (beta_a, alpha_a, phi_a, beta_b, alpha_b,
    phi_b, gamma_c, cmat [2x2], eta_x, etap_x, eta_y, etap_y, eta_a, etap_a,
    eta_b, etap_b when m is 6x6).  Raises ValueError in a stop band or for
    an unstable mode."""
    m = np.asarray(m, dtype=np.float64)
    t4 = m[:4, :4]
    M, mm, n, N = t4[0:2, 0:2], t4[0:2, 2:4], t4[2:4, 0:2], t4[2:4, 2:4]
    H = mm + _symp_conj2(n)
    t = np.trace(M) - np.trace(N)
    D = t * t + 4.0 * np.linalg.det(H)
    if np.max(np.abs(H)) == 0.0:
        gamma, C = 1.0, np.zeros((2, 2))
    else:
        if D <= 0.0:
            raise ValueError("stop band: the modes cannot be decoupled")
        gamma = math.sqrt(0.5 + 0.5 * math.sqrt(t * t / D))
        C = -math.copysign(1.0, t) * H / (gamma * math.sqrt(D))
    V = np.zeros((4, 4))
    V[0:2, 0:2] = V[2:4, 2:4] = gamma * np.eye(2)
    V[0:2, 2:4] = C
    V[2:4, 0:2] = -_symp_conj2(C)
    Vc = np.zeros((4, 4))  # symplectic conjugate of V
    Vc[0:2, 0:2] = Vc[2:4, 2:4] = gamma * np.eye(2)
    Vc[0:2, 2:4] = -C
    Vc[2:4, 0:2] = _symp_conj2(C)
    U = Vc @ t4 @ V
    ba, aa, pa = twiss_from_mat2(U[0:2, 0:2])
    bb, ab, pb = twiss_from_mat2(U[2:4, 2:4])
    out = dict(beta_a=ba, alpha_a=aa, phi_a=pa, beta_b=bb, alpha_b=ab, phi_b=pb, gamma_c=gamma, cmat=C)
    if m.shape == (6, 6):
        eta = np.linalg.solve(np.eye(4) - t4, m[0:4, 5])
        ev = Vc @ eta
        out.update(eta_x=eta[0], etap_x=eta[1], eta_y=eta[2], etap_y=eta[3], eta_a=ev[0], etap_a=ev[1],
                   eta_b=ev[2], etap_b=ev[3])
    return out


def closed_orbit_4d(one_turn, pz, guess=None, tol=1e-16, max_iter=40):
    """Transverse closed orbit at fixed pz by Newton steps on the exact
    one-turn map: one_turn(v) -> (out [6], J [6, 6]).  Returns v [6]."""
    v = np.zeros(6) if guess is None else np.array(guess, dtype=np.float64)
    v[5] = pz
    for _ in range(max_iter):
        out, J = one_turn(v)
        dv = np.linalg.solve(J[:4, :4] - np.eye(4), -(out[:4] - v[:4]))
        v[:4] += dv
        if np.max(np.abs(dv)) <= tol * max(1.0, np.max(np.abs(v[:4]))):
            break
    return v


def chromatic_twiss(one_turn, pz=0.0, delta=1e-4):
    """d(beta, alpha)/d pz of both normal modes as chrom_calc forms them
    (bmad/code/chrom_calc.f90:72-137): the rings' Twiss (rf off) at the
    transverse closed orbits of pz +- delta (delta_e, default 1e-4), and
    central differences over 2 delta.  Also the chromaticities (tune
    change / (2 delta)).  one_turn as closed_orbit_4d."""
    tw = {}
    for s in (1, -1):
        v = closed_orbit_4d(one_turn, pz + s * delta)
        tw[s] = twiss_coupled(one_turn(v)[1])
    out = {}
    for k in ("beta", "alpha"):
        for mode in ("a", "b"):
            out[f"d{k}_dpz_{mode}"] = (tw[1][f"{k}_{mode}"] - tw[-1][f"{k}_{mode}"]) / (2 * delta)
    for mode in ("a", "b"):
        dphi = tw[1][f"phi_{mode}"] - tw[-1][f"phi_{mode}"]
        dphi = (dphi + math.pi) % (2 * math.pi) - math.pi
        out[f"chrom_{mode}"] = dphi / (2 * math.pi) / (2 * delta)
    return out


def twiss_coupled_from_tracking(lat, orbit=None):
    """Normal-mode periodic Twiss at the lattice start from the device's
    exact one-turn matrix about orbit (tracker.transfer_matrix; the closed
    orbit, zero when None).  Upload the lattice without rf for the ring's
    optics, as Bmad's Twiss calculation does."""
    from . import tracker
    v = np.zeros(6) if orbit is None else np.asarray(orbit, dtype=np.float64)
    _, mat, st = tracker.transfer_matrix(lat, v)
    if st[0] != 1:
        raise ValueError("the orbit was lost within one turn")
    return twiss_coupled(mat[0]), mat[0]


def matched_gaussian(n, twiss, eps_x, eps_y, sigma_z, sigma_pz, seed):
    """(6, n) float64 array of a Twiss-matched 6-D Gaussian bunch."""
    bx, ax, by, ay = twiss[:4]
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.standard_normal((6, n))
    c = np.empty((6, n))
    c[0] = math.sqrt(eps_x * bx) * r[0]
    c[1] = math.sqrt(eps_x / bx) * (r[1] - ax * r[0])
    c[2] = math.sqrt(eps_y * by) * r[2]
    c[3] = math.sqrt(eps_y / by) * (r[3] - ay * r[2])
    c[4] = sigma_z * r[4]
    c[5] = sigma_pz * r[5]
    return c


def c0_bunch(eles_cell, n=10_000, seed=20261203):
    """The 1 GeV FODO-ring bunch: eps_x = 10 nm, eps_y = 0.1 nm, sigma_z = 1 mm,
    sigma_pz = 1e-3, matched to the periodic Twiss of one cell."""
    tw = periodic_twiss(eles_cell)
    return matched_gaussian(n, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed), tw
