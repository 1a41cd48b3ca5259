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
