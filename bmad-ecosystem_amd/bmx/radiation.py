"""Radiation maps of a lattice's elements about the design orbit, built on
libbmx's exact transfer matrices.

radiation_map_setup (bmad/modules/radiation_mod.f90:160-232) makes, per
element, the two maps track1_radiation applies (rm0 at the entrance, rm1 at
the exit): each half of the element (create_element_slice) gets the damping
and stochastic matrices of rad1_damp_and_stoc_mats
(bmad/modules/rad_6d_mod.f90:406-718), a Romberg integral along the half of
the radiation at each point (calc_rad_at_pt, :590-686) transported back by
the inverse body matrix, referred to the half's entrance or exit and the
stochastic one Cholesky-factored (tracking_rad_map_setup,
bmad/low_level/tracking_rad_map_setup.f90:19-84).  A Bmad caller has these
maps already (ele%rad_map); this module builds them for a lattice given as
element records, for the case the reference orbit is the design orbit: only
sector bends radiate there (every other element keeps the orbit's angle and
gets no map, as radiation_map_setup returns early for it).  The transfer
matrices of the bend's slices come from bmx_transfer_matrix (the device's
forward-mode derivatives of its own tracking).  Refused with ValueError:
misaligned or ref-tilted bends, bends with kicks or electric multipoles, and
any element that moves the design orbit.
"""
import math

import numpy as np

from . import abi, tracker

R_E = 2.8179403227e-15                                      # physical_constants.f90:35
H_BAR = 4.135667696e-15 / (2 * math.pi)                     # :38-39
SQRT3 = 1.732050807568877293527446341505872
RAD_MAP = 84  # ref_orb[6], damp_dmat[36], xfer_damp_vec[6], stoc_mat[36]


def _symp_conj(m):
    """mat_symp_conj: -S M^T S (sim_utils/matrix/mat_symp_conj.f90)."""
    s = np.zeros((6, 6))
    for i in range(0, 6, 2):
        s[i, i + 1] = 1.0
        s[i + 1, i] = -1.0
    return -s @ m.T @ s


def _polint(xa, ya, x=0.0):
    """Neville extrapolation to x with its error estimate (super_polint,
    sim_utils/math/super_recipes_mod.f90)."""
    n = len(xa)
    c = np.array(ya, dtype=np.float64)
    d = c.copy()
    ho = (np.asarray(xa, dtype=np.float64) - x).reshape((-1,) + (1,) * (c.ndim - 1))  # broadcast over the values
    ns = int(np.argmin(np.abs(x - np.asarray(xa))))
    y = c[ns]
    ns -= 1
    dy = 0.0 * y
    for m in range(1, n):
        den = (c[1:n - m + 1] - d[:n - m]) / (ho[:n - m] - ho[m:n])
        d = np.concatenate([ho[m:n] * den, d[n - m:]])
        c = np.concatenate([ho[:n - m] * den, c[n - m:]])
        if 2 * (ns + 1) < n - m:
            dy = c[ns + 1]
        else:
            dy = d[ns]
            ns -= 1
        y = y + dy
    return y, dy


def _slice(e, l_slice, inc_entrance, inc_exit):
    """A slice of a sector bend with the ends kept as asked
    (makeup_super_slave1, bmad/modules/bookkeeper_mod.f90): fringe_at and
    the edge angles cut to the ends kept, the multipoles scaled by the
    length ratio; no aperture (the slice carries the design orbit)."""
    s = abi.EleRec.from_buffer_copy(e)
    coef = 1.0 if abs(l_slice) >= abs(e.l) else l_slice / e.l
    ds_step = abs(e.l) / max(e.num_steps, 1)
    s.l = l_slice
    s.num_steps = max(1, int(math.floor(abs(l_slice / ds_step) + 0.5)))
    ifr = e.fringe_at
    if inc_entrance and inc_exit:
        pass
    elif inc_entrance:
        s.fringe_at = abi.ENTRANCE_END if ifr in (abi.ENTRANCE_END, abi.BOTH_ENDS) else abi.NO_END
    elif inc_exit:
        s.fringe_at = abi.EXIT_END if ifr in (abi.EXIT_END, abi.BOTH_ENDS) else abi.NO_END
    else:
        s.fringe_at = abi.NO_END
    for n in range(abi.N_POLE):
        s.a[n] = e.a[n] * coef
        s.b[n] = e.b[n] * coef
    s.x1_limit = s.x2_limit = s.y1_limit = s.y2_limit = 0.0
    s.delta_ref_time = e.delta_ref_time * coef
    if s.fringe_at in (abi.NO_END, abi.EXIT_END):
        s.e1 = s.fint = s.hgap = 0.0
    if s.fringe_at in (abi.NO_END, abi.ENTRANCE_END):
        s.e2 = s.fintx = s.hgapx = 0.0
    return s


class _Matrices:
    """Transfer matrices at the design orbit of single-element lattices,
    one upload and one bmx_transfer_matrix call each, remembered by slice."""

    def __init__(self, ctx, species):
        self.ctx, self.species, self.cache = ctx, species, {}

    def __call__(self, rec, key):
        if key in self.cache:
            return self.cache[key]
        if rec.l == 0:
            m = np.eye(6)
        else:
            lat = self.ctx.lattice([rec], self.species, aperture_limit_on=False)
            out, mat, state = tracker.transfer_matrix(lat, np.zeros(6))
            lat.close()
            if state[0] != abi.ALIVE or np.max(np.abs(out)) > 1e-15:
                raise ValueError("the element moves the design orbit")
            m = mat[0]
        self.cache[key] = m
        return m


def _bend_half(e, half, mats, kd, kf, gamma0, g2_tol, g3_tol):
    """rad1_damp_and_stoc_mats of one half of a bend on the design orbit,
    then tracking_rad_map_setup's referral and Cholesky factor."""
    length = 0.5 * e.l
    runt = _slice(e, length, half == 0, half == 1)
    m_half = mats(runt, ("half", half))
    # offsets and fringe at the upstream end: the entrance edge of the first
    # half (the body matrix of the half times it is the half's matrix)
    if half == 0:
        mat0 = np.linalg.solve(mats(_slice(e, length, False, False), ("body", length)), m_half)
    else:
        mat0 = np.eye(6)
    # the field on the design orbit: B_y = (g + dg) p0c / (c q), gradient k1
    # (em_field_calc's bend field and its multipoles, :580-707); g of the
    # particle and its x, y derivatives (g_bending_strength_from_em_field)
    gx = e.g + e.dg
    g1 = abs(gx)
    g2 = g1 * g1
    g3 = g1 * g2
    dg2_dx = 2.0 * gx * e.k1
    dg2_dy = 0.0
    dg2_dx = dg2_dx + g2 * e.g  # the path-length factor (1 + g x) of the sector bend, x = 0
    w_vec = np.array([0.0, 0.0, 0.0, 0.0, 0.0, g2])

    def at(z):
        mb = _symp_conj(mats(_slice(e, z, False, False), ("body", z)))
        m2, m4, m6 = mb[:, 1], mb[:, 3], mb[:, 5]
        dd = np.zeros((6, 6))
        dd[:, 0] = -kd * (m6 * dg2_dx)
        dd[:, 1] = -kd * (m2 * g2)
        dd[:, 2] = -kd * (m6 * dg2_dy)
        dd[:, 3] = -kd * (m4 * g2)
        dd[:, 5] = -kd * (m6 * 2.0 * g2)
        dv = -kd * (mb @ w_vec)
        kfv = kf * g3
        ss = kfv * np.outer(m6, m6)
        kv = kf * 13.0 * g1 / (55.0 * gamma0 * gamma0)  # the photons' opening angle
        ss = ss + kv * (gx * gx * np.outer(m4, m4))
        return dd, dv, ss

    eps_damp = kd * g2_tol / length
    eps_stoc = kf * g3_tol / length
    qi = [dict(h=4.0, D=np.zeros((6, 6)), v=np.zeros(6), S=np.zeros((6, 6)))]
    for j in range(1, 11):
        if j == 1:
            pts, del_z = [0.0, length], length
        else:
            n_pts = 2 ** (j - 2)
            del_z = length / n_pts
            pts = [del_z / 2 + k * del_z for k in range(n_pts)]
        sd, sv, ss = np.zeros((6, 6)), np.zeros(6), np.zeros((6, 6))
        for z in pts:
            dd, dv, s1 = at(z)
            sd, sv, ss = sd + dd, sv + dv, ss + s1
        prev = qi[-1]
        qi.append(dict(h=0.25 * prev["h"], D=0.5 * (prev["D"] + del_z * sd), v=0.5 * (prev["v"] + del_z * sv),
                       S=0.5 * (prev["S"] + del_z * ss)))
        if len(qi) > 5:
            qi.pop(0)
        if j < 3:
            continue
        hs = [q["h"] for q in qi[1:]]
        D1, ddamp = _polint(hs, [q["D"] for q in qi[1:]])
        v1, ddvec = _polint(hs, [q["v"] for q in qi[1:]])
        S1, dstoc = _polint(hs, [q["S"] for q in qi[1:]])
        d_max = max(eps_damp * np.max(np.abs(ddamp)), eps_damp * np.max(np.abs(ddvec)),
                    eps_stoc * np.max(np.abs(dstoc)))
        if d_max < 1:
            break
    # bend edges (add_bend_edge_to_damp_mat): the entrance edge at the body's
    # entrance orbit, the exit edge at the orbit the integration saved at its
    # first point (the body's entrance as well), both the design orbit here
    for e_edge in ((runt.e1, runt.e2) if e.orientation == 1 else (runt.e2, runt.e1)):
        if e_edge == 0:
            continue
        g2x = -math.tan(e_edge) * (e.g + e.dg) ** 2
        dm = np.zeros((6, 6))
        dm[5, 0] = -kd * g2x
        dm[5, 2] = -kd * -g2x
        D1 = D1 + dm
    mat0_inv = _symp_conj(mat0)
    t = m_half @ mat0_inv
    vec = t @ v1
    damp = t @ D1 @ mat0
    stoc = m_half @ (mat0_inv @ S1 @ mat0_inv.T) @ m_half.T
    if half == 0:  # referred to the upstream end
        m_inv = _symp_conj(m_half)
        damp = m_inv @ damp
        vec = m_inv @ vec
        stoc = m_inv @ stoc @ m_inv.T
    try:
        stoc = np.tril(np.linalg.cholesky(stoc))
    except np.linalg.LinAlgError:
        stoc = np.zeros((6, 6))
    out = np.zeros(RAD_MAP)
    out[6:42] = damp.reshape(36)
    out[42:48] = vec
    out[48:84] = stoc.reshape(36)
    return out


def design_orbit_maps(ctx, eles, species, g2_integral=0.0, g3_integral=0.0):
    """Radiation maps [n_ele, 2, 84] (rm0, rm1 in the bmx_rad_map layout) of
    every element about the design orbit, for bmx_lattice_upload_rad.
    g2_integral, g3_integral: the branch's integrals of g^2 and g^3, which set
    the Romberg tolerance (1e-4 / total length of each) as the reference's
    branch%param does; 0 stops at the third refinement."""
    arr = list(eles)
    n = len(arr)
    maps = np.zeros((n, 2, RAD_MAP))
    total = sum(e.l for e in arr)
    tol = 1e-4 / total if total else 0.0
    mc2 = abi.MASS[species]
    mats = _Matrices(ctx, species)
    done = {}
    for i, e in enumerate(arr):
        if e.l == 0 or e.key != abi.KEY_SBEND:
            if e.hkick != 0 or e.vkick != 0 or e.kick != 0:
                raise ValueError(f"element {i}: a kick moves the design orbit")
            continue
        if (e.x_offset or e.y_offset or e.z_offset or e.x_pitch or e.y_pitch or e.ref_tilt or e.hkick or e.vkick
                or e.has_elec_poles):
            raise ValueError(f"element {i}: a misaligned, tilted or kicked bend is not on the design orbit")
        if e.p0c < 2.0 * mc2:
            continue
        key = bytes(abi.EleRec.from_buffer_copy(e))
        if key not in done:
            gamma0 = math.sqrt(e.p0c ** 2 + mc2 ** 2) / mc2
            q = abi.CHARGE[species]
            radi = R_E * abi.M_ELECTRON * q * q / mc2
            kd = 2.0 / 3.0 * radi * gamma0 ** 3
            kf = 55.0 * H_BAR * abi.C_LIGHT / (24.0 * SQRT3) * radi * gamma0 ** 5 / mc2
            mats.cache = {}
            done[key] = np.stack([_bend_half(e, h, mats, kd, kf, gamma0, tol * g2_integral, tol * g3_integral)
                                  for h in (0, 1)])
        maps[i] = done[key]
    return maps
