"""Exact transfer matrices on the device (bmx_transfer_matrix) against the
oracle's dual-number Jacobian of the same map (oracle.jacobian, itself
pinned by the symplectic invariants and fringe_test's matrix rows):

* every element kind alone, at random orbits, at two energies;
* the one-turn matrix of the FODO ring, of the 4000-element ring with and
  without rf, at matched orbits: equal to the oracle's and symplectic,
  ||M^T S M - S|| <= 1e-12;
* sub-ranges and several turns chain like products of the pieces;
* a lost orbit returns its loss state and a zero matrix.
"""
import numpy as np
import pytest

from bmx import abi, beam, lattice, tracker
import oracle
from helpers import NAMES, coord_errors
from test_gpu_parity import ELEMENT_CASES, SCALE, _coords

pytestmark = pytest.mark.gpu

S = np.zeros((6, 6))
for _k in range(3):
    S[2 * _k, 2 * _k + 1], S[2 * _k + 1, 2 * _k] = 1.0, -1.0


def _symp(M):
    return float(np.max(np.abs(M.T @ S @ M - S)))


def _rel(A, B):
    return float(np.max(np.abs(A - B)) / max(np.max(np.abs(B)), 1.0))


@pytest.mark.parametrize("name", sorted(ELEMENT_CASES))
@pytest.mark.parametrize("e_tot", [1e9, 3e6])
def test_element_matrices_match_oracle(ctx, name, e_tot):
    eles = lattice.set_reference([lattice.EleRec_copy(e) for e in ELEMENT_CASES[name]], abi.ELECTRON, e_tot=e_tot)
    coords = _coords(16, 7, [s * 0.25 for s in SCALE])
    lat = ctx.lattice(eles, abi.ELECTRON)
    out, M, st = tracker.transfer_matrix(lat, coords.T)
    lat.close()
    worst = 0.0
    n_ok = 0
    ok, ref = [], []
    for i in range(coords.shape[1]):
        try:
            o, J = oracle.jacobian(eles, abi.ELECTRON, 1, coords[:, i])
        except RuntimeError:
            assert st[i] != abi.ALIVE
            assert np.all(M[i] == 0.0)
            continue
        assert st[i] == abi.ALIVE
        n_ok += 1
        worst = max(worst, _rel(M[i], J))
        ok.append(i)
        ref.append(o)
    # the tracked orbits by the parity metric of the tracking tests (z against
    # max(|z|, 1e-3 x the path length): z is a small difference of path lengths)
    errs = coord_errors({k: out[ok, c] for c, k in enumerate(NAMES)},
                        {k: np.array(ref)[:, c] for c, k in enumerate(NAMES)}, slice(None), sum(e.l for e in eles))
    print(f"{name} {e_tot:.0e}: {n_ok} orbits, max |M - J| / max(|J|, 1) = {worst:.2e}, orbit errors "
          f"{ {k: f'{v:.1e}' for k, v in errs.items()} }", flush=True)
    assert n_ok > 0
    assert worst <= 1e-12
    assert max(errs.values()) <= 1e-12, errs


@pytest.mark.parametrize("ring", ["c0", "c1_no_rf", "c1"])
def test_one_turn_matrix_exact_and_symplectic(ctx, ring):
    if ring == "c0":
        eles = lattice.fodo_ring_c0()
        coords, _ = beam.c0_bunch(eles[:7], n=4, seed=9)
    else:
        eles = lattice.ring_c1(n_rf=0 if ring == "c1_no_rf" else 8)
        tw = beam.periodic_twiss(lattice.ring_c1(n_rf=0)[:8])
        coords = beam.matched_gaussian(4, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed=9)
    lat = ctx.lattice(eles, abi.ELECTRON)
    out, M, st = tracker.transfer_matrix(lat, coords.T)
    lat.close()
    assert np.all(st == abi.ALIVE)
    for i in range(coords.shape[1]):
        o, J = oracle.jacobian(eles, abi.ELECTRON, 1, coords[:, i])
        err = _rel(M[i], J)
        sd, so = _symp(M[i]), oracle.symp_error(J)
        print(f"{ring} orbit {i}: |M - J| / max|J| = {err:.2e}, symplectic error device {sd:.2e} oracle {so:.2e}",
              flush=True)
        assert err <= 1e-11
        assert sd <= 1e-12


def test_ranges_and_turns_chain(ctx):
    """M(0 -> 30) M(30 -> 0 through the end) = M(one turn), and three turns
    = the product of the three one-turn matrices along the orbit."""
    eles = lattice.fodo_ring_c0(n_cells=10)
    coords, _ = beam.c0_bunch(eles[:7], n=2, seed=4)
    lat = ctx.lattice(eles, abi.ELECTRON)
    v = coords[:, 0]
    o1, A, _ = tracker.transfer_matrix(lat, v, 0, 30)
    o2, B, _ = tracker.transfer_matrix(lat, o1[0], 30, 0)
    oT, T, _ = tracker.transfer_matrix(lat, v)
    assert np.allclose(o2[0], oT[0], rtol=1e-13, atol=1e-17)
    assert _rel(B[0] @ A[0], T[0]) <= 1e-13
    o3, T3, _ = tracker.transfer_matrix(lat, v, n_turns=3)
    P, w = np.eye(6), v
    for _ in range(3):
        w2, Mi, _ = tracker.transfer_matrix(lat, w)
        P = Mi[0] @ P
        w = w2[0]
    assert np.allclose(o3[0], w, rtol=1e-13, atol=1e-17)
    assert _rel(T3[0], P) <= 1e-12
    _, J3 = oracle.jacobian(eles, abi.ELECTRON, 3, v)
    assert _rel(T3[0], J3) <= 1e-11
    lat.close()


def test_lost_orbit_zero_matrix(ctx):
    eles = lattice.fodo_ring_c0(n_cells=4)
    lat = ctx.lattice(eles, abi.ELECTRON)
    v = np.array([[0.0, 0, 0, 0, 0, 0], [0.5, 0, 0, 0, 0, 0]])
    out, M, st = tracker.transfer_matrix(lat, v)
    lat.close()
    assert st[0] == abi.ALIVE and st[1] != abi.ALIVE
    assert np.all(M[1] == 0.0) and _symp(M[0]) <= 1e-13


def test_mat6_golden_rows_on_the_device(ctx):
    """mat6_calc_method_test's Bmad_Standard matrices (make_mat6 with the
    make_matrix branches), vec0 and symplectic errors of the 23 elements of
    this path, from bmx_transfer_matrix on the device at particle_start."""
    import test_oracle_golden as tog
    x0 = np.array(tog.M6["start"])
    for name in sorted(tog.M6["elements"]):
        e = tog._m6_element(name)
        lat = ctx.lattice([e], abi.POSITRON)
        out, M, st = tracker.transfer_matrix(lat, x0)
        lat.close()
        assert st[0] == abi.ALIVE, name

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

        tog.m6_check(name, M[0], out[0], fd)


def test_coupled_and_chromatic_twiss_of_the_nonlin_test_ring(ctx):
    """The Twiss nonlin_test's small ring (rf off): the device's one-turn
    matrices give the coupled Twiss Bmad recorded at its start to 1e-12 and
    chrom_calc's dbeta/dalpha_dpz of both modes to 1e-9 relative."""
    import test_oracle_golden as tog
    eles = tog._small_ring()
    lat = ctx.lattice(eles, abi.POSITRON)
    tw, T = beam.twiss_coupled_from_tracking(lat)
    tog.check_coupled_twiss(T)

    def one_turn(v):
        out, M, st = tracker.transfer_matrix(lat, v)
        assert st[0] == abi.ALIVE
        return out[0], M[0]

    res = beam.chromatic_twiss(one_turn)
    lat.close()
    g = {k: v["value"] for k, v in tog.NL["twiss"].items()}
    for k in ("dbeta_dpz_a", "dbeta_dpz_b", "dalpha_dpz_a", "dalpha_dpz_b"):
        print(f"{k}: device {res[k]:.12g} recorded {g[k]:.12g}", flush=True)
        assert abs(res[k] - g[k]) <= 1e-9 * abs(g[k]), (k, res[k], g[k])
