"""The oracle's regularised inverse (spd_inverse.py) and its eigen solver against
numpy's LAPACK eigh and against closed forms; and the reference's recorded lambda
spectrum on a plane-strain lattice (SOLID.py:350-353)."""
import numpy as np
import pytest

from oracle import oracle as O
from sigmasph_amd.scenes import free_block
from helpers import layout, oracle_state


def weight(e, tol):
    u = min(1.0, max(0.0, (e - tol) / tol))
    b = u * u * (3 - 2 * u)
    return b / max(e, tol) + (1 - b)


def test_weight_closed_form():
    tol = 0.02
    assert O.spd_weight(0.0, tol) == 1.0          # starved: no correction
    assert O.spd_weight(tol, tol) == 1.0
    assert O.spd_weight(2 * tol, tol) == pytest.approx(1.0 / (2 * tol))
    assert O.spd_weight(0.9, tol) == pytest.approx(1.0 / 0.9)
    for e in np.linspace(0, 0.1, 41):
        assert O.spd_weight(e, tol) == pytest.approx(weight(e, tol), rel=1e-14)


def random_spd(rng, lo, hi):
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    ev = rng.uniform(lo, hi, size=3)
    return Q @ np.diag(ev) @ Q.T, ev


@pytest.mark.parametrize("seed", range(20))
def test_spd_inverse3_matches_eigh(seed):
    rng = np.random.default_rng(seed)
    tol = 0.02
    E, ev = random_spd(rng, 0.0, 1.2)
    E = E + rng.normal(scale=1e-3, size=(3, 3))   # unsymmetric noise: the oracle symmetrises
    L6, me = O.spd_inverse3(E, tol)
    Es = 0.5 * (E + E.T)
    w, V = np.linalg.eigh(Es)
    Lref = (V * np.array([weight(e, tol) for e in w])) @ V.T
    np.testing.assert_allclose(O.L6_to_mat(L6), Lref, rtol=1e-10, atol=1e-10)
    assert me == pytest.approx(w.min(), abs=1e-12)


def test_spd_inverse3_is_inverse_when_resolved():
    rng = np.random.default_rng(7)
    for _ in range(20):
        E, ev = random_spd(rng, 0.1, 2.0)          # every eigenvalue >= 2 tol
        L6, _ = O.spd_inverse3(E, 0.02)
        np.testing.assert_allclose(O.L6_to_mat(L6) @ E, np.eye(3), atol=1e-11)


def test_spd_inverse3_degenerate():
    # triple-degenerate E = lambda I (the interior) and a rank-deficient E
    L6, me = O.spd_inverse3(0.997 * np.eye(3), 0.02)
    np.testing.assert_allclose(O.L6_to_mat(L6), np.eye(3) / 0.997, atol=1e-14)
    assert me == pytest.approx(0.997)
    E = np.diag([0.8, 0.8, 0.0])
    L6, me = O.spd_inverse3(E, 0.02)
    np.testing.assert_allclose(O.L6_to_mat(L6), np.diag([1 / 0.8, 1 / 0.8, 1.0]), atol=1e-14)
    assert me == 0.0


@pytest.mark.parametrize("seed", range(10))
def test_spd_inverse2_matches_eigh(seed):
    rng = np.random.default_rng(100 + seed)
    tol = 0.02
    a = rng.normal(size=(2, 2))
    E2 = a @ a.T * rng.uniform(0.01, 1.0)
    L6 = O.spd_inverse2(E2, tol)
    w, V = np.linalg.eigh(E2)
    Lref = (V * np.array([weight(e, tol) for e in w])) @ V.T
    L = O.L6_to_mat(L6)
    np.testing.assert_allclose(L[:2, :2], Lref, rtol=1e-9, atol=1e-12)
    assert L[2, 2] == 1.0 and L[0, 2] == 0.0 and L[1, 2] == 0.0


def test_lambda_spectrum_plane_strain_lattice():
    """The reference's measured E spectrum on a plane-strain lattice at support 3 dx0:
    interior 0.997, flat free edge 0.499, convex corner 0.254 (SOLID.py:350-353)."""
    sc = free_block(n=(20, 20, 1), dx=0.01, jitter=0.0, two_d="planeStrain", pst=False, delta=0.0)
    lay, fs, P = layout(sc)
    s = oracle_state(lay, fs)
    O.fluid_step(lay.dims, P, s)
    x = s.x
    xi = np.round((x[:, 0] - x[:, 0].min()) / 0.01).astype(int)
    yi = np.round((x[:, 1] - x[:, 1].min()) / 0.01).astype(int)
    lam = s.lam
    assert round(float(lam[(xi == 10) & (yi == 10)][0]), 3) == 0.997
    assert round(float(lam[(xi == 10) & (yi == 0)][0]), 3) == 0.499
    assert round(float(lam[(xi == 0) & (yi == 0)][0]), 3) == 0.254
