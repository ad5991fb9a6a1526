"""AIC assembly, LU factorisation (single-workgroup and MFMA blocked paths)
and the refined circulation solve against the oracle and LAPACK."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import wake_harness as H
from tests.geometry import rings_from_grids

pytestmark = pytest.mark.gpu


def _ctx():
    from volcanor_amd import vlm
    return vlm.Context(0)


def _matrix(n, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((n, n)))


@pytest.mark.parametrize("n", [40, 256, 257, 777])
def test_lu_factor_matches_oracle(n):
    """n <= 256 runs the one-workgroup path, n > 256 the blocked path whose
    trailing update runs on MFMA; 257 and 777 leave ragged last panels."""
    A = _matrix(n, n)
    LU, piv = _ctx().lu_factor(A)
    LUo, pivo, info = orc.lu_factor(A)
    assert info == 0
    np.testing.assert_array_equal(piv, pivo)
    err = np.max(np.abs(LU - LUo)) / np.max(np.abs(LUo))
    assert err < 1e-11, err


@pytest.mark.parametrize("n", [2176, 2600])
def test_lu_factor_large_matches_lapack(n):
    """2176: the LU microbenchmark size (panels held in registers and LDS);
    2600: the first panels exceed on-chip capacity and take the tall-panel
    kernel."""
    from scipy.linalg import lu_factor
    A = _matrix(n, 5)
    LU, piv = _ctx().lu_factor(A)
    LUl, pivl = lu_factor(A)
    np.testing.assert_array_equal(piv, pivl)
    err = np.max(np.abs(LU - LUl)) / np.max(np.abs(LUl))
    assert err < 1e-10, err


@pytest.mark.parametrize("n", [1, 7, 40, 64, 65, 90, 96, 97, 100, 600])
def test_lu_solve_refine_reaches_fp64_residual(n):
    """n <= 64 runs the one-wave solve with a row per lane, 65 <= n <= 96
    the one-wave solve with two rows per lane (90: a three-bladed rotor of
    30 panels per blade), larger n the multi-launch path."""
    A = _matrix(n, 11 + n)
    b = np.random.default_rng(3).standard_normal(n)
    ctx = _ctx()
    LU, piv = ctx.lu_factor(A)
    x, iters = ctx.lu_solve_refine(A, LU, piv, b, max_iter=3, tol=1e-15)
    r = b - A @ x
    assert np.max(np.abs(r)) <= 1e-13 * np.max(np.abs(b)) * n
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
    assert 0 <= iters <= 3
    x0, it0 = ctx.lu_solve_refine(A, LU, piv, b, max_iter=0, tol=1e-15)
    assert it0 == 0
    np.testing.assert_allclose(x0, x, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("n", [33, 64, 65, 90, 96, 97, 130])
def test_triangular_solve_matches_oracle(n):
    """Without refinement the solve is the pivoted forward and back
    substitution of or_lu_solve, row by row in the same order, on every path
    (one and two rows per lane, multi-launch); the device contracts a*b + c
    into one fused operation, so the agreement is to rounding, not bitwise."""
    A = _matrix(n, 23 + n)
    b = np.random.default_rng(n).standard_normal(n)
    ctx = _ctx()
    LU, piv = ctx.lu_factor(A)
    x0, it0 = ctx.lu_solve_refine(A, LU, piv, b, max_iter=0, tol=1e-15)
    xo = orc.lu_solve(np.asfortranarray(LU), piv, b)
    assert it0 == 0
    np.testing.assert_allclose(x0, xo, rtol=1e-10, atol=1e-12 * np.max(np.abs(xo)))


@pytest.mark.parametrize("n,col", [(50, 16), (300, 100), (1300, 260)])
def test_lu_zero_pivot_is_an_error(n, col):
    """50: one-workgroup path; 300: the zero column in a 32-column panel held
    in registers (masking design); 1300: column 260 lies in the 16-column
    panel at 256, of 1044 rows (retiring design), where every candidate is 0
    and retired or empty slots must lose the tie to the live rows."""
    from volcanor_amd import vlm
    A = _matrix(n, 2)
    A[:, col] = 0.0
    with pytest.raises(vlm.VLMError):
        _ctx().lu_factor(A)


def test_rotor_aic_and_solve_match_oracle():
    case = H.Case(nb=2, nc=3, ns=8, nNwake=4, nFwake=0)
    gpu, cpu = _ctx(), orc.System()
    ig, ic = H.setup(gpu, case)[0], H.setup(cpu, case)[0]
    grids, _ = H.blade_grids(case, 0.0)
    corners = rings_from_grids(grids).reshape(-1, 4, 3)
    cp = corners.mean(axis=1)
    d1 = corners[:, 2] - corners[:, 0]
    d2 = corners[:, 3] - corners[:, 1]
    ncap = np.cross(d1, d2)
    ncap /= np.linalg.norm(ncap, axis=1)[:, None]
    n = len(cp)
    A = gpu.rotor_calc_aic(ig, cp, ncap, n)
    Ao = cpu.rotor_calc_aic(ic, cp, ncap, n)
    assert np.max(np.abs(A - Ao)) <= 1e-14 * np.max(np.abs(Ao))
    rhs = np.random.default_rng(9).standard_normal(n)
    g, _ = gpu.rotor_solve(ig, rhs)
    np.testing.assert_allclose(g, np.linalg.solve(Ao, rhs), rtol=1e-10, atol=1e-12)
