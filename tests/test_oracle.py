"""Pins of the CPU oracle against closed forms, independent formulas and
numerical libraries (CPU only)."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.geometry import blade_grid, rings_from_grids


def classical_vatistas(P, A, B, rc):
    """Independent form: (r1 x r2)(|r1|+|r2|) / (|r1||r2|(|r1||r2| + r1.r2))
    times the Vatistas n=2 factor h^2/sqrt(rc^4 + h^4), h = |r1 x r2|/|r0|."""
    r1, r2 = P - A, P - B
    c = np.cross(r1, r2)
    n1, n2 = np.linalg.norm(r1), np.linalg.norm(r2)
    r0 = np.linalg.norm(B - A)
    h2 = np.dot(c, c) / r0**2
    k = h2 / np.sqrt(rc**4 + h2**2)
    return c * (n1 + n2) / (n1 * n2 * (n1 * n2 + np.dot(r1, r2))) / (4 * np.pi) * k


def test_infinite_line_closed_form():
    L = 1e4
    for h, rc in [(0.1, 0.0), (0.1, 0.05), (1.0, 0.3), (0.02, 0.03)]:
        v = orc.vf_vind([0, h, 0], [-L, 0, 0], [L, 0, 0], rc)
        expect = h / (2 * np.pi * np.sqrt(rc**4 + h**4))
        assert v[0] == 0.0 and v[1] == 0.0
        assert v[2] == pytest.approx(expect, rel=1e-8)


def test_square_ring_centre_closed_form():
    for a, rc in [(1.0, 0.0), (2.0, 0.1), (0.5, 0.2)]:
        c = np.array([[-a / 2, -a / 2, 0], [a / 2, -a / 2, 0], [a / 2, a / 2, 0], [-a / 2, a / 2, 0]])
        v = sum(orc.vf_vind([0, 0, 0], c[k], c[(k + 1) % 4], rc) for k in range(4))
        per_side = np.sqrt(2) * a / (8 * np.pi * np.sqrt(rc**4 + a**4 / 16))
        assert v[2] == pytest.approx(4 * per_side, rel=1e-13)
        assert abs(v[0]) < 1e-15 and abs(v[1]) < 1e-15
    # rc = 0 reduces to 2 sqrt(2) / (pi a)
    a = 1.3
    c = np.array([[-a / 2, -a / 2, 0], [a / 2, -a / 2, 0], [a / 2, a / 2, 0], [-a / 2, a / 2, 0]])
    v = sum(orc.vf_vind([0, 0, 0], c[k], c[(k + 1) % 4], 0.0) for k in range(4))
    assert v[2] == pytest.approx(2 * np.sqrt(2) / (np.pi * a), rel=1e-13)


def test_guards_collinear_endpoint_and_antisymmetry():
    A, B = np.array([0.0, 0.0, 0.0]), np.array([1.0, 2.0, 3.0])
    assert np.all(orc.vf_vind(2.5 * B, A, B, 0.1) == 0.0)      # collinear beyond B
    assert np.all(orc.vf_vind(A, A, B, 0.1) == 0.0)            # on an end point
    assert np.all(orc.vf_vind(0.5 * B, A, B, 0.0) == 0.0)      # on the filament
    P = np.array([0.3, -0.7, 1.1])
    v1 = orc.vf_vind(P, A, B, 0.2)
    v2 = orc.vf_vind(P, B, A, 0.2)
    assert np.array_equal(v1, -v2)


def test_matches_independent_formula_random():
    rng = np.random.default_rng(7)
    for _ in range(200):
        A, B, P = rng.normal(size=(3, 3))
        rc = abs(rng.normal()) * 0.2
        v = orc.vf_vind(P, A, B, rc)
        w = classical_vatistas(P, A, B, rc)
        np.testing.assert_allclose(v, w, rtol=1e-10, atol=1e-13)


def test_rigid_motion_invariance():
    rng = np.random.default_rng(3)
    a, b = rng.normal(size=(50, 3)), rng.normal(size=(50, 3))
    g, rc = rng.normal(size=50), np.full(50, 0.05)
    p = rng.normal(size=(20, 3)) * 2
    v = orc.filament_sum(a, b, g, rc, p)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    t = rng.normal(size=3)
    w = orc.filament_sum(a @ q.T + t, b @ q.T + t, g, rc, p @ q.T + t)
    np.testing.assert_allclose(w, v @ q.T, rtol=1e-9, atol=1e-12)


def test_filament_sum_threads_identical():
    rng = np.random.default_rng(5)
    a, b = rng.normal(size=(300, 3)), rng.normal(size=(300, 3))
    g, rc = rng.normal(size=300), np.full(300, 0.01)
    p = rng.normal(size=(64, 3))
    assert np.array_equal(orc.filament_sum(a, b, g, rc, p), orc.filament_sum(a, b, g, rc, p, threads=4))


def test_lu_matches_lapack():
    from scipy.linalg import lu_factor
    rng = np.random.default_rng(11)
    for n in (1, 3, 17, 64):
        A = rng.normal(size=(n, n))
        LU, piv, info = orc.lu_factor(A)
        assert info == 0
        LU2, piv2 = lu_factor(A)
        assert np.array_equal(piv, piv2)
        np.testing.assert_allclose(LU, LU2, rtol=1e-12, atol=1e-12)
        b = rng.normal(size=n)
        np.testing.assert_allclose(orc.lu_solve(LU, piv, b), np.linalg.solve(A, b), rtol=1e-10)
    _, _, info = orc.lu_factor(np.zeros((3, 3)))
    assert info == 1


def test_inv2_3x3_reference_fixture():
    # tests/libMath_test.f90:64-78 of the reference: A and its inverse
    A = np.array([[1.0, 2.0, 3.0], [0.0, 1.0, 4.0], [5.0, 6.0, 0.0]], order="F")
    Ainv = np.array([[-24.0, 18.0, 5.0], [20.0, -15.0, -4.0], [-5.0, 4.0, 1.0]])
    LU, piv, info = orc.lu_factor(A)
    assert info == 0
    inv = np.column_stack([orc.lu_solve(LU, piv, e) for e in np.eye(3)])
    np.testing.assert_allclose(inv, Ainv, atol=1e-12)


MLP_NP = 45 * 64 + 64 + 2 * (64 * 64 + 64) + 3 * 64 + 3


def test_mlp_gradient_finite_difference():
    rng = np.random.default_rng(2)
    params = rng.normal(size=MLP_NP) * 0.2
    x = rng.normal(size=(6, 45))
    t = rng.normal(size=(6, 3))
    loss, g = orc.mlp_loss_grad(params, x, t)
    idx = rng.choice(MLP_NP, 60, replace=False)
    for k in idx:
        h = 1e-6
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (orc.mlp_loss_grad(pp, x, t)[0] - orc.mlp_loss_grad(pm, x, t)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9)
    y = orc.mlp_forward(params, x)
    assert loss == pytest.approx(0.5 * np.mean(np.sum((y - t) ** 2, axis=1)), rel=1e-12)


def test_ring_form_wing_sum_equals_filament_list():
    """or_rotor_vind(WING) is the sum over ring filaments with ring gam."""
    grids = [blade_grid(3, 5, seed=1)]
    corners = rings_from_grids(grids)
    gam = np.random.default_rng(1).normal(size=(1, 5, 3))
    rc = np.full((1, 5, 3, 4), 0.02)
    s = orc.System()
    ir = s.rotor_create(1, 3, 5, 4, 0, spanwiseCore=0.03, streamwiseCoreVec=np.full(6, 0.03))
    s.rotor_set_wing(ir, corners, rc, gam)
    p = np.random.default_rng(2).normal(size=(10, 3))
    v = s.rotor_induced_velocity(ir, 0, p)
    a = corners.reshape(-1, 4, 3)
    A = a.reshape(-1, 3)
    B = np.roll(a, -1, axis=1).reshape(-1, 3)
    G = np.repeat(gam.ravel(), 4)
    w = orc.filament_sum(A, B, G, np.full(len(G), 0.02), p)
    np.testing.assert_allclose(v, w, rtol=1e-12, atol=1e-14)
