"""Pins of the oracle against the reference's own unit-test data and against closed forms (CPU).

Golden vectors (tests/golden/*.json, extracted by tests/golden/make_fixtures.py from the reference's
input/testing/unit/**.data files):
  nodes_tp.json      GL ("GLe") / GLL ("GLoL") tensor-product nodes and weights (test_unit_nodes)
  operators_tp.json  tensor-product operators: Kronecker ordering and application to b (test_unit_operators_tp.c:75-176)
  matrix.json        mm / mv / mm_diag (test_unit_containers)
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _tp_nodes(x1, w1, d):
    """Tensor-product nodes with r fastest, as the reference lists them."""
    n = x1.size
    pts, wts = [], []
    for k in range(n if d > 2 else 1):
        for j in range(n if d > 1 else 1):
            for i in range(n):
                idx = [i, j, k][:d]
                pts.append([x1[a] for a in idx])
                wts.append(np.prod([w1[a] for a in idx]))
    return np.array(pts), np.array(wts)


@pytest.mark.parametrize("key,fam", [("Nodes_d1_p3_GLe", 0), ("Nodes_d1_p3_GLoL", 1), ("Nodes_d2_p4_GLe", 0),
                                     ("Nodes_d2_p4_GLoL", 1), ("Nodes_d3_p3_GLe", 0), ("Nodes_d3_p3_GLoL", 1)])
def test_nodes_match_reference(key, fam):
    """orc_nodes_1d (Golub-Welsch) reproduces the reference's GL/GLL node and weight tables."""
    ref = _load("nodes_tp.json")[key]
    d = int(key.split("_")[1][1:])
    p = int(key.split("_")[2][1:])
    x, w = O.nodes_1d(fam, p)
    pts, wts = _tp_nodes(x, w, d)
    # the data file prints 16 decimals; the eigenvalue route is accurate to a few ulps
    assert np.allclose(pts, np.array(ref["rst"]), atol=1e-15, rtol=0)
    assert np.allclose(wts, np.array(ref["w"]), atol=1e-15, rtol=0)


@pytest.mark.parametrize("fam,p", [(0, p) for p in range(0, 8)] + [(1, p) for p in range(1, 8)])
def test_quadrature_exactness(fam, p):
    """GL with n = p+1 points integrates x^k exactly for k <= 2n-1, GLL for k <= 2n-3."""
    x, w = O.nodes_1d(fam, p)
    n = p + 1
    kmax = 2 * n - 1 if fam == 0 else 2 * n - 3
    for k in range(kmax + 1):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-14
    # one degree beyond is not exact: the rule is the Gauss(-Lobatto) one and not a wider one
    k = kmax + 1
    exact = 0.0 if k % 2 else 2.0 / (k + 1)
    assert abs(np.dot(w, x ** k) - exact) > 1e-6


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_lagrange_operators_closed_form(p):
    """orc_lagrange_1d: interpolation reproduces degree-p polynomials, derivative is exact, D 1 = 0."""
    xs, _ = O.nodes_1d(0, p)
    y = np.linspace(-1.0, 1.0, 7)
    V = O.lagrange_1d(xs, y, 0)
    Dm = O.lagrange_1d(xs, y, 1)
    for k in range(p + 1):
        assert np.allclose(V @ xs ** k, y ** k, atol=1e-13)
        dk = k * y ** (k - 1) if k > 0 else 0 * y
        assert np.allclose(Dm @ xs ** k, dk, atol=1e-12)
    assert np.allclose(Dm.sum(axis=1), 0.0, atol=1e-13)


@pytest.mark.parametrize("elem", ["quad", "hex", "wedge"])
def test_tensor_product_operators_match_reference(elem):
    """orc_kron reproduces op_std of the reference's tensor-product test, and op_std b = c."""
    data = _load("operators_tp.json")
    ops = [np.array(a) for a in data[f"{elem}_ops_tp"] if a is not None]
    op = O.kron(ops)
    ref_op = np.array(data[f"{elem}_op_std"])
    assert op.shape == ref_op.shape
    # op_std is printed with 4 significant digits in the data file
    assert np.allclose(op, ref_op, rtol=6e-4, atol=1e-6)
    b = np.array(data[f"{elem}_b"])
    c = np.array(data[f"{elem}_c"])
    assert np.allclose(op @ b, c, rtol=0, atol=5e-14 * max(1.0, np.abs(c).max()))


def test_dense_products_match_reference():
    """orc_mm against the reference's container unit test (mm, mv, mm_diag)."""
    m = _load("matrix.json")
    a, b, c = (np.array(m[k]) for k in ("a_Matrix", "b_Matrix", "c_Matrix"))
    assert np.allclose(O.mm(a, b), c, atol=1e-14)
    bv, cv = np.array(m["b_Vector"]), np.array(m["c_Vector"])
    assert np.allclose(O.mm(a, bv[:, None])[:, 0], cv, atol=1e-14)
    bl, br = np.array(m["bl_Vector"]), np.array(m["br_Vector"])
    assert np.allclose(O.mm(np.diag(bl), a), np.array(m["cl_Matrix"]), atol=1e-14)
    assert np.allclose(O.mm(a, np.diag(br)), np.array(m["cr_Matrix"]), atol=1e-14)


# --- pointwise physics ---------------------------------------------------------------------------------

def _state(d, rho, vel, p):
    u = np.empty(d + 2)
    u[0] = rho
    u[1:1 + d] = rho * np.asarray(vel[:d])
    u[d + 1] = p / 0.4 + 0.5 * rho * float(np.dot(vel[:d], vel[:d]))
    return u


@pytest.mark.parametrize("d", [2, 3])
def test_euler_flux_properties(d):
    """Flux of a state at rest is pure pressure; homogeneity F(a u) = a F(u); rotation covariance."""
    rng = np.random.default_rng(1)
    u0 = _state(d, 1.3, np.zeros(3), 2.1)
    f0 = O.flux_point(d, u0)
    expect = np.zeros((d, d + 2))
    for k in range(d):
        expect[k, 1 + k] = 2.1
    assert np.allclose(f0, expect, atol=1e-15)
    u = _state(d, 0.9, rng.normal(size=3), 1.7)
    assert np.allclose(O.flux_point(d, 2.5 * u), 2.5 * O.flux_point(d, u), rtol=1e-14, atol=1e-14)
    if d == 2:
        th = 0.37
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        ur = u.copy()
        ur[1:3] = R @ u[1:3]
        F, Fr = O.flux_point(2, u), O.flux_point(2, ur)
        # mass and energy fluxes rotate as vectors, momentum flux as a tensor
        for eq in (0, 3):
            assert np.allclose(Fr[:, eq], R @ F[:, eq], atol=1e-13)
        assert np.allclose(Fr[:, 1:3], R @ F[:, 1:3] @ R.T, atol=1e-13)


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("flux", [0, 1])
def test_numerical_flux_consistency(d, flux):
    """F*(u, u, n) = F(u) . n for LF and Roe-Pike (consistency)."""
    rng = np.random.default_rng(3 + d + flux)
    for _ in range(5):
        u = _state(d, rng.uniform(0.5, 2.0), rng.normal(size=3), rng.uniform(0.5, 3.0))
        n = rng.normal(size=d)
        n /= np.linalg.norm(n)
        nf = O.numflux_point(d, flux, u, u, n)
        assert np.allclose(nf, n @ O.flux_point(d, u), rtol=1e-13, atol=1e-13)


def test_lax_friedrichs_symmetry():
    """LF is conservative under exchange: F*(a, b, n) = -F*(b, a, -n)."""
    rng = np.random.default_rng(5)
    for d in (2, 3):
        a = _state(d, 1.1, rng.normal(size=3), 1.5)
        b = _state(d, 0.7, rng.normal(size=3), 2.5)
        n = rng.normal(size=d)
        n /= np.linalg.norm(n)
        assert np.allclose(O.numflux_point(d, 0, a, b, n), -O.numflux_point(d, 0, b, a, -n), atol=1e-13)


@pytest.mark.parametrize("d", [2, 3])
def test_roe_exact_for_stationary_contact(d):
    """Roe-Pike resolves a stationary contact exactly (zero mass flux, pressure flux), LF does not."""
    n = np.zeros(d)
    n[0] = 1.0
    vt = np.zeros(3)
    vt[1] = 0.4                                  # tangential velocity, zero normal velocity
    uL = _state(d, 1.0, vt, 1.0)
    uR = _state(d, 0.25, vt, 1.0)
    exact = np.zeros(d + 2)
    exact[1] = 1.0
    assert np.allclose(O.numflux_point(d, 1, uL, uR, n), exact, atol=1e-14)
    assert not np.allclose(O.numflux_point(d, 0, uL, uR, n), exact, atol=1e-3)


@pytest.mark.parametrize("d", [2, 3])
def test_roe_rotation_covariance(d):
    """Rotating both states and the normal rotates the momentum flux (checks index/sign of each
    velocity component in the Roe-Pike dissipation)."""
    rng = np.random.default_rng(11)
    uL = _state(d, 1.2, rng.normal(size=3), 1.9)
    uR = _state(d, 0.8, rng.normal(size=3), 1.1)
    n = rng.normal(size=d)
    n /= np.linalg.norm(n)
    A = rng.normal(size=(d, d))
    R, _ = np.linalg.qr(A)
    rot = lambda u: np.concatenate([[u[0]], R @ u[1:1 + d], [u[d + 1]]])
    f = O.numflux_point(d, 1, uL, uR, n)
    fr = O.numflux_point(d, 1, rot(uL), rot(uR), R @ n)
    assert np.allclose(fr[0], f[0], atol=1e-13)
    assert np.allclose(fr[1:1 + d], R @ f[1:1 + d], atol=1e-13)
    assert np.allclose(fr[d + 1], f[d + 1], atol=1e-13)
