"""Element set-up of the batching form (dpgsolver_amd.mixed): cubature, bases, face maps.

Pins:
- cubature: the collapsed Gauss-Jacobi rules integrate every monomial up to the strength of the
  reference's Witherden-Vincent rules held in its unit-test data (nodes_si.data: TRI strength 6, TET
  strength 4; nodes_pyr.data: PYR strength 4) exactly as those rules do, after the affine map from this
  build's reference element to the reference's (vertices taken from the same data files);
- bases: orthonormal under an independent, finer rule; the spaces have the expected dimensions and
  contain every polynomial of degree <= p (TET: exactly P_p);
- divergence identity on every reference element: sum_q w_q d phi_j/d xi_d equals the face integrals of
  phi_j n_d (catches a wrong collapse, face vertex list or normal - a pyramid collapsed towards a corner
  instead of its apex fails it by O(1));
- WEDGE = TRI x LINE with the TRI index fastest: the fixture's standard operator is kron(line, tri)
  (operators_tp.data wedge_*, test_unit_operators_tp.c) and the WEDGE basis at the WEDGE nodes has exactly
  that Kronecker structure;
- face maps: on a mixed mesh the face cubature nodes of both sides coincide in physical space for every
  face and correspondence case.
"""
import itertools
import json
import os

import numpy as np
import pytest

from dpgsolver_amd import mixed as M
from dpgsolver_amd import mixed_mesh as MM

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _monomials(dim, deg):
    return [e for e in itertools.product(range(deg + 1), repeat=dim) if sum(e) <= deg]


def _affine(src, dst):
    """x = c + A xi mapping the vertices src (k, dim) onto dst (k, dim)."""
    S = np.hstack([np.ones((src.shape[0], 1)), src])
    sol, *_ = np.linalg.lstsq(S, dst, rcond=None)
    assert np.abs(S @ sol - dst).max() < 1e-13
    return sol[0], sol[1:].T


def _rows_present(pts, rows):
    for r in rows:
        assert np.min(np.linalg.norm(pts - r, axis=1)) < 1e-12, r


@pytest.mark.parametrize("case", ["tri", "tet", "pyr"])
def test_cubature_exactness_vs_reference_wv(case):
    if case == "tri":
        fx = _fixture("nodes_si.json")
        wv, strength = fx["Nodes_d2_p6_WV"], 6
        verts = np.array([[-1.0, -1 / np.sqrt(3)], [1.0, -1 / np.sqrt(3)], [0.0, 2 / np.sqrt(3)]])
        _rows_present(np.array(fx["Nodes_d2_p3_EQ"]["rst"]), verts)
        mine = np.array([[-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]])
        n = strength // 2 + 1
        xi, w = M.face_rule(M.F_TRI, n)
    elif case == "tet":
        fx = _fixture("nodes_si.json")
        wv, strength = fx["Nodes_d3_p4_WV"], 4
        verts = np.array([[-1.0, -1 / np.sqrt(3), -1 / np.sqrt(6)], [1.0, -1 / np.sqrt(3), -1 / np.sqrt(6)],
                          [0.0, 2 / np.sqrt(3), -1 / np.sqrt(6)], [0.0, 0.0, 3 / np.sqrt(6)]])
        _rows_present(np.array(fx["Nodes_d3_p2_AO"]["rst"]), verts)
        mine = M.VERTS[M.TET]
        n = strength // 2 + 1
        xi, w = M.volume_rule(M.TET, n)
    else:
        fx = _fixture("nodes_pyr.json")
        wv, strength = fx["Nodes_d3_p4_WV"], 4
        zb, za = -np.sqrt(2) / 5, 4 * np.sqrt(2) / 5
        verts = np.array([[-1, -1, zb], [1, -1, zb], [-1, 1, zb], [1, 1, zb], [0, 0, za]], dtype=float)
        _rows_present(np.array(fx["Nodes_d3_p3_GLoL"]["rst"]), verts)
        mine = M.VERTS[M.PYR]
        n = strength // 2 + 1
        xi, w = M.volume_rule(M.PYR, n)
    c, A = _affine(mine, verts)
    x = c + xi @ A.T
    wm = w * abs(np.linalg.det(A))
    xr, wr = np.array(wv["rst"]), np.array(wv["w"])
    dim = xr.shape[1]
    for e in _monomials(dim, strength):
        f_ref = np.prod([xr[:, k] ** e[k] for k in range(dim)], axis=0)
        f_mine = np.prod([x[:, k] ** e[k] for k in range(dim)], axis=0)
        a, b = wr @ f_ref, wm @ f_mine
        assert abs(a - b) <= 1e-12 * max(1.0, abs(a)), (case, e, a, b)


@pytest.mark.parametrize("etype", [M.HEX, M.WEDGE, M.TET, M.PYR])
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_bases_orthonormal_and_spaces(etype, p):
    # orthonormal under a finer rule (other nodes than the rule that built the factor)
    xi, w = M.volume_rule(etype, p + 3)
    B = M.basis(etype, p, xi)
    assert np.abs(B.T @ (w[:, None] * B) - np.eye(B.shape[1])).max() < 1e-12
    # dimensions
    dims = {M.HEX: (p + 1) ** 3, M.WEDGE: (p + 1) ** 2 * (p + 2) // 2, M.TET: (p + 1) * (p + 2) * (p + 3) // 6,
            M.PYR: (p + 1) * (p + 2) * (2 * p + 3) // 6}
    assert B.shape[1] == dims[etype] == M.n_modes(etype, p)
    # every polynomial of degree <= p is in the space: projection reproduces it at other points
    rng = np.random.default_rng(p)
    bary = rng.dirichlet(np.ones(len(M.VERTS[etype])), size=40)
    xt = bary @ M.VERTS[etype]
    Bt = M.basis(etype, p, xt)
    for e in _monomials(3, p):
        f = np.prod([xi[:, k] ** e[k] for k in range(3)], axis=0)
        coef = B.T @ (w * f)
        ft = np.prod([xt[:, k] ** e[k] for k in range(3)], axis=0)
        assert np.abs(Bt @ coef - ft).max() < 1e-11, e


@pytest.mark.parametrize("etype", [M.HEX, M.WEDGE, M.TET, M.PYR])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_divergence_identity(etype, p):
    _, G, w = M.volume_operators(etype, p)
    lhs = np.einsum("dqj,q->dj", G, w)
    rhs = np.zeros_like(lhs)
    V = M.VERTS[etype]
    cen = V.mean(axis=0)
    for lf, fv in enumerate(M.FACES[etype]):
        ft = M.face_type(len(fv))
        _, wf = M.face_rule(ft, p + 1)
        X = V[list(fv)]
        nv = np.cross(X[1] - X[0], X[2] - X[0])
        area = np.linalg.norm(nv) * (0.5 if ft == M.F_TRI else 1.0)
        nv /= np.linalg.norm(nv)
        if np.dot(nv, X.mean(axis=0) - cen) < 0:
            nv = -nv
        Bf = M.face_operator(etype, p, lf, 0, p)
        rhs += np.outer(nv, (wf * area / M.REF_FACE_AREA[ft]) @ Bf)
    assert np.abs(lhs - rhs).max() < 1e-12


def test_wedge_kron_order():
    fx = _fixture("operators_tp.json")
    tri, _, line = (None if m is None else np.array(m) for m in fx["wedge_ops_tp"])
    std = np.array(fx["wedge_op_std"])
    # the reference's standard WEDGE operator: line index slowest, TRI index fastest
    assert np.abs(np.kron(line, tri) - std).max() < 1e-12
    assert np.abs(std @ np.array(fx["wedge_b"]) - np.array(fx["wedge_c"])).max() < 1e-8
    # the WEDGE basis at the WEDGE nodes is kron(line factor, tri factor) in the same order
    for p in (1, 2, 3):
        B, _, _ = M.volume_operators(M.WEDGE, p)
        n, nt = p + 1, (p + 1) * (p + 2) // 2
        R = B.reshape(n, n * n, n, nt).transpose(0, 2, 1, 3).reshape(n * n, n * n * nt)
        s = np.linalg.svd(R, compute_uv=False)
        assert s[1] <= 1e-12 * s[0], p


def test_face_nodes_coincide_on_mixed_mesh():
    mesh = MM.build_mixed_mesh((4, 3, 3), seed=5)
    cases = set()
    for f in range(mesh.n_face):
        pts = []
        for side in range(2):
            e = int(mesh.face_elem[f, side])
            t, lf = int(mesh.etype[e]), int(mesh.face_lf[f, side])
            case = 0 if side == 0 else int(mesh.face_case[f])
            ft = M.face_type(len(M.FACES[t][lf]))
            pts2, _ = M.face_rule(ft, 3)
            xi = M.face_map(t, lf, M.perms_of(ft)[case], pts2)
            x = mesh.c[e] + xi @ mesh.A[e].T
            pts.append(x)
        d = pts[0] - pts[1]
        d -= np.round(d)              # periodic images
        assert np.abs(d).max() < 1e-13, f
        cases.add((int(mesh.face_type[f]), int(mesh.face_case[f])))
    # the mesh exercises several correspondence cases of both face types
    assert len({c for t, c in cases if t == M.F_TRI}) >= 2
    assert len({c for t, c in cases if t == M.F_QUAD}) >= 2


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_tensor_product_factors(p):
    """The HEX and WEDGE volume operators are the Kronecker products of the factors the kernels apply by
    sum factorization (r fastest; WEDGE: TRI index fastest, LINE slowest)."""
    B1, D1, w1 = M.line_operators(p)
    B, G, w = M.volume_operators(M.HEX, p)
    assert np.abs(np.kron(np.kron(B1, B1), B1) - B).max() < 1e-12
    assert np.abs(np.kron(np.kron(B1, B1), D1) - G[0]).max() < 1e-10
    assert np.abs(np.kron(np.kron(B1, D1), B1) - G[1]).max() < 1e-10
    assert np.abs(np.kron(np.kron(D1, B1), B1) - G[2]).max() < 1e-10
    assert np.abs(np.kron(np.kron(w1, w1), w1) - w).max() < 1e-15
    Bt, Gr, Gs, wt = M.tri_operators(p)
    B, G, w = M.volume_operators(M.WEDGE, p)
    # the factor signs are fixed by the Cholesky normalization (positive leading coefficients)
    assert np.abs(np.kron(B1, Bt) - B).max() < 1e-12
    assert np.abs(np.kron(B1, Gr) - G[0]).max() < 1e-10
    assert np.abs(np.kron(B1, Gs) - G[1]).max() < 1e-10
    assert np.abs(np.kron(D1, Bt) - G[2]).max() < 1e-10
    assert np.abs(np.kron(w1, wt) - w).max() < 1e-15


@pytest.mark.parametrize("etype", [M.TET, M.PYR])
@pytest.mark.parametrize("p", [1, 2, 3, 5])
def test_collapsed_factors(etype, p):
    """The TET / PYR volume operators are the scaled products of one factor per collapsed coordinate,
    and their reference gradients follow from the collapse's chain rule: the tables the sum-factorized
    kernels contract rebuild the dense operators."""
    B, G, w = M.volume_operators(etype, p)
    Bc, Gc = M.collapsed_dense(etype, p)
    assert np.abs(Bc - B).max() < 1e-12
    assert np.abs(Gc - G).max() < 1e-10 * max(1.0, np.abs(G).max())
    assert np.abs(M.collapsed_factors(etype, p)[9] - w).max() < 1e-15


@pytest.mark.parametrize("p", [1, 3, 5])
def test_collapsed_tables_tet_compact(p):
    """The TET factor block keeps C, Cd only at the modes (row = mode in mode_list order, then r): the
    padded entries it drops are zero, and the block's length is the kernels' (2n^2 + 2n^3 + 2 Ns n + 3n +
    n^3 + Ns)."""
    n = p + 1
    A, Ad, B, Bd, C, Cd, xa, xb, xc, w, s = M.collapsed_factors(M.TET, p)
    tab = M.collapsed_tables(M.TET, p)
    modes = M.mode_list(M.TET, p)
    ns = len(modes)
    assert tab.size == 2 * n * n + 2 * n ** 3 + 2 * ns * n + 3 * n + n ** 3 + ns
    o = 2 * n * n + 2 * n ** 3
    Cc = tab[o:o + ns * n].reshape(ns, n)
    Cdc = tab[o + ns * n:o + 2 * ns * n].reshape(ns, n)
    full, fulld = np.zeros_like(C), np.zeros_like(Cd)
    for row, (i, j, k) in enumerate(modes):
        full[i, j, :, k] = Cc[row]
        fulld[i, j, :, k] = Cdc[row]
    assert np.array_equal(full, C) and np.array_equal(fulld, Cd)
    assert np.array_equal(tab[o + 2 * ns * n:], np.concatenate([xa, xb, xc, w.ravel(), s]))
