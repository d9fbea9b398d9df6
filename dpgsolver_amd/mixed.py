"""Element operators of the batching form: HEX, WEDGE, TET and PYR volumes of any order p <= 5.

The reference builds every element's operators from an orthonormal basis on the reference element and
cubature rules (src/element/bases.c, nodes_operators.c, element_operators.c; WEDGE as TRI x LINE,
element_operators_tp.c:90-241), then uses them in the same residual for every element type
(compute_volume_rlhs_T.c:145-177, compute_face_rlhs_T.c:297-359).  Here the solution is represented in
the orthonormal (modal) basis of the element's polynomial space itself (the same space and the same
Galerkin residual as a Lagrange basis over it; the mass matrix of an affine element is then J times the
identity), with collapsed-coordinate Gauss-Jacobi cubature (Karniadakis & Sherwin, Spectral/hp Element
Methods, 2nd ed., sec. 4.1-4.2 and appendix B):

  element  space                          basis (collapsed coordinates a, b, c)                   Ns
  HEX      Q_p                            L_i(r) L_j(s) L_k(t)                                    (p+1)^3
  WEDGE    P_p(TRI) x P_p(LINE)           P_i(a) P_j^(2i+1,0)(s) (1-s)^i L_k(t)                   (p+1)^2(p+2)/2
  TET      P_p                            P_i(a) P_j^(2i+1,0)(b)(1-b)^i P_k^(2i+2j+2,0)(c)(1-c)^(i+j)  (p+1)(p+2)(p+3)/6
  PYR      pyramid space (Chan & Warburton 2016: i, j <= p, k <= p - max(i, j))
                                          P_i(a) P_j(b) (1-c)^m P_k^(2m+2,0)(c), m = max(i, j)    (p+1)(p+2)(2p+3)/6

The formulas fix the space; each basis is then made orthonormal under its volume rule by a Cholesky
factor of its Gram matrix (exact: the rules integrate every product of two basis functions).
Volume rules use p+1 points per collapsed direction (Gauss-Legendre, Gauss-Jacobi (1,0) and (2,0)),
exact for degree 2p+1; face rules use q+1 points per direction of the face with q = max(p_L, p_R).

Reference elements (vertex lists, the local faces as vertex lists in the order that fixes their face
coordinates):
  HEX   [-1,1]^3, TP vertex order, faces r-, r+, s-, s+, t-, t+
  WEDGE TRI (-1,-1),(1,-1),(-1,1) x [-1,1]; faces: 3 QUAD sides then the TRI bottom and top
  TET   (-1,-1,-1),(1,-1,-1),(-1,1,-1),(-1,-1,1)
  PYR   base [-1,1]^2 at t = -1, apex (0,0,1); faces: the QUAD base then 4 TRI sides
"""
import math
from functools import lru_cache

import numpy as np

HEX, WEDGE, TET, PYR = 0, 1, 2, 3
TYPE_NAMES = {HEX: "HEX", WEDGE: "WEDGE", TET: "TET", PYR: "PYR"}
F_TRI, F_QUAD = 0, 1
P_MAX = 5

VERTS = {
    HEX: np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [1, 1, -1],
                   [-1, -1, 1], [1, -1, 1], [-1, 1, 1], [1, 1, 1]], dtype=float),
    WEDGE: np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [-1, 1, 1]], dtype=float),
    TET: np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=float),
    PYR: np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [1, 1, -1], [0, 0, 1]], dtype=float),
}
# local faces as vertex lists: TRI faces (v0, v1, v2), QUAD faces (v0, v1, v2, v3) in tensor-product order
FACES = {
    HEX: [(0, 2, 4, 6), (1, 3, 5, 7), (0, 1, 4, 5), (2, 3, 6, 7), (0, 1, 2, 3), (4, 5, 6, 7)],
    WEDGE: [(1, 2, 4, 5), (0, 2, 3, 5), (0, 1, 3, 4), (0, 1, 2), (3, 4, 5)],
    TET: [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)],
    PYR: [(0, 1, 2, 3), (0, 1, 4), (2, 3, 4), (0, 2, 4), (1, 3, 4)],
}
REF_FACE_AREA = {F_TRI: 2.0, F_QUAD: 4.0}


def face_type(nv):
    return F_TRI if nv == 3 else F_QUAD


# -------------------------------------------------------------------------------------------------------
# 1-D polynomials and rules
# -------------------------------------------------------------------------------------------------------

def jacobi(n, a, b, x):
    """P_n^(a,b)(x) by the three-term recurrence (any real or complex x array)."""
    x = np.asarray(x)
    p0 = np.ones_like(x)
    if n == 0:
        return p0
    p1 = 0.5 * (a - b + (a + b + 2) * x)
    for k in range(2, n + 1):
        c1 = 2 * k * (k + a + b) * (2 * k + a + b - 2)
        c2 = (2 * k + a + b - 1) * (a * a - b * b)
        c3 = (2 * k + a + b - 2) * (2 * k + a + b - 1) * (2 * k + a + b)
        c4 = 2 * (k + a - 1) * (k + b - 1) * (2 * k + a + b)
        p0, p1 = p1, ((c2 + c3 * x) * p1 - c4 * p0) / c1
    return p1


@lru_cache(maxsize=None)
def gauss_jacobi(n, a, b):
    """n-point Gauss-Jacobi rule for the weight (1-x)^a (1+x)^b on [-1,1] (Golub-Welsch)."""
    k = np.arange(n, dtype=float)
    ab = a + b
    diag = np.empty(n)
    for i in range(n):
        den = (2 * i + ab) * (2 * i + ab + 2)
        diag[i] = (b * b - a * a) / den if den != 0 else (b - a) / (ab + 2)
    off = np.empty(max(n - 1, 0))
    for i in range(1, n):
        num = 4 * i * (i + a) * (i + b) * (i + ab)
        den = (2 * i + ab) ** 2 * (2 * i + ab + 1) * (2 * i + ab - 1)
        off[i - 1] = math.sqrt(num / den)
    T = np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)
    x, V = np.linalg.eigh(T)
    mu0 = 2 ** (ab + 1) * math.gamma(a + 1) * math.gamma(b + 1) / math.gamma(ab + 2)
    w = mu0 * V[0, :] ** 2
    del k
    return x, w


# -------------------------------------------------------------------------------------------------------
# cubature
# -------------------------------------------------------------------------------------------------------

@lru_cache(maxsize=None)
def volume_rule(etype, n):
    """Points (Nc, 3) and weights of the n-per-direction collapsed rule of the reference element."""
    xg, wg = gauss_jacobi(n, 0.0, 0.0)
    x1, w1 = gauss_jacobi(n, 1.0, 0.0)
    x2, w2 = gauss_jacobi(n, 2.0, 0.0)
    pts, wts = [], []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                if etype == HEX:
                    pts.append((xg[i], xg[j], xg[k]))
                    wts.append(wg[i] * wg[j] * wg[k])
                elif etype == WEDGE:
                    a, b, t = xg[i], x1[j], xg[k]
                    pts.append(((1 + a) * (1 - b) / 2 - 1, b, t))
                    wts.append(wg[i] * w1[j] * wg[k] / 2)
                elif etype == TET:
                    a, b, c = xg[i], x1[j], x2[k]
                    pts.append(((1 + a) * (1 - b) * (1 - c) / 4 - 1, (1 + b) * (1 - c) / 2 - 1, c))
                    wts.append(wg[i] * w1[j] * w2[k] / 8)
                else:
                    a, b, c = xg[i], xg[j], x2[k]
                    pts.append((a * (1 - c) / 2, b * (1 - c) / 2, c))
                    wts.append(wg[i] * wg[j] * w2[k] / 4)
    return np.array(pts), np.array(wts)


@lru_cache(maxsize=None)
def face_rule(ftype, n):
    """Points (Nf, 2) on the reference TRI (-1,-1),(1,-1),(-1,1) or QUAD [-1,1]^2, and weights."""
    xg, wg = gauss_jacobi(n, 0.0, 0.0)
    x1, w1 = gauss_jacobi(n, 1.0, 0.0)
    pts, wts = [], []
    for j in range(n):
        for i in range(n):
            if ftype == F_QUAD:
                pts.append((xg[i], xg[j]))
                wts.append(wg[i] * wg[j])
            else:
                a, b = xg[i], x1[j]
                pts.append(((1 + a) * (1 - b) / 2 - 1, b))
                wts.append(wg[i] * w1[j] / 2)
    return np.array(pts), np.array(wts)


# -------------------------------------------------------------------------------------------------------
# bases
# -------------------------------------------------------------------------------------------------------

def _leg(n, x):
    return jacobi(n, 0.0, 0.0, x)


def _collapse(etype, r, s, t):
    """Collapsed coordinates, with the degenerate denominators guarded (their limits multiply by 0)."""
    if etype == TET:
        den = -s - t
        a = np.where(np.abs(den) > 0, 2 * (1 + r) / np.where(np.abs(den) > 0, den, 1) - 1, -1)
        den2 = 1 - t
        b = np.where(np.abs(den2) > 0, 2 * (1 + s) / np.where(np.abs(den2) > 0, den2, 1) - 1, -1)
        return a, b, t
    if etype == PYR:
        den = 1 - t
        ok = np.abs(den) > 0
        a = np.where(ok, 2 * r / np.where(ok, den, 1), 0)
        b = np.where(ok, 2 * s / np.where(ok, den, 1), 0)
        return a, b, t
    if etype == WEDGE:
        den = 1 - s
        ok = np.abs(den) > 0
        a = np.where(ok, 2 * (1 + r) / np.where(ok, den, 1) - 1, -1)
        return a, s, t
    return r, s, t


def mode_list(etype, p):
    if etype == HEX:
        return [(i, j, k) for k in range(p + 1) for j in range(p + 1) for i in range(p + 1)]
    if etype == WEDGE:
        return [(i, j, k) for k in range(p + 1) for i in range(p + 1) for j in range(p + 1 - i)]
    if etype == TET:
        return [(i, j, k) for i in range(p + 1) for j in range(p + 1 - i) for k in range(p + 1 - i - j)]
    return [(i, j, k) for i in range(p + 1) for j in range(p + 1) for k in range(p + 1 - max(i, j))]


def n_modes(etype, p):
    return len(mode_list(etype, p))


def raw_basis(etype, p, x):
    """Un-normalized basis (N, Ns) at points x (N, 3), real or complex."""
    r, s, t = x[:, 0], x[:, 1], x[:, 2]
    a, b, c = _collapse(etype, r, s, t)
    cols = []
    for (i, j, k) in mode_list(etype, p):
        if etype == HEX:
            cols.append(_leg(i, r) * _leg(j, s) * _leg(k, t))
        elif etype == WEDGE:
            cols.append(_leg(i, a) * jacobi(j, 2 * i + 1.0, 0.0, s) * ((1 - s) / 2) ** i * _leg(k, t))
        elif etype == TET:
            cols.append(_leg(i, a) * jacobi(j, 2 * i + 1.0, 0.0, b) * ((1 - b) / 2) ** i
                        * jacobi(k, 2 * i + 2 * j + 2.0, 0.0, c) * ((1 - c) / 2) ** (i + j))
        else:
            m = max(i, j)
            cols.append(_leg(i, a) * _leg(j, b) * ((1 - c) / 2) ** m * jacobi(k, 2 * m + 2.0, 0.0, c))
    return np.stack(cols, axis=1)


@lru_cache(maxsize=None)
def _orth(etype, p):
    """The upper factor R with raw_basis @ inv(R) orthonormal under the element's volume rule."""
    x, w = volume_rule(etype, p + 1)
    V = raw_basis(etype, p, x)
    G = V.T @ (w[:, None] * V)
    L = np.linalg.cholesky(G)
    return np.linalg.inv(L.T)


def basis(etype, p, x):
    """Orthonormal basis (N, Ns) at points x (N, 3)."""
    return raw_basis(etype, p, np.asarray(x, dtype=float)) @ _orth(etype, p)


def basis_grad(etype, p, x):
    """d basis / d(r, s, t): (3, N, Ns), by complex-step differentiation (exact to round-off)."""
    x = np.asarray(x, dtype=float)
    h = 1e-30
    R = _orth(etype, p)
    out = np.empty((3, x.shape[0], R.shape[1]))
    for d in range(3):
        xc = x.astype(complex)
        xc[:, d] += 1j * h
        out[d] = (raw_basis(etype, p, xc) @ R).imag / h
    return out


# -------------------------------------------------------------------------------------------------------
# faces
# -------------------------------------------------------------------------------------------------------

def face_map(etype, lf, perm, pts2):
    """Reference-element coordinates (N, 3) of face points given in face coordinates pts2 (N, 2) of a face
    whose vertices are listed as FACES[etype][lf] permuted by `perm` (the order the face's owner lists
    them: face vertex k of the owner's list is vertex FACES[etype][lf][perm[k]])."""
    fv = FACES[etype][lf]
    V = VERTS[etype][[fv[q] for q in perm]]
    xi, eta = pts2[:, 0], pts2[:, 1]
    if len(fv) == 3:
        return V[0] + np.outer((xi + 1) / 2, V[1] - V[0]) + np.outer((eta + 1) / 2, V[2] - V[0])
    N = [(1 - xi) * (1 - eta) / 4, (1 + xi) * (1 - eta) / 4, (1 - xi) * (1 + eta) / 4, (1 + xi) * (1 + eta) / 4]
    return sum(np.outer(N[q], V[q]) for q in range(4))


# vertex permutations of a face as seen by its right volume: case k maps the left's face vertex i to the
# right's face vertex perm[i]
TRI_PERMS = [(0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (1, 0, 2)]
QUAD_PERMS = [(0, 1, 2, 3), (1, 0, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0), (0, 2, 1, 3), (2, 0, 3, 1), (1, 3, 0, 2),
              (3, 1, 2, 0)]


def perms_of(ftype):
    return TRI_PERMS if ftype == F_TRI else QUAD_PERMS


@lru_cache(maxsize=None)
def face_operator(etype, p, lf, case, q):
    """Basis values (Nf, Ns) of the element at the points of the face rule of order q, the face listed by
    the face's left volume; `case` is the correspondence case of this element's vertex list (0 for the
    left volume itself)."""
    ft = face_type(len(FACES[etype][lf]))
    pts2, _ = face_rule(ft, q + 1)
    perm = perms_of(ft)[case]
    return basis(etype, p, face_map(etype, lf, perm, pts2))


@lru_cache(maxsize=None)
def volume_operators(etype, p):
    """(B (Nc, Ns) basis values, G (3, Nc, Ns) reference gradients, w (Nc) weights) of the volume rule."""
    x, w = volume_rule(etype, p + 1)
    return basis(etype, p, x), basis_grad(etype, p, x), w


# -------------------------------------------------------------------------------------------------------
# tensor-product factors (the kernels apply HEX = LINE x LINE x LINE and WEDGE = TRI x LINE by sum
# factorization; element_operators_tp.c:90-241 builds the reference's TP operators the same way)
# -------------------------------------------------------------------------------------------------------

@lru_cache(maxsize=None)
def line_operators(p):
    """(B1 [n][n] orthonormal Legendre at the n = p+1 GL points, D1 [n][n] their derivatives, w1 [n])."""
    n = p + 1
    x, w = gauss_jacobi(n, 0.0, 0.0)
    B = np.empty((n, n))
    D = np.empty((n, n))
    h = 1e-30
    for k in range(n):
        c = math.sqrt((2 * k + 1) / 2.0)
        B[:, k] = c * _leg(k, x)
        D[:, k] = (c * _leg(k, x.astype(complex) + 1j * h)).imag / h
    return B, D, w


def _tri_raw(p, x):
    r, s = x[:, 0], x[:, 1]
    den = 1 - s
    ok = np.abs(den) > 0
    a = np.where(ok, 2 * (1 + r) / np.where(ok, den, 1) - 1, -1)
    cols = []
    for i in range(p + 1):
        for j in range(p + 1 - i):
            cols.append(_leg(i, a) * jacobi(j, 2 * i + 1.0, 0.0, s) * ((1 - s) / 2) ** i)
    return np.stack(cols, axis=1)


@lru_cache(maxsize=None)
def tri_operators(p):
    """(B [Nq][Nt] orthonormal TRI basis at the (p+1)^2 collapsed points of face_rule(F_TRI, p+1),
    Gr, Gs [Nq][Nt] its r and s derivatives, w [Nq]); modes in the WEDGE's TRI order."""
    x, w = face_rule(F_TRI, p + 1)
    V = _tri_raw(p, x)
    R = np.linalg.inv(np.linalg.cholesky(V.T @ (w[:, None] * V)).T)
    h = 1e-30
    G = []
    for d in range(2):
        xc = x.astype(complex)
        xc[:, d] += 1j * h
        G.append((_tri_raw(p, xc) @ R).imag / h)
    return V @ R, G[0], G[1], w


@lru_cache(maxsize=None)
def line_traces(p):
    """(L_i(-1), L_i(+1)) of the orthonormal Legendre basis of order p."""
    out = []
    for x in (-1.0, 1.0):
        out.append(np.array([math.sqrt((2 * k + 1) / 2.0) * _leg(k, np.array([x]))[0] for k in range(p + 1)]))
    return out[0], out[1]


def hex_face_modal(p, lf, Bf):
    """The face-modal rows BF [Nq][n^2] of a HEX face operator Bf [Nq][n^3] on local face lf:
    Bf[q][(i, a, b)] = tau_i BF[q][a + n b] with tau the 1-D basis traced to the face's side along its
    normal direction (a, b: the other two directions, lower fastest).  Checks the factorization."""
    n = p + 1
    dn, side = lf >> 1, lf & 1
    tau = line_traces(p)[side]
    B4 = Bf.reshape(-1, n, n, n)                      # [q][k][j][i]
    axis = {0: 3, 1: 2, 2: 1}[dn]
    BF = np.take(B4, 0, axis=axis) / tau[0]           # [q][b][a]: the remaining axes, higher first
    rebuilt = np.expand_dims(BF, axis) * tau.reshape([-1 if ax == axis else 1 for ax in range(4)])
    assert np.abs(rebuilt - B4).max() < 1e-12, "face operator is not separable along the face normal"
    return BF.reshape(BF.shape[0], n * n)


# -------------------------------------------------------------------------------------------------------
# collapsed-coordinate factors of TET and PYR (Sherwin & Karniadakis sec. 4.1-4.2): the raw bases are
# orthogonal, so the orthonormal basis is the raw one times a per-mode scale s, and the raw basis is a
# product of one function of each collapsed coordinate:
#   TET  phi_ijk = s_ijk P_i(a) [P_j^(2i+1,0)(b) ((1-b)/2)^i] [P_k^(2i+2j+2,0)(c) ((1-c)/2)^(i+j)]
#   PYR  phi_ijk = s_ijk P_i(a) P_j(b) [P_k^(2m+2,0)(c) ((1-c)/2)^m],  m = max(i, j)
# and the reference gradient follows from the collapse's chain rule at each node.
# -------------------------------------------------------------------------------------------------------

def _cplx_d(f, x):
    h = 1e-30
    return (f(x.astype(complex) + 1j * h)).imag / h


@lru_cache(maxsize=None)
def collapsed_factors(etype, p):
    """Factor tables of TET / PYR order p at the n = p+1 point rule of each collapsed direction:
    A, Ad [n][n] (P_i(a_q), its derivative), B, Bd (TET: [i][q][j] with j <= p-i, zero-padded;
    PYR: [q][j]), C, Cd (TET: [i][j][r][k], k <= p-i-j; PYR: [m][r][k], k <= p-m), the collapsed
    coordinates xa, xb, xc [n], the node weights w [n^3] (volume_rule order, a fastest) and the mode scale
    s [Ns] (mode_list order)."""
    assert etype in (TET, PYR)
    n = p + 1
    xa, _ = gauss_jacobi(n, 0.0, 0.0)
    if etype == TET:
        xb, _ = gauss_jacobi(n, 1.0, 0.0)
    else:
        xb, _ = gauss_jacobi(n, 0.0, 0.0)
    xc, _ = gauss_jacobi(n, 2.0, 0.0)
    A = np.stack([_leg(i, xa) for i in range(n)], axis=1)
    Ad = np.stack([_cplx_d(lambda x, i=i: _leg(i, x), xa) for i in range(n)], axis=1)
    if etype == TET:
        fb = lambda i, j: (lambda x: jacobi(j, 2 * i + 1.0, 0.0, x) * ((1 - x) / 2) ** i)
        B = np.zeros((n, n, n))
        Bd = np.zeros((n, n, n))
        for i in range(n):
            for j in range(n - i):
                B[i, :, j] = fb(i, j)(xb)
                Bd[i, :, j] = _cplx_d(fb(i, j), xb)
        fc = lambda i, j, k: (lambda x: jacobi(k, 2 * i + 2 * j + 2.0, 0.0, x) * ((1 - x) / 2) ** (i + j))
        C = np.zeros((n, n, n, n))
        Cd = np.zeros((n, n, n, n))
        for i in range(n):
            for j in range(n - i):
                for k in range(n - i - j):
                    C[i, j, :, k] = fc(i, j, k)(xc)
                    Cd[i, j, :, k] = _cplx_d(fc(i, j, k), xc)
    else:
        B = np.stack([_leg(j, xb) for j in range(n)], axis=1)
        Bd = np.stack([_cplx_d(lambda x, j=j: _leg(j, x), xb) for j in range(n)], axis=1)
        fc = lambda m, k: (lambda x: ((1 - x) / 2) ** m * jacobi(k, 2 * m + 2.0, 0.0, x))
        C = np.zeros((n, n, n))
        Cd = np.zeros((n, n, n))
        for m in range(n):
            for k in range(n - m):
                C[m, :, k] = fc(m, k)(xc)
                Cd[m, :, k] = _cplx_d(fc(m, k), xc)
    _, w = volume_rule(etype, n)
    s = np.diag(_orth(etype, p)).copy()
    return A, Ad, B, Bd, C, Cd, xa, xb, xc, w, s


def collapsed_tables(etype, p):
    """The factor block of a TET / PYR group as the sum-factorized kernels read it: collapsed_factors
    flattened in order, TET's C and Cd kept only at its modes, row (mode (i, j, k) in mode_list order)
    times n + r."""
    A, Ad, B, Bd, C, Cd, xa, xb, xc, w, s = collapsed_factors(etype, p)
    if etype == TET:
        modes = mode_list(etype, p)
        C = np.array([C[i, j, :, k] for (i, j, k) in modes])
        Cd = np.array([Cd[i, j, :, k] for (i, j, k) in modes])
    return np.concatenate([x.ravel() for x in (A, Ad, B, Bd, C, Cd, xa, xb, xc, w, s)])


def collapsed_dense(etype, p):
    """The dense volume operators (B [Nc][Ns], G [3][Nc][Ns]) rebuilt from collapsed_factors by the
    chain rule of the collapse - the kernels' sum-factorized TET / PYR paths apply exactly this."""
    A, Ad, Bf, Bfd, Cf, Cfd, xa, xb, xc, w, s = collapsed_factors(etype, p)
    n = p + 1
    modes = mode_list(etype, p)
    Bm = np.zeros((n ** 3, len(modes)))
    Ga = np.zeros_like(Bm)
    Gb = np.zeros_like(Bm)
    Gc = np.zeros_like(Bm)
    for col, (i, j, k) in enumerate(modes):
        for r in range(n):
            for q in range(n):
                for pa in range(n):
                    row = pa + n * q + n * n * r
                    if etype == TET:
                        fa, fad = A[pa, i], Ad[pa, i]
                        fbv, fbd = Bf[i, q, j], Bfd[i, q, j]
                        fcv, fcd = Cf[i, j, r, k], Cfd[i, j, r, k]
                    else:
                        m = max(i, j)
                        fa, fad = A[pa, i], Ad[pa, i]
                        fbv, fbd = Bf[q, j], Bfd[q, j]
                        fcv, fcd = Cf[m, r, k], Cfd[m, r, k]
                    Bm[row, col] = s[col] * fa * fbv * fcv
                    Ga[row, col] = s[col] * fad * fbv * fcv
                    Gb[row, col] = s[col] * fa * fbd * fcv
                    Gc[row, col] = s[col] * fa * fbv * fcd
    G = np.zeros((3, n ** 3, len(modes)))
    a = np.tile(xa, n * n)
    b = np.tile(np.repeat(xb, n), n)
    c = np.repeat(xc, n * n)
    for row in range(n ** 3):
        da, db = collapse_chain(etype, a[row], b[row], c[row])
        for d in range(3):
            G[d, row] = da[d] * Ga[row] + db[d] * Gb[row] + (1.0 if d == 2 else 0.0) * Gc[row]
    return Bm, G


def collapse_chain(etype, a, b, c):
    """(d a / d(r, s, t), d b / d(r, s, t)) of the collapse at a node (d c / d t = 1, d c / d r = d c / d s
    = 0).  TET: r = (1+a)(1-b)(1-c)/4 - 1, s = (1+b)(1-c)/2 - 1, t = c;  PYR: r = a(1-c)/2, s = b(1-c)/2,
    t = c."""
    if etype == TET:
        da = (4.0 / ((1 - b) * (1 - c)), 2.0 * (1 + a) / ((1 - b) * (1 - c)), 2.0 * (1 + a) / ((1 - b) * (1 - c)))
        db = (0.0, 2.0 / (1 - c), (1 + b) / (1 - c))
    else:
        da = (2.0 / (1 - c), 0.0, a / (1 - c))
        db = (0.0, 2.0 / (1 - c), b / (1 - c))
    return da, db


def tri_basis(p, x):
    """The orthonormal TRI basis of tri_operators at points x (N, 2)."""
    xq, w = face_rule(F_TRI, p + 1)
    V = _tri_raw(p, xq)
    R = np.linalg.inv(np.linalg.cholesky(V.T @ (w[:, None] * V)).T)
    return _tri_raw(p, np.asarray(x, dtype=float)) @ R


def _wedge_edge(lf):
    """Endpoints A, B (r, s) of the TRI edge under WEDGE quad face lf, in the face's vertex order."""
    fv = FACES[WEDGE][lf]
    return VERTS[WEDGE][fv[0]][:2], VERTS[WEDGE][fv[1]][:2]


@lru_cache(maxsize=None)
def wedge_edge_matrix(p, lf):
    """E [n][Nt]: the TRI modes restricted to the edge of WEDGE quad face lf, in the orthonormal
    Legendre basis of the edge parameter xi in [-1, 1] (A -> B): psi_m(edge(xi)) = sum_a E[a][m] L_a(xi)."""
    n = p + 1
    A, B = _wedge_edge(lf)
    xg, wg = gauss_jacobi(n, 0.0, 0.0)
    pts = A[None, :] + np.outer((xg + 1) / 2, B - A)
    psi = tri_basis(p, pts)                                   # [g][m]
    L = np.stack([math.sqrt((2 * a + 1) / 2.0) * _leg(a, xg) for a in range(n)], axis=1)   # [g][a]
    return L.T @ (wg[:, None] * psi)


def wedge_face_modal(p, lf, case, q, Bf):
    """The face-modal rows of a WEDGE face operator Bf [Nq][Ns] (modes k Nt + m, TRI fastest):
    quad face (lf 0..2): BF [Nq][n^2], BF[q][a + n k] = L_a(xi_q) L_k(t_q), with
    Bf[q][(k, m)] = sum_a BF[q][a + n k] E[a][m] (E = wedge_edge_matrix);  TRI face (lf 3, 4: t = -1, +1):
    BF [Nq][Nt] = psi_m at the face nodes, with Bf[q][(k, m)] = tau_k BF[q][m], tau = L_k(-1) or L_k(+1).
    Checks the factorization."""
    n = p + 1
    nt = n * (n + 1) // 2
    ft = face_type(len(FACES[WEDGE][lf]))
    pts2, _ = face_rule(ft, q + 1)
    x = face_map(WEDGE, lf, perms_of(ft)[case], pts2)          # reference coordinates of the face nodes
    Lt = np.stack([math.sqrt((2 * k + 1) / 2.0) * _leg(k, x[:, 2]) for k in range(n)], axis=1)   # [q][k]
    B3 = Bf.reshape(-1, n, nt)                                 # [q][k][m]
    if lf < 3:
        A, Bv = _wedge_edge(lf)
        d = Bv - A
        xi = 2.0 * ((x[:, :2] - A) @ d) / (d @ d) - 1.0
        La = np.stack([math.sqrt((2 * a + 1) / 2.0) * _leg(a, xi) for a in range(n)], axis=1)   # [q][a]
        BF = (La[:, :, None] * Lt[:, None, :]).transpose(0, 2, 1).reshape(-1, n * n)        # [q][k][a] -> a + n k
        E = wedge_edge_matrix(p, lf)
        rebuilt = np.einsum("qka,am->qkm", BF.reshape(-1, n, n), E)
    else:
        BF = tri_basis(p, x[:, :2])
        tau = line_traces(p)[lf - 3]
        rebuilt = tau[None, :, None] * BF[:, None, :]
    assert np.abs(rebuilt - B3).max() < 1e-11, "WEDGE face operator does not factor"
    return BF
