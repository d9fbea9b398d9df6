"""Seeded periodic conforming mixed mesh of the unit cube (HEX, WEDGE, TET, PYR) for the batching form.

The unit cube is cut into nx x ny x nz boxes (periodic in x, y, z).  Every box face is a quad ('Q') or
split into two triangles along a diagonal ('S'; diagonal 0 joins the face's corners 00 and 11 in the
face's tensor-product order), and every box is filled according to its faces:
  * HEX box     all faces Q                                  -> 1 HEX
  * WEDGE box   z faces S (one diagonal per z column), x, y Q  -> 2 WEDGEs (split along the vertical diagonal plane)
  * TET box     all faces S with diagonal 0                  -> 6 TETs (the Kuhn subdivision along 000-111)
  * centre box  any faces                                    -> a PYR on every Q face, 2 TETs on every S face,
                                                                apex at the box centre
Boxes are arranged in x-layers of one kind: the TET layers form one block wrapped in a centre layer on
each side (a centre box turns towards the TET block with split faces and towards the other layers with
quads; its faces to other centre boxes are drawn at random), and HEX and WEDGE layers (whose shared faces
are quads) fill the rest in a seeded random order.  The layer counts follow the target element fractions
(default 40 % HEX, 30 % WEDGE, 20 % TET, 10 % PYR); every face is conforming by construction.  Per-element
orders are drawn uniformly from {1..p_max} (seed).  All elements are straight-sided (affine maps from
their reference elements, positively oriented).

Faces are matched through their sorted vertex ids (periodic ids wrap; at least 3 boxes per direction);
the first element met in element order is a face's left volume (the reference's face construction,
face.c:109-146), and `case` is the correspondence of the face's vertices as listed by the right volume
(the reference's neigh_info[1].ind_ord rule, set_ind_ord, face.c:496-552, for TRI and QUAD faces).
"""
from dataclasses import dataclass

import numpy as np

from . import mixed as M


@dataclass
class MixedMesh:
    etype: np.ndarray      # [E] int8 element type
    order: np.ndarray      # [E] int8 solution order p
    xv: list               # [E] (nv_e, 3) vertex coordinates (unwrapped, the element's own copy)
    A: np.ndarray          # [E][3][3] affine map d x / d xi
    c: np.ndarray          # [E][3] x = c + A xi
    jdet: np.ndarray       # [E] det A
    metrics: np.ndarray    # [E][3][3] m(i, j) = J d xi_j / d x_i (cofactor form)
    face_elem: np.ndarray  # [F][2]
    face_lf: np.ndarray    # [F][2]
    face_case: np.ndarray  # [F] right volume's correspondence case
    face_type: np.ndarray  # [F] F_TRI / F_QUAD
    normals: np.ndarray    # [F][3] unit normal, outward from the left volume
    face_jac: np.ndarray   # [F] physical face area / reference face area
    n_cubes: int

    @property
    def n_elem(self):
        return int(self.etype.size)

    @property
    def n_face(self):
        return int(self.face_elem.shape[0])

    def counts(self):
        return {M.TYPE_NAMES[t]: int((self.etype == t).sum()) for t in (M.HEX, M.WEDGE, M.TET, M.PYR)}


# cube corners: offset (a, b, c) -> local corner index a + 2 b + 4 c; faces of the cube in the order
# x-, x+, y-, y+, z-, z+, each as its 4 corners in tensor-product order of the face plane
_CUBE_FACES = [(0, 2, 4, 6), (1, 3, 5, 7), (0, 1, 4, 5), (2, 3, 6, 7), (0, 1, 2, 3), (4, 5, 6, 7)]
KIND_HEX, KIND_WEDGE, KIND_TET, KIND_CENTER = 0, 1, 2, 3
# the 6 tets of the Kuhn subdivision of a box along its 000-111 diagonal: corner paths 0 -> e_a -> e_a + e_b -> 7
_KUHN = [(0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7)]


def layer_counts(nx, fractions):
    """x-layer counts (HEX, WEDGE, TET, centre) for nx layers: two centre layers wrap the TET block; with
    q = 0 lateral splits a centre box holds 5 PYR + 2 TET, a TET box 6 TET, a WEDGE box 2, a HEX box 1."""
    f_hex, f_wedge, f_tet, f_pyr = fractions
    n_c = 2
    per = nx - n_c
    # elements per layer-box: HEX 1, WEDGE 2, TET 6; choose integer layers nearest to the targets
    best = None
    for n_t in range(1, per + 1):
        for n_w in range(0, per - n_t + 1):
            n_h = per - n_t - n_w
            if n_h < 0:
                continue
            e = {"h": n_h, "w": 2 * n_w, "t": 6 * n_t + 2 * n_c, "p": 5 * n_c}
            tot = sum(e.values())
            err = ((e["h"] / tot - f_hex) ** 2 + (e["w"] / tot - f_wedge) ** 2 + (e["t"] / tot - f_tet) ** 2
                   + (e["p"] / tot - f_pyr) ** 2)
            if best is None or err < best[0]:
                best = (err, n_h, n_w, n_t)
    return best[1], best[2], best[3], n_c


def build_mixed_mesh(n, seed=0, p_max=5, fractions=(0.4, 0.3, 0.2, 0.1), p_fixed=None):
    """The mixed mesh on n = (nx, ny, nz) boxes (nx layers along x, see layer_counts); p_fixed: one order
    for every element instead of random orders."""
    nx, ny, nz = n
    assert min(nx, ny, nz) >= 3
    rng = np.random.default_rng(seed)
    n_h, n_w, n_t, n_c = layer_counts(nx, fractions)
    others = [KIND_HEX] * n_h + [KIND_WEDGE] * n_w
    rng.shuffle(others)
    cut = int(rng.integers(0, len(others) + 1))
    layers = others[:cut] + [KIND_CENTER] + [KIND_TET] * n_t + [KIND_CENTER] + others[cut:]
    start = int(rng.integers(0, nx))
    layers = layers[start:] + layers[:start]              # periodic: any rotation of the cycle
    wdiag = rng.integers(0, 2, size=(ny, nx))             # wedge split diagonal of each z column (j, i)
    lat = rng.random((3, nx, ny, nz))                     # lateral faces of centre boxes: split when < q
    ldiag = rng.integers(0, 2, size=(3, nx, ny, nz))
    q_split = 0.0 if fractions[3] > 0 else 0.5
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    ncell = nx * ny * nz

    def vid(i, j, k):
        return (i % nx) + nx * ((j % ny) + ny * (k % nz))

    def face_label(i, j, k, f):
        """(split, diagonal) of face f of box (i, j, k), the same from both boxes of the face."""
        dn, side = f >> 1, f & 1
        kind = layers[i]
        if dn == 0:
            # x faces: between layers i and i +- 1
            other = layers[(i + (1 if side else -1)) % nx]
            kinds = {kind, other}
            if KIND_TET in kinds:
                return 1, 0
            return 0, 0
        if kind == KIND_TET:
            return 1, 0
        if kind == KIND_CENTER:
            # lateral face between two centre boxes (layer i): drawn once, stored at its lower box
            ijk = [i, j, k]
            if side:
                ijk[dn] += 1
            ijk = [ijk[0] % nx, ijk[1] % ny, ijk[2] % nz]
            return (1 if lat[dn, ijk[0], ijk[1], ijk[2]] < q_split else 0), int(ldiag[dn, ijk[0], ijk[1], ijk[2]])
        if kind == KIND_WEDGE and dn == 2:
            return 1, int(wdiag[j % ny, i % nx])
        return 0, 0

    etypes, verts, coords = [], [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                cell = i + nx * (j + ny * k)
                cid = [vid(i + a, j + b, k + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                cx = [h * np.array([i + a, j + b, k + c]) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                kind = layers[i]
                if kind == KIND_HEX:
                    etypes.append(M.HEX)
                    verts.append(cid)
                    coords.append(np.array(cx))
                elif kind == KIND_WEDGE:
                    tris = [(0, 1, 3), (0, 3, 2)] if wdiag[j, i] == 0 else [(0, 1, 2), (1, 3, 2)]
                    for t in tris:
                        t = _ccw(t, cx)
                        etypes.append(M.WEDGE)
                        verts.append([cid[q] for q in t] + [cid[q + 4] for q in t])
                        coords.append(np.array([cx[q] for q in t] + [cx[q + 4] for q in t]))
                elif kind == KIND_TET:
                    for t in _KUHN:
                        t = _orient_tet(list(t), cx)
                        etypes.append(M.TET)
                        verts.append([cid[q] for q in t])
                        coords.append(np.array([cx[q] for q in t]))
                else:
                    centre_id = nx * ny * nz + cell
                    centre = h * np.array([i + 0.5, j + 0.5, k + 0.5])
                    for f in range(6):
                        split, diag = face_label(i, j, k, f)
                        cf = _CUBE_FACES[f]
                        if not split:
                            base = list(cf)
                            if np.dot(np.cross(cx[base[1]] - cx[base[0]], cx[base[2]] - cx[base[0]]),
                                      centre - cx[base[0]]) < 0:
                                base = [base[0], base[2], base[1], base[3]]
                            etypes.append(M.PYR)
                            verts.append([cid[q] for q in base] + [centre_id])
                            coords.append(np.array([cx[q] for q in base] + [centre]))
                        else:
                            c00, c10, c01, c11 = cf
                            tris = ([(c00, c10, c11), (c00, c11, c01)] if diag == 0 else
                                    [(c00, c10, c01), (c10, c11, c01)])
                            for t in tris:
                                t = list(t)
                                if np.dot(np.cross(cx[t[1]] - cx[t[0]], cx[t[2]] - cx[t[0]]), centre - cx[t[0]]) < 0:
                                    t = [t[0], t[2], t[1]]
                                etypes.append(M.TET)
                                verts.append([cid[q] for q in t] + [centre_id])
                                coords.append(np.array([cx[q] for q in t] + [centre]))
    E = len(etypes)
    etype = np.array(etypes, dtype=np.int8)
    order = (np.full(E, p_fixed, dtype=np.int8) if p_fixed is not None
             else rng.integers(1, p_max + 1, size=E).astype(np.int8))
    # affine maps
    A = np.empty((E, 3, 3))
    c = np.empty((E, 3))
    for e in range(E):
        R = M.VERTS[int(etype[e])]
        X = coords[e]
        Rh = np.hstack([np.ones((R.shape[0], 1)), R])
        sol, *_ = np.linalg.lstsq(Rh, X, rcond=None)
        c[e] = sol[0]
        A[e] = sol[1:].T
        assert np.abs(Rh @ sol - X).max() < 1e-12, "element is not affine"
    jdet = np.linalg.det(A)
    assert (jdet > 0).all()
    Ainv = np.linalg.inv(A)                                # d xi / d x: Ainv[j][i] = d xi_j / d x_i
    metrics = jdet[:, None, None] * np.transpose(Ainv, (0, 2, 1))
    # faces
    keys = {}
    fe, fl, fcase, ftype, nrm, fjac = [], [], [], [], [], []
    for e in range(E):
        t = int(etype[e])
        for lf, fvs in enumerate(M.FACES[t]):
            ids = [verts[e][q] for q in fvs]
            key = tuple(sorted(ids))
            if key in keys:
                f = keys.pop(key)
                eL, lfL = fe[f][0], fl[f][0]
                idsL = [verts[eL][q] for q in M.FACES[int(etype[eL])][lfL]]
                match = tuple(ids.index(x) for x in idsL)
                fe[f][1], fl[f][1] = e, lf
                fcase[f] = M.perms_of(M.face_type(len(fvs))).index(match)
            else:
                keys[key] = len(fe)
                fe.append([e, -1])
                fl.append([lf, -1])
                fcase.append(-1)
                ft = M.face_type(len(fvs))
                ftype.append(ft)
                X = coords[e][list(fvs)]
                nv_ = np.cross(X[1] - X[0], X[2] - X[0])
                area = np.linalg.norm(nv_) * (0.5 if ft == M.F_TRI else 1.0)
                nv_ = nv_ / np.linalg.norm(nv_)
                cen = coords[e].mean(axis=0)
                if np.dot(nv_, X.mean(axis=0) - cen) < 0:
                    nv_ = -nv_
                nrm.append(nv_)
                fjac.append(area / M.REF_FACE_AREA[ft])
    assert not keys, f"{len(keys)} faces without a neighbour"
    return MixedMesh(etype, order, coords, A, c, jdet, metrics, np.array(fe, dtype=np.int32),
                     np.array(fl, dtype=np.int8), np.array(fcase, dtype=np.int8), np.array(ftype, dtype=np.int8),
                     np.array(nrm), np.array(fjac), ncell)


def _ccw(t, cx):
    """The triangle (indices into cube corners) counter-clockwise seen from +z."""
    a, b, cc = (cx[q] for q in t)
    if np.cross(b - a, cc - a)[2] < 0:
        return (t[0], t[2], t[1])
    return t


def _orient_tet(t, cx):
    """Vertex order with (v1 - v0) x (v2 - v0) pointing towards v3 (positive orientation)."""
    if np.dot(np.cross(cx[t[1]] - cx[t[0]], cx[t[2]] - cx[t[0]]), cx[t[3]] - cx[t[0]]) < 0:
        t = [t[0], t[2], t[1], t[3]]
    return t


NV = 5
GAMMA = 1.4


def dof_offsets(mesh):
    """Offsets of every element's coefficients ([var][mode], nv Ns values) in mesh order, and Ns."""
    ns = np.array([M.n_modes(int(t), int(p)) for t, p in zip(mesh.etype, mesh.order)], dtype=np.int64)
    off = np.zeros(mesh.n_elem + 1, dtype=np.int64)
    off[1:] = np.cumsum(NV * ns)
    return off, ns


def entropy_wave(x, t=0.0, amp=0.2):
    """Conserved variables of the advected entropy wave rho = 1 + amp sin(2 pi (x + y + z - 3 t)),
    velocity (1, 1, 1), pressure 1 (an exact solution of the Euler equations)."""
    rho = 1.0 + amp * np.sin(2 * np.pi * (x[:, 0] + x[:, 1] + x[:, 2] - 3.0 * t))
    E = 1.0 / (GAMMA - 1.0) + 0.5 * rho * 3.0
    return np.stack([rho, rho, rho, rho, E], axis=1)


def entropy_wave_dt(x, t=0.0, amp=0.2):
    """Time derivative of the conserved variables of entropy_wave."""
    drho = -3.0 * amp * 2 * np.pi * np.cos(2 * np.pi * (x[:, 0] + x[:, 1] + x[:, 2] - 3.0 * t))
    return np.stack([drho, drho, drho, drho, 1.5 * drho], axis=1)


def project(mesh, fn):
    """L2 projection of fn(x (N, 3)) -> (N, nv) onto every element's orthonormal basis (volume rule):
    c = sum_q w_q phi(xi_q) fn(x(xi_q)) (the physical mass matrix is J I)."""
    off, _ = dof_offsets(mesh)
    u = np.empty(int(off[-1]))
    groups = {}
    for e in range(mesh.n_elem):
        groups.setdefault((int(mesh.etype[e]), int(mesh.order[e])), []).append(e)
    for (t, p), elems in groups.items():
        elems = np.array(elems)
        B, _, w = M.volume_operators(t, p)
        xi, _ = M.volume_rule(t, p + 1)
        x = mesh.c[elems][:, None, :] + np.matmul(mesh.A[elems], xi.T).transpose(0, 2, 1)
        vals = fn(x.reshape(-1, 3)).reshape(len(elems), -1, NV)          # (n, Nc, nv)
        coef = np.matmul(vals.transpose(0, 2, 1), B * w[:, None])                # (n, nv, Ns)
        idx = off[elems][:, None] + np.arange(coef.shape[1] * coef.shape[2])[None, :]
        u[idx.ravel()] = coef.reshape(-1)
    return u


REF_VOLUME = {M.HEX: 8.0, M.WEDGE: 4.0, M.TET: 4.0 / 3.0, M.PYR: 8.0 / 3.0}


def dt_cfl(mesh, cfl=0.5, speed=None):
    """Time step cfl * min_e h_e / ((2 p_e + 1) s), h_e the cube root of the element volume and s the
    largest wave speed |v| + c (default: that of entropy_wave, sqrt(3) + sqrt(gamma / 0.8))."""
    if speed is None:
        speed = np.sqrt(3.0) + np.sqrt(GAMMA / 0.8)
    vol = mesh.jdet * np.array([REF_VOLUME[int(t)] for t in mesh.etype])
    h = np.cbrt(vol)
    return float(cfl * np.min(h / (2 * mesh.order.astype(np.float64) + 1)) / speed)
