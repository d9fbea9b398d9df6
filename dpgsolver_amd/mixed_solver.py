"""Host side of the batching form: tables of a mixed mesh for include/dgx_mixed.h and the solver mirror.

build_tables(mesh) groups the elements by (type, p) (the C-ABI wants them sorted by group), builds every
group's volume operators (basis at the volume cubature nodes, weighted reference gradients), one face
operator per (type, p, local face, correspondence case, face order) that occurs and one weight table per
(face type, face order), all from dpgsolver_amd.mixed.  The face order is max(p_L, p_R)
(compute_face_reference_order, face_solver.c:123-132).

MixedSolver keeps the user-facing state in mesh element order ([var][mode] per element, the layout of
dpgsolver_amd.mixed_mesh.project) and maps it to group order at the boundary.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _ffi
from . import mixed as M
from .mixed_mesh import NV, dof_offsets


@dataclass
class MixedTables:
    perm: np.ndarray          # group-order position -> mesh element
    gidx: np.ndarray          # group-order flat state index -> mesh-order flat index
    arrays: dict              # contiguous arrays kept alive for the C struct
    n_group: int
    group_keys: list          # [(type, p)]
    n_face_op: int
    n_face_wt: int


def build_tables(mesh, sum_factorized=True):
    """sum_factorized: give every group its factor tables (the kernels' sum-factorized paths: HEX and
    WEDGE tensor products, TET and PYR collapsed coordinates) or not (the dense path for every group)."""
    E = mesh.n_elem
    keys = [(int(t), int(p)) for t, p in zip(mesh.etype, mesh.order)]
    gkeys = sorted(set(keys))
    gid = {k: i for i, k in enumerate(gkeys)}
    perm = np.array(sorted(range(E), key=lambda e: (gid[keys[e]], e)), dtype=np.int64)
    inv = np.empty(E, dtype=np.int64)
    inv[perm] = np.arange(E)
    off_mesh, ns_mesh = dof_offsets(mesh)
    gidx = np.concatenate([off_mesh[e] + np.arange(NV * ns_mesh[e]) for e in perm])
    # groups and volume operators
    g_type, g_p, g_ns, g_nc, g_start, vol_off, vol_tp_off, vol = [], [], [], [], [0], [], [], []
    pos = 0
    for (t, p) in gkeys:
        B, G, w = M.volume_operators(t, p)
        g_type.append(t)
        g_p.append(p)
        g_ns.append(B.shape[1])
        g_nc.append(B.shape[0])
        g_start.append(g_start[-1] + sum(1 for k in keys if k == (t, p)))
        vol_off.append(pos)
        blk = np.concatenate([B.ravel(), (G * w[None, :, None]).ravel()])
        vol.append(blk)
        pos += blk.size
        if t in (M.TET, M.PYR) and sum_factorized:
            tp = M.collapsed_tables(t, p)
            vol_tp_off.append(pos)
            vol.append(tp)
            pos += tp.size
        elif t in (M.HEX, M.WEDGE) and sum_factorized:
            B1, D1, w1 = M.line_operators(p)
            tp = [B1.ravel(), D1.ravel(), w1]
            if t == M.HEX:
                tm, tq = M.line_traces(p)
                tp += [tm, tq]
            if t == M.WEDGE:
                Bt, Gr, Gs, wt = M.tri_operators(p)
                tm, tq = M.line_traces(p)
                tp += [Bt.ravel(), Gr.ravel(), Gs.ravel(), wt,
                       np.concatenate([M.wedge_edge_matrix(p, e).ravel() for e in range(3)]), tm, tq]
            tp = np.concatenate(tp)
            vol_tp_off.append(pos)
            vol.append(tp)
            pos += tp.size
        else:
            vol_tp_off.append(-1)
    # faces
    fop_index, fop_nq, fop_ns, fop_off, fop_red_off, fop_red_n, fop_data = {}, [], [], [], [], [], []
    fw_index, fw_n, fw_off, fw_data = {}, [], [], []
    fpos = wpos = 0
    F = mesh.n_face
    face_op = np.empty((F, 2), dtype=np.int32)
    face_wt = np.empty(F, dtype=np.int32)
    for f in range(F):
        eL, eR = (int(x) for x in mesh.face_elem[f])
        tL, pL, tR, pR = int(mesh.etype[eL]), int(mesh.order[eL]), int(mesh.etype[eR]), int(mesh.order[eR])
        q = max(pL, pR)
        sides = ((tL, pL, int(mesh.face_lf[f, 0]), 0), (tR, pR, int(mesh.face_lf[f, 1]), int(mesh.face_case[f])))
        for s, (t, p, lf, case) in enumerate(sides):
            k = (t, p, lf, case, q)
            if k not in fop_index:
                Bf = M.face_operator(t, p, lf, case, q)
                fop_index[k] = len(fop_nq)
                fop_nq.append(Bf.shape[0])
                fop_ns.append(Bf.shape[1])
                fop_off.append(fpos)
                fop_data.append(Bf.ravel())
                fpos += Bf.size
                if t in (M.HEX, M.WEDGE) and sum_factorized:
                    BF = (M.hex_face_modal(p, lf, Bf) if t == M.HEX else M.wedge_face_modal(p, lf, case, q, Bf))
                    fop_red_off.append(fpos)
                    fop_red_n.append(BF.shape[1])
                    fop_data.append(BF.ravel())
                    fpos += BF.size
                else:
                    fop_red_off.append(-1)
                    fop_red_n.append(0)
            face_op[f, s] = fop_index[k]
        ft = int(mesh.face_type[f])
        kw = (ft, q)
        if kw not in fw_index:
            _, wf = M.face_rule(ft, q + 1)
            fw_index[kw] = len(fw_n)
            fw_n.append(wf.size)
            fw_off.append(wpos)
            fw_data.append(wf)
            wpos += wf.size
        face_wt[f] = fw_index[kw]
    arrays = {
        "group_type": np.array(g_type, dtype=np.int32), "group_p": np.array(g_p, dtype=np.int32),
        "group_ns": np.array(g_ns, dtype=np.int32), "group_nc": np.array(g_nc, dtype=np.int32),
        "group_start": np.array(g_start, dtype=np.int64), "vol_ops": np.concatenate(vol),
        "vol_off": np.array(vol_off, dtype=np.int64), "vol_tp_off": np.array(vol_tp_off, dtype=np.int64),
        "jdet": np.ascontiguousarray(mesh.jdet[perm], dtype=np.float64),
        "metrics": np.ascontiguousarray(mesh.metrics[perm].reshape(E, 9), dtype=np.float64),
        "face_elem": np.ascontiguousarray(inv[mesh.face_elem.astype(np.int64)], dtype=np.int32),
        "face_op": face_op, "face_wt": face_wt,
        "face_normal": np.ascontiguousarray(mesh.normals, dtype=np.float64),
        "face_jac": np.ascontiguousarray(mesh.face_jac, dtype=np.float64),
        "fop_nq": np.array(fop_nq, dtype=np.int32), "fop_ns": np.array(fop_ns, dtype=np.int32),
        "fop_off": np.array(fop_off, dtype=np.int64), "fop_red_off": np.array(fop_red_off, dtype=np.int64),
        "fop_red_n": np.array(fop_red_n, dtype=np.int32),
        "fop_data": np.concatenate(fop_data),
        "face_lf": np.ascontiguousarray(mesh.face_lf, dtype=np.int32),
        "fw_n": np.array(fw_n, dtype=np.int32), "fw_off": np.array(fw_off, dtype=np.int64),
        "fw_data": np.concatenate(fw_data),
    }
    return MixedTables(perm, gidx, arrays, len(gkeys), gkeys, len(fop_nq), len(fw_n))


def _c_mesh(tab, mesh):
    a = tab.arrays
    cm = _ffi.DgxMixedMesh()
    cm.n_elem = mesh.n_elem
    cm.n_face = mesh.n_face
    cm.n_group = tab.n_group
    cm.reserved = 0
    cm.n_face_op = tab.n_face_op
    cm.n_face_wt = tab.n_face_wt
    for name, arr in a.items():
        if arr.dtype == np.float64:
            setattr(cm, name, arr.ctypes.data_as(_ffi.c_double_p))
        elif arr.dtype == np.int64:
            setattr(cm, name, arr.ctypes.data_as(_ffi.c_int64_p))
        else:
            setattr(cm, name, arr.ctypes.data_as(_ffi.c_int32_p))
    return cm


class MixedSolver:
    """The batching form on the GPU through include/dgx_mixed.h (the reference's explicit solver loop,
    solve_explicit.c, over a mixed mesh)."""

    def __init__(self, mesh, num_flux="lf", tables=None):
        self._lib = _ffi.load()
        self.mesh = mesh
        self.tab = tables if tables is not None else build_tables(mesh)
        flux = {"lf": _ffi.FLUX_LAX_FRIEDRICHS, "roe": _ffi.FLUX_ROE_PIKE}[num_flux]
        cm = _c_mesh(self.tab, mesh)
        ctx = ctypes.c_void_p()
        st = self._lib.dgx_mixed_create(ctypes.byref(cm), flux, ctypes.byref(ctx))
        if st != _ffi.DGX_OK:
            m = self._lib.dgx_mixed_last_error(None)
            raise _ffi.DgxError(st, m.decode() if m else "")
        self._ctx = ctx
        self.n_dof = int(self._lib.dgx_mixed_n_dof(ctx))
        assert self.n_dof == self.tab.gidx.size

    def _check(self, st):
        if st != _ffi.DGX_OK:
            m = self._lib.dgx_mixed_last_error(self._ctx)
            raise _ffi.DgxError(st, m.decode() if m else "")

    def close(self):
        if self._ctx:
            self._lib.dgx_mixed_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state in mesh order <-> group order
    def to_group(self, u):
        return np.ascontiguousarray(np.asarray(u, dtype=np.float64)[self.tab.gidx])

    def from_group(self, ug):
        out = np.empty_like(ug)
        out[self.tab.gidx] = ug
        return out

    def set_state(self, u):
        ug = self.to_group(u)
        self._check(self._lib.dgx_mixed_set_state(self._ctx, _ffi.PTR_HOST, ug.ctypes.data_as(ctypes.c_void_p)))

    def get_state(self):
        ug = np.empty(self.n_dof)
        self._check(self._lib.dgx_mixed_get_state(self._ctx, _ffi.PTR_HOST, ug.ctypes.data_as(ctypes.c_void_p)))
        return self.from_group(ug)

    def rhs(self, apply_minv=True):
        rg = np.empty(self.n_dof)
        self._check(self._lib.dgx_mixed_rhs(self._ctx, int(bool(apply_minv)), _ffi.PTR_HOST,
                                            rg.ctypes.data_as(ctypes.c_void_p)))
        return self.from_group(rg)

    def step(self, dt, nsteps=1):
        self._check(self._lib.dgx_mixed_step(self._ctx, float(dt), int(nsteps)))

    def set_launch(self, mode="persistent", tile_elems=None, blocks_per_cu=None):
        m = {"persistent": _ffi.MIXED_LAUNCH_PERSISTENT, "per_group": _ffi.MIXED_LAUNCH_PER_GROUP,
             "grouped": _ffi.MIXED_LAUNCH_GROUPED}[mode]
        cur_mode, tiles, bpc = self.get_launch()
        t = None
        if tile_elems is not None:
            t = np.ascontiguousarray(tile_elems, dtype=np.int32)
        self._check(self._lib.dgx_mixed_set_launch(self._ctx, m, None if t is None else _ffi.i32ptr(t),
                                                   int(bpc if blocks_per_cu is None else blocks_per_cu)))

    def set_face_maps(self, tensor=True):
        """Quad face maps of the HEX / WEDGE sides as tensor products (True) or as the dense face-modal
        rows (False); both give the same residual to round-off."""
        self._check(self._lib.dgx_mixed_set_face_maps(self._ctx, int(bool(tensor))))

    def tensor_face_ops(self):
        """Number of face operators the library applies in tensor form."""
        return int(self._lib.dgx_mixed_tensor_face_ops(self._ctx))

    def get_launch(self):
        mode = ctypes.c_int()
        tiles = np.empty(self.tab.n_group, dtype=np.int32)
        bpc = ctypes.c_int32()
        self._check(self._lib.dgx_mixed_get_launch(self._ctx, ctypes.byref(mode), _ffi.i32ptr(tiles),
                                                   ctypes.byref(bpc)))
        names = {_ffi.MIXED_LAUNCH_PERSISTENT: "persistent", _ffi.MIXED_LAUNCH_PER_GROUP: "per_group",
                 _ffi.MIXED_LAUNCH_GROUPED: "grouped"}
        return names[mode.value], tiles, bpc.value

    def autotune(self, reps=5):
        """Runs the library's autotuner; returns the best microseconds per launch of every group."""
        t = np.empty(self.tab.n_group)
        self._check(self._lib.dgx_mixed_autotune(self._ctx, int(reps), _ffi.dptr(t)))
        return t

    def persistent_waves(self):
        """Occupancy variant (waves per SIMD) of the persistent volume kernel the autotuner chose."""
        return int(self._lib.dgx_mixed_persistent_waves(self._ctx))

    def time_stage(self, dt, nstages=30):
        """(stage, face kernel, volume kernel) mean microseconds."""
        us = np.empty(3)
        self._check(self._lib.dgx_mixed_time_stage(self._ctx, float(dt), int(nstages), _ffi.dptr(us)))
        return us
