"""ctypes binding of the oracle (oracle.h).  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's cpu_baseline leg import this module, and only as the
checker; the product path never does.  The library is built by ``make -C oracle`` (also done by
``__graft_entry__.build()``).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i8p = ctypes.POINTER(ctypes.c_int8)
_i64p = ctypes.POINTER(ctypes.c_int64)

_SIGS = {
    "orc_nodes_1d": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "orc_lagrange_1d": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp]),
    "orc_kron": (ctypes.c_int, [ctypes.c_int, _i32p, _i32p, ctypes.POINTER(_dp), _dp]),
    "orc_mm": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
    "orc_create": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                     _dp, _i32p, _i8p, ctypes.c_int, ctypes.c_int]),
    "orc_destroy": (None, [ctypes.c_void_p]),
    "orc_geometry": (None, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "orc_set_geometry": (None, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "orc_set_viscous": (None, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "orc_gradient": (None, [ctypes.c_void_p, _dp, _dp, _dp]),
    "orc_set_penalty": (None, [ctypes.c_void_p, ctypes.c_double]),
    "orc_set_threads": (None, [ctypes.c_void_p, ctypes.c_int]),
    "orc_set_sum_factorized": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "orc_operator": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _dp]),
    "orc_residual": (ctypes.c_double, [ctypes.c_void_p, _dp, _dp, ctypes.c_int]),
    "orc_jvp_cs": (None, [ctypes.c_void_p, _dp, _dp, ctypes.c_double, _dp, ctypes.c_int]),
    "orc_positivity": (ctypes.c_int64, [ctypes.c_void_p, _dp]),
    "orc_rk_stage": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, _dp, _i64p]),
    "orc_set_time_stepper": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "orc_face_corr": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p]),
    "orc_set_face_ord": (ctypes.c_int, [ctypes.c_void_p, _i8p]),
    "orc_n_stages": (ctypes.c_int, [ctypes.c_void_p]),
    "orc_flux_point": (None, [ctypes.c_int, _dp, _dp]),
    "orc_numflux_point": (None, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
    "orc_flux_batch": (None, [ctypes.c_int, ctypes.c_int64, _dp, _dp]),
    "orc_numflux_batch": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp, _dp]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built ({LIB_PATH}); run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def nodes_1d(family, p):
    lib = load()
    x, w = np.empty(p + 1), np.empty(p + 1)
    if lib.orc_nodes_1d(family, p, _d(x), _d(w)) != 0:
        raise ValueError("orc_nodes_1d failed")
    return x, w


def lagrange_1d(xs, y, der=0):
    lib = load()
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty((y.size, xs.size))
    if lib.orc_lagrange_1d(xs.size, _d(xs), y.size, _d(y), der, _d(out)) != 0:
        raise ValueError("orc_lagrange_1d failed")
    return out


def kron(mats):
    """Dense tensor-product operator of 1-D factors [A_r, A_s, (A_t)] (r fastest)."""
    lib = load()
    mats = [np.ascontiguousarray(a, dtype=np.float64) for a in mats]
    d = len(mats)
    m = np.array([a.shape[0] for a in mats], dtype=np.int32)
    n = np.array([a.shape[1] for a in mats], dtype=np.int32)
    arr = (_dp * d)(*[_d(a) for a in mats])
    out = np.empty((int(np.prod(m)), int(np.prod(n))))
    lib.orc_kron(d, m.ctypes.data_as(_i32p), n.ctypes.data_as(_i32p), arr, _d(out))
    return out


def mm(A, B):
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    C = np.empty((A.shape[0], B.shape[1]))
    lib.orc_mm(A.shape[0], A.shape[1], B.shape[1], _d(A), _d(B), _d(C))
    return C


def flux_point(d, u):
    lib = load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.empty(d * (d + 2))
    lib.orc_flux_point(d, _d(u), _d(f))
    return f.reshape(d + 2, d).T  # [dim][eq]


def face_corr(d_face, n1, ord_):
    out = np.empty(n1 ** d_face, dtype=np.int32)
    if load().orc_face_corr(d_face, n1, ord_, out.ctypes.data_as(_i32p)) != 0:
        raise ValueError("unsupported face correspondence")
    return out


def numflux_point(d, num_flux, ul, ur, n):
    lib = load()
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    ur = np.ascontiguousarray(ur, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    nf = np.empty(d + 2)
    lib.orc_numflux_point(d, num_flux, _d(ul), _d(ur), _d(n), _d(nf))
    return nf


def flux_batch(d, u):
    """Euler fluxes of n states u [n][nv]: f [n][d][nv]."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    n = u.shape[0]
    f = np.empty((n, d + 2, d))
    load().orc_flux_batch(d, n, _d(u), _d(f))
    return f.transpose(0, 2, 1)


def numflux_batch(d, num_flux, ul, ur, nrm):
    """Numerical fluxes n.F* of n face points: nf [n][nv]."""
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    ur = np.ascontiguousarray(ur, dtype=np.float64)
    nrm = np.ascontiguousarray(nrm, dtype=np.float64)
    nf = np.empty_like(ul)
    load().orc_numflux_batch(d, num_flux, ul.shape[0], _d(ul), _d(ur), _d(nrm), _d(nf))
    return nf


class Oracle:
    """Dense CPU restatement of the explicit residual path on one mesh (see oracle.h)."""

    def __init__(self, mesh, p, num_flux=0, positivity=True, geometry=None, viscous=None):
        lib = load()
        self.d, self.p = mesh.d, p
        self.n_elem, self.n_face = mesh.n_elem, mesh.n_face
        self.Ns = (p + 1) ** mesh.d
        self.Nf = (p + 1) ** (mesh.d - 1)
        self.nv = mesh.d + 2
        # with injected geometry the oracle's own (dense, slow) geometry construction is skipped
        xyz = None if geometry is not None else _d(mesh.xyz_g)
        self._ptr = lib.orc_create(mesh.d, p, mesh.p_g, mesh.n_elem, mesh.n_face, xyz,
                                   mesh.face_elem.ctypes.data_as(_i32p), mesh.face_lf.ctypes.data_as(_i8p),
                                   int(num_flux), int(bool(positivity)))
        if not self._ptr:
            raise ValueError("orc_create failed")
        if geometry is not None:
            self.set_geometry(geometry)
        if viscous is not None:
            mu, pr = viscous
            lib.orc_set_viscous(self._ptr, float(mu), float(pr))
        fo = getattr(mesh, "face_ord", None)
        if fo is not None and np.any(fo != 0):
            self.set_face_ord(fo)

    def __del__(self):
        if getattr(self, "_ptr", None):
            load().orc_destroy(self._ptr)
            self._ptr = None

    def geometry(self):
        d, E, F, Nc, Nf = self.d, self.n_elem, self.n_face, self.Ns, self.Nf
        m = np.empty((E, d, d, Nc))
        j = np.empty((E, Nc))
        n = np.empty((F, Nf, d))
        jf = np.empty((F, Nf))
        load().orc_geometry(self._ptr, _d(m), _d(j), _d(n), _d(jf))
        return m, j, n, jf

    def set_geometry(self, geo):
        """Use the given dgx Geometry (metrics_vc, jdet_vc, normals_fc, jdet_fc) for the residual."""
        load().orc_set_geometry(self._ptr, _d(geo.metrics_vc), _d(geo.jdet_vc), _d(geo.normals_fc),
                                _d(geo.jdet_fc))

    def operator(self, name):
        out = np.empty(self.Ns * self.Ns)
        cnt = load().orc_operator(self._ptr, name.encode(), _d(out))
        if cnt < 0:
            raise KeyError(name)
        return out[:cnt].copy()

    def residual(self, u, apply_minv=True):
        u = np.ascontiguousarray(u, dtype=np.float64)
        rhs = np.empty_like(u)
        mx = load().orc_residual(self._ptr, _d(u), _d(rhs), int(bool(apply_minv)))
        return rhs, mx

    def set_penalty(self, eta):
        """Test hook: BR2 penalty eta instead of 1.01 x faces per element."""
        load().orc_set_penalty(self._ptr, float(eta))

    def set_threads(self, n):
        """OpenMP threads of residual / rk_stage (1 = the reference's sequential order; more threads
        change the face-term summation order, i.e. round-off only)."""
        load().orc_set_threads(self._ptr, int(n))

    def set_sum_factorized(self, on=True):
        """Sum-factorized application of the element operators (the same operators up to summation order;
        a stronger CPU baseline than the dense per-element operators)."""
        if load().orc_set_sum_factorized(self._ptr, int(bool(on))) != 0:
            raise ValueError("sum-factorized application needs a collocated (GL) basis")

    def gradient(self, u):
        """(grad_v, grad), each [E][d][nv][Ns]: weak volume gradient and the lifted gradient."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        gv = np.empty((self.n_elem, self.d, self.nv, self.Ns))
        g = np.empty_like(gv)
        load().orc_gradient(self._ptr, _d(u), _d(gv), _d(g))
        return gv, g

    def jvp_cs(self, u, w, h=1e-30, apply_minv=False):
        u = np.ascontiguousarray(u, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = np.empty_like(u)
        load().orc_jvp_cs(self._ptr, _d(u), _d(w), h, _d(out), int(bool(apply_minv)))
        return out

    def positivity(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        bad = load().orc_positivity(self._ptr, _d(u))
        return u, bad

    def rk_stage(self, stage, dt, u, up):
        bad = ctypes.c_int64(-1)
        mx = load().orc_rk_stage(self._ptr, stage, dt, _d(u), _d(up), ctypes.byref(bad))
        return mx, bad.value

    def set_face_ord(self, face_ord):
        """neigh_info[1].ind_ord per face (None: all aligned)."""
        if face_ord is None:
            rc = load().orc_set_face_ord(self._ptr, None)
        else:
            self._face_ord = np.ascontiguousarray(face_ord, dtype=np.int8)
            rc = load().orc_set_face_ord(self._ptr, self._face_ord.ctypes.data_as(_i8p))
        if rc:
            raise ValueError(f"orc_set_face_ord: {rc}")

    def set_time_stepper(self, scheme):
        """0 SSP-RK3, 1 forward Euler, 2 LS-RK54 (or the names "ssp_rk_33", "euler", "ls_rk_54")."""
        scheme = {"ssp_rk_33": 0, "euler": 1, "ls_rk_54": 2}.get(scheme, scheme)
        if load().orc_set_time_stepper(self._ptr, int(scheme)) != 0:
            raise ValueError(f"bad time stepper {scheme}")

    @property
    def n_stages(self):
        return load().orc_n_stages(self._ptr)

    def step(self, dt, u, up, nsteps=1):
        mx = 0.0
        for _ in range(nsteps):
            for s in range(self.n_stages):
                mx, bad = self.rk_stage(s, dt, u, up)
                if bad >= 0:
                    raise RuntimeError(f"positivity failure in element {bad}")
        return mx
