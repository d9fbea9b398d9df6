"""Host-side problem set-up over the C-ABI of include/dgx_setup.h: meshes, geometry, initial states.

Mirrors the reference's set-up chain (mesh -> solver elements -> geometry -> initial solution;
src/simulation/simulation.c:128-144, src/simulation/geometry/geometry_T.c:165-185,
src/simulation/test_case/solution/solution_T.c:127-156) for the structured periodic QUAD/HEX meshes the
explicit path is measured on.  All numerical work happens in the C library; this module only owns the
numpy buffers.
"""
from dataclasses import dataclass
import ctypes
import math

import numpy as np

from . import _ffi


@dataclass
class Mesh:
    d: int
    p_g: int
    n: tuple
    lo: tuple
    hi: tuple
    xyz_g: np.ndarray       # [E][d][(p_g+1)^d]
    face_elem: np.ndarray   # [F][2] int32
    face_lf: np.ndarray     # [F][2] int8
    face_ord: np.ndarray    # [F] int8
    h: float = 0.0          # smallest element edge when the mesh is not a structured box (0: from n, lo, hi)

    @property
    def n_elem(self):
        return self.xyz_g.shape[0]

    @property
    def n_face(self):
        return self.face_elem.shape[0]

    @property
    def h_min(self):
        if self.h > 0.0:
            return self.h
        return min((self.hi[k] - self.lo[k]) / self.n[k] for k in range(self.d))

    def c_struct(self):
        m = _ffi.DgxMesh()
        m.d = self.d
        m.p_g = self.p_g
        m.n_elem = self.n_elem
        m.n_face = self.n_face
        m.xyz_g = _ffi.dptr(self.xyz_g)
        m.face_elem = _ffi.i32ptr(self.face_elem)
        m.face_lf = _ffi.i8ptr(self.face_lf)
        m.face_ord = _ffi.i8ptr(self.face_ord)
        return m


@dataclass
class Geometry:
    metrics_vc: np.ndarray  # [E][d][d][Nc]
    jdet_vc: np.ndarray     # [E][Nc]
    normals_fc: np.ndarray  # [F][Nf][d]
    jdet_fc: np.ndarray     # [F][Nf]

    def c_struct(self):
        g = _ffi.DgxGeometry()
        g.metrics_vc = _ffi.dptr(self.metrics_vc)
        g.jdet_vc = _ffi.dptr(self.jdet_vc)
        g.normals_fc = _ffi.dptr(self.normals_fc)
        g.jdet_fc = _ffi.dptr(self.jdet_fc)
        return g


def structured_mesh(d, n, lo, hi, p_g=1, amp=0.0, seed=0):
    """Periodic structured QUAD/HEX mesh (see dgx_mesh_structured in include/dgx_setup.h)."""
    lib = _ffi.load()
    n = tuple(int(v) for v in n)
    na = np.array(n, dtype=np.int32)
    E = lib.dgx_mesh_structured_n_elem(d, _ffi.i32ptr(na))
    F = lib.dgx_mesh_structured_n_face(d, _ffi.i32ptr(na))
    Ng = (p_g + 1) ** d
    xyz = np.empty((E, d, Ng))
    fe = np.empty((F, 2), dtype=np.int32)
    fl = np.empty((F, 2), dtype=np.int8)
    fo = np.empty((F,), dtype=np.int8)
    lo_a = np.array(lo, dtype=np.float64)
    hi_a = np.array(hi, dtype=np.float64)
    _ffi.check(lib.dgx_mesh_structured(d, _ffi.i32ptr(na), _ffi.dptr(lo_a), _ffi.dptr(hi_a), p_g, float(amp),
                                       int(seed), _ffi.dptr(xyz), _ffi.i32ptr(fe), _ffi.i8ptr(fl), _ffi.i8ptr(fo)))
    return Mesh(d, p_g, n, tuple(lo), tuple(hi), xyz, fe, fl, fo)


def read_gmsh(path, d):
    """Mesh from a Gmsh 2.2 ASCII file of QUAD (d=2) / HEX (d=3) volumes with periodic boundaries, the
    reference's mesh input (dgx_mesh_read_gmsh in include/dgx_setup.h).  Returns (mesh, conn) with conn
    the reference's connectivity {"v_to_v", "v_to_lf", "node_nums"}; boundary faces that are not
    periodic leave the volume without a face record (dgx_create then rejects the mesh)."""
    lib = _ffi.load()
    bpath = str(path).encode()
    E, F = ctypes.c_int64(0), ctypes.c_int64(0)
    _ffi.check(lib.dgx_mesh_read_gmsh(bpath, d, ctypes.byref(E), ctypes.byref(F), None, None, None, None, None,
                                      None, None))
    E, F = E.value, F.value
    nv = 2 ** d
    xyz = np.empty((E, d, nv))
    fe = np.empty((F, 2), dtype=np.int32)
    fl = np.empty((F, 2), dtype=np.int8)
    fo = np.empty((F,), dtype=np.int8)
    vv = np.empty((E, 2 * d), dtype=np.int32)
    vl = np.empty((E, 2 * d), dtype=np.int32)
    nn = np.empty((E, nv), dtype=np.int32)
    e_, f_ = ctypes.c_int64(0), ctypes.c_int64(0)
    _ffi.check(lib.dgx_mesh_read_gmsh(bpath, d, ctypes.byref(e_), ctypes.byref(f_), _ffi.dptr(xyz), _ffi.i32ptr(fe),
                                      _ffi.i8ptr(fl), _ffi.i8ptr(fo), _ffi.i32ptr(vv), _ffi.i32ptr(vl),
                                      _ffi.i32ptr(nn)))
    # smallest edge: vertex pairs that differ in one tensor-product bit
    h = min(float(np.min(np.linalg.norm(xyz[:, :, v | (1 << k)] - xyz[:, :, v], axis=1)))
            for k in range(d) for v in range(nv) if not v & (1 << k))
    lo = tuple(float(xyz[:, k].min()) for k in range(d))
    hi = tuple(float(xyz[:, k].max()) for k in range(d))
    mesh = Mesh(d, 1, (0,) * d, lo, hi, xyz, fe, fl, fo, h)
    return mesh, {"v_to_v": vv, "v_to_lf": vl, "node_nums": nn}


def geometry(mesh, p):
    lib = _ffi.load()
    d = mesh.d
    Nc, Nf = (p + 1) ** d, (p + 1) ** (d - 1)
    g = Geometry(np.empty((mesh.n_elem, d, d, Nc)), np.empty((mesh.n_elem, Nc)),
                 np.empty((mesh.n_face, Nf, d)), np.empty((mesh.n_face, Nf)))
    ms = mesh.c_struct()
    _ffi.check(lib.dgx_geometry_build(ms, p, _ffi.dptr(g.metrics_vc), _ffi.dptr(g.jdet_vc),
                                      _ffi.dptr(g.normals_fc), _ffi.dptr(g.jdet_fc)))
    return g


def geometry_gpu(mesh, p, device=False, timing=None):
    """geometry(mesh, p) computed on the GPU (dgx_geometry_build_gpu; p, p_g <= 4).  device=False: numpy
    arrays (a Geometry); device=True: a Geometry of float64 torch cuda tensors, never copied to the host.
    timing: an optional dict that receives the kernel time ("kernel_ms")."""
    lib = _ffi.load()
    d = mesh.d
    Nc, Nf = (p + 1) ** d, (p + 1) ** (d - 1)
    shapes = ((mesh.n_elem, d, d, Nc), (mesh.n_elem, Nc), (mesh.n_face, Nf, d), (mesh.n_face, Nf))
    if device:
        import torch
        arrs = [torch.empty(sh, dtype=torch.float64, device="cuda") for sh in shapes]
        ptrs = [ctypes.c_void_p(a.data_ptr()) for a in arrs]
        kind = _ffi.PTR_DEVICE
    else:
        arrs = [np.empty(sh) for sh in shapes]
        ptrs = [ctypes.c_void_p(a.ctypes.data) for a in arrs]
        kind = _ffi.PTR_HOST
    kms = ctypes.c_double(0.0)
    _ffi.check(lib.dgx_geometry_build_gpu(mesh.c_struct(), p, kind, *ptrs, ctypes.byref(kms)))
    if timing is not None:
        timing["kernel_ms"] = kms.value
    return Geometry(*arrs)


def write_restart(path, mesh, p, u):
    """Restart file of the state u [E][d+2][Ns] in the reference's format (dgx_restart_write)."""
    lib = _ffi.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    _ffi.check(lib.dgx_restart_write(str(path).encode(), mesh.c_struct(), p, _ffi.dptr(u)))


def read_restart(path, mesh, p):
    """The state [E][d+2][Ns] of a restart file written for this mesh and order (dgx_restart_read)."""
    lib = _ffi.load()
    u = np.empty((mesh.n_elem, mesh.d + 2, (p + 1) ** mesh.d))
    _ffi.check(lib.dgx_restart_read(str(path).encode(), mesh.d, p, mesh.n_elem, _ffi.dptr(u)))
    return u


def write_vtu(path, mesh, p, u):
    """VTK XML solution file of the Euler state u (dgx_vtu_write)."""
    lib = _ffi.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    _ffi.check(lib.dgx_vtu_write(str(path).encode(), mesh.c_struct(), p, _ffi.dptr(u)))


def solution_nodes_xyz(mesh, p):
    lib = _ffi.load()
    out = np.empty((mesh.n_elem, mesh.d, (p + 1) ** mesh.d))
    _ffi.check(lib.dgx_solution_nodes_xyz(mesh.c_struct(), p, _ffi.dptr(out)))
    return out


def ic_isentropic_vortex(mesh, p, beta=5.0, xc=0.0, yc=0.0):
    lib = _ffi.load()
    u = np.empty((mesh.n_elem, 4, (p + 1) ** 2))
    _ffi.check(lib.dgx_ic_isentropic_vortex(mesh.c_struct(), p, beta, xc, yc, _ffi.dptr(u)))
    return u


def ic_taylor_green(mesh, p, mach=0.1):
    lib = _ffi.load()
    u = np.empty((mesh.n_elem, 5, (p + 1) ** 3))
    _ffi.check(lib.dgx_ic_taylor_green(mesh.c_struct(), p, mach, _ffi.dptr(u)))
    return u


def ic_uniform_flow(mesh, p, mach=2.0, angle=0.5, amp=1e-2, gamma=1.4):
    """Uniform flow (rho = 1, p = 1/gamma, so c = 1) at speed `mach` in direction (cos a, sin a) (2-D) or
    (cos a cos a/2, sin a cos a/2, sin a/2) (3-D), plus smooth periodic perturbations of amplitude `amp`
    on [0, 2 pi]^d, each varying along every coordinate (so that no face sees an exactly continuous
    normal velocity, where the Roe-Pike eigenvalue selection ties).  The steady state of a periodic domain
    is a uniform flow; the perturbation gives the implicit solver a residual to drive to zero."""
    x = solution_nodes_xyz(mesh, p)
    d = mesh.d
    X, Y = x[:, 0], x[:, 1]
    Z = x[:, 2] if d == 3 else 0.0
    if d == 2:
        vdir = (math.cos(angle), math.sin(angle))
    else:
        vdir = (math.cos(angle) * math.cos(0.5 * angle), math.sin(angle) * math.cos(0.5 * angle),
                math.sin(0.5 * angle))
    rho = 1.0 + amp * np.sin(X + 2 * Y + 3 * Z)
    pr = 1.0 / gamma + amp * np.cos(2 * X + Y - Z)
    vel = [mach * vdir[0] + amp * np.sin(2 * X - Y + Z), mach * vdir[1] + amp * np.cos(X + 3 * Y - 2 * Z)]
    if d == 3:
        vel.append(mach * vdir[2] + amp * np.sin(3 * X + Y + 2 * Z))
    ke = 0.5 * rho * sum(v * v for v in vel)
    return np.stack([rho] + [rho * v for v in vel] + [pr / (gamma - 1.0) + ke], axis=1)


def dt_cfl(mesh, p, u, cfl=0.5):
    lib = _ffi.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib.dgx_dt_cfl(mesh.d, p, mesh.n_elem, _ffi.dptr(u), mesh.h_min, cfl)


def randn(seed, shape):
    lib = _ffi.load()
    out = np.empty(int(np.prod(shape)))
    lib.dgx_randn(int(seed), out.size, _ffi.dptr(out))
    return out.reshape(shape)


def ops1d(p):
    lib = _ffi.load()
    o = _ffi.DgxOps1d()
    _ffi.check(lib.dgx_ops1d_build(p, ctypes_byref(o)))
    n = p + 1
    return {
        "x": np.array(o.x[:n]), "w": np.array(o.w[:n]),
        "D": np.array(o.D[:n * n]).reshape(n, n), "lm": np.array(o.lm[:n]), "lp": np.array(o.lp[:n]),
        "sb": np.array(o.sb[:n * n]).reshape(n, n), "bs": np.array(o.bs[:n * n]).reshape(n, n),
    }


def nodes_1d(family, p):
    lib = _ffi.load()
    x, w = np.empty(p + 1), np.empty(p + 1)
    _ffi.check(lib.dgx_nodes_1d(family, p, _ffi.dptr(x), _ffi.dptr(w)))
    return x, w


def ctypes_byref(o):
    return ctypes.byref(o)


# --- the measured problems ---------------------------------------------------------------------------

def vortex_2d(n=32, p=2):
    """2-D Euler, 32x32 QUAD on periodic [-5,5]^2, Shu isentropic vortex (beta=5) on (1,1,1,1)."""
    mesh = structured_mesh(2, (n, n), (-5.0, -5.0), (5.0, 5.0), p_g=1)
    geo = geometry(mesh, p)
    u0 = ic_isentropic_vortex(mesh, p, beta=5.0)
    return mesh, geo, u0


def tgv_3d(n=64, p=3, seed=0, amp=0.05, mach=0.1):
    """3-D Euler, n^3 curved HEX on periodic [0,2pi]^3 (sinusoidal perturbation 0.05h), Taylor-Green M=0.1."""
    L = 2.0 * math.pi
    mesh = structured_mesh(3, (n, n, n), (0.0, 0.0, 0.0), (L, L, L), p_g=p, amp=amp, seed=seed)
    geo = geometry(mesh, p)
    u0 = ic_taylor_green(mesh, p, mach)
    return mesh, geo, u0
