"""ctypes declarations of the C-ABI in include/dgx.h, include/dgx_setup.h and include/dgx_mixed.h.

The shared library ``libdgx.so`` sits next to this file (built by ``__graft_entry__.build()`` or
``make -C dpgsolver_amd/csrc``).  Loading fails loudly when it is missing: there is no fallback path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdgx.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int8_p = ctypes.POINTER(ctypes.c_int8)
c_int64_p = ctypes.POINTER(ctypes.c_int64)

DGX_OK = 0
DGX_ERR_INVALID_ARG = 1
STATUS_NAMES = {
    0: "DGX_OK",
    1: "DGX_ERR_INVALID_ARG",
    2: "DGX_ERR_HIP",
    3: "DGX_ERR_POSITIVITY",
    4: "DGX_ERR_NONFINITE",
    5: "DGX_ERR_GEOMETRY",
    6: "DGX_ERR_NO_DEVICE",
}
FLUX_LAX_FRIEDRICHS, FLUX_ROE_PIKE = 0, 1
PDE_EULER, PDE_NAVIER_STOKES = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
LAYOUT_REF, LAYOUT_SOA = 0, 1
PTR_HOST, PTR_DEVICE = 0, 1
NODES_GL, NODES_GLL = 0, 1
KRON_R_FASTEST = 0
SOLVER_E_SSP_RK_33, SOLVER_E_EULER, SOLVER_E_LS_RK_54 = 0, 1, 2


class DgxConfig(ctypes.Structure):
    _fields_ = [("d", ctypes.c_int32), ("p", ctypes.c_int32), ("pde", ctypes.c_int32),
                ("num_flux", ctypes.c_int32), ("positivity", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("mu", ctypes.c_double), ("Pr", ctypes.c_double), ("solver_type_e", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class DgxMesh(ctypes.Structure):
    _fields_ = [("d", ctypes.c_int32), ("p_g", ctypes.c_int32), ("n_elem", ctypes.c_int64),
                ("n_face", ctypes.c_int64), ("xyz_g", c_double_p), ("face_elem", c_int32_p),
                ("face_lf", c_int8_p), ("face_ord", c_int8_p)]


class DgxGeometry(ctypes.Structure):
    _fields_ = [("metrics_vc", c_double_p), ("jdet_vc", c_double_p), ("normals_fc", c_double_p),
                ("jdet_fc", c_double_p)]


class DgxOps1d(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int32), ("n", ctypes.c_int32), ("x", ctypes.c_double * 8),
                ("w", ctypes.c_double * 8), ("D", ctypes.c_double * 64), ("lm", ctypes.c_double * 8),
                ("lp", ctypes.c_double * 8), ("sb", ctypes.c_double * 64), ("bs", ctypes.c_double * 64)]


class DgxOperators(ctypes.Structure):
    _fields_ = [("node_type", ctypes.c_int32), ("kron_order", ctypes.c_int32), ("line", DgxOps1d)]


class DgxMixedMesh(ctypes.Structure):
    _fields_ = [("n_elem", ctypes.c_int64), ("n_face", ctypes.c_int64), ("n_group", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("group_type", c_int32_p), ("group_p", c_int32_p),
                ("group_ns", c_int32_p), ("group_nc", c_int32_p), ("group_start", c_int64_p),
                ("vol_ops", c_double_p), ("vol_off", c_int64_p), ("vol_tp_off", c_int64_p), ("jdet", c_double_p),
                ("metrics", c_double_p),
                ("face_elem", c_int32_p), ("face_lf", c_int32_p), ("face_op", c_int32_p), ("face_wt", c_int32_p),
                ("face_normal", c_double_p), ("face_jac", c_double_p), ("n_face_op", ctypes.c_int32),
                ("n_face_wt", ctypes.c_int32), ("fop_nq", c_int32_p), ("fop_ns", c_int32_p),
                ("fop_off", c_int64_p), ("fop_red_off", c_int64_p), ("fop_red_n", c_int32_p), ("fop_data", c_double_p), ("fw_n", c_int32_p), ("fw_off", c_int64_p),
                ("fw_data", c_double_p)]


MIXED_LAUNCH_PERSISTENT, MIXED_LAUNCH_PER_GROUP, MIXED_LAUNCH_GROUPED = 0, 1, 2


class DgxError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")


# Every exported symbol with (restype, argtypes); tests check that the library exports all of them.
SIGNATURES = {
    # dgx.h
    "dgx_create": (ctypes.c_int, [ctypes.POINTER(DgxConfig), ctypes.POINTER(DgxMesh),
                                  ctypes.POINTER(DgxGeometry), ctypes.POINTER(DgxOperators),
                                  ctypes.POINTER(ctypes.c_void_p)]),
    "dgx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "dgx_jacobian_blocks": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "dgx_jacobian_blocks_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "dgx_block_inverse_f32": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dgx_block_inverse": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dgx_block_apply": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]),
    "dgx_block_apply_sel": (ctypes.c_int, [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dgx_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "dgx_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "dgx_get_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "dgx_state_size": (ctypes.c_int64, [ctypes.c_void_p]),
    "dgx_set_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "dgx_get_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "dgx_rhs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_void_p, c_double_p]),
    "dgx_rk_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, c_double_p]),
    "dgx_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, c_double_p]),
    "dgx_stages_per_step": (ctypes.c_int, [ctypes.c_void_p]),
    "dgx_sync": (ctypes.c_int, [ctypes.c_void_p, c_int64_p]),
    "dgx_vjp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_void_p]),
    "dgx_jvp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_void_p]),
    "dgx_last_elapsed_ms": (ctypes.c_double, [ctypes.c_void_p]),
    # dgx_setup.h
    "dgx_nodes_1d": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_double_p, c_double_p]),
    "dgx_lagrange_1d": (ctypes.c_int, [ctypes.c_int, c_double_p, ctypes.c_int, c_double_p, ctypes.c_int,
                                       c_double_p]),
    "dgx_ops1d_build": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(DgxOps1d)]),
    "dgx_mesh_structured_n_elem": (ctypes.c_int64, [ctypes.c_int, c_int32_p]),
    "dgx_mesh_structured_n_face": (ctypes.c_int64, [ctypes.c_int, c_int32_p]),
    "dgx_mesh_structured": (ctypes.c_int, [ctypes.c_int, c_int32_p, c_double_p, c_double_p, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_uint64, c_double_p, c_int32_p, c_int8_p,
                                           c_int8_p]),
    "dgx_mesh_read_gmsh": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, c_int64_p, c_int64_p, c_double_p,
                                          c_int32_p, c_int8_p, c_int8_p, c_int32_p, c_int32_p, c_int32_p]),
    "dgx_geometry_build": (ctypes.c_int, [ctypes.POINTER(DgxMesh), ctypes.c_int, c_double_p, c_double_p,
                                          c_double_p, c_double_p]),
    "dgx_geometry_build_gpu": (ctypes.c_int, [ctypes.POINTER(DgxMesh), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_double_p]),
    "dgx_geometry_build_gpu_noalloc": (ctypes.c_int, [ctypes.POINTER(DgxMesh), ctypes.c_int, ctypes.c_void_p,
                                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                      ctypes.c_void_p, ctypes.c_void_p]),
    "dgx_restart_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(DgxMesh), ctypes.c_int, c_double_p]),
    "dgx_restart_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64, c_double_p]),
    "dgx_vtu_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(DgxMesh), ctypes.c_int, c_double_p]),
    "dgx_solution_nodes_xyz": (ctypes.c_int, [ctypes.POINTER(DgxMesh), ctypes.c_int, c_double_p]),
    "dgx_ic_isentropic_vortex": (ctypes.c_int, [ctypes.POINTER(DgxMesh), ctypes.c_int, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, c_double_p]),
    "dgx_ic_taylor_green": (ctypes.c_int, [ctypes.POINTER(DgxMesh), ctypes.c_int, ctypes.c_double,
                                           c_double_p]),
    "dgx_dt_cfl": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, c_double_p, ctypes.c_double,
                                     ctypes.c_double]),
    "dgx_randn": (None, [ctypes.c_uint64, ctypes.c_int64, c_double_p]),
    # dgx_mixed.h
    "dgx_mixed_create": (ctypes.c_int, [ctypes.POINTER(DgxMixedMesh), ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "dgx_mixed_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "dgx_mixed_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "dgx_mixed_n_dof": (ctypes.c_int64, [ctypes.c_void_p]),
    "dgx_mixed_set_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "dgx_mixed_get_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "dgx_mixed_rhs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "dgx_mixed_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]),
    "dgx_mixed_set_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int32_p, ctypes.c_int32]),
    "dgx_mixed_get_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), c_int32_p,
                                            c_int32_p]),
    "dgx_mixed_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_double_p]),
    "dgx_mixed_persistent_waves": (ctypes.c_int, [ctypes.c_void_p]),
    "dgx_mixed_set_face_maps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dgx_mixed_tensor_face_ops": (ctypes.c_int32, [ctypes.c_void_p]),
    "dgx_mixed_time_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, c_double_p]),
}

_lib = None


def load(path=None):
    """Load libdgx.so once and declare every signature.  Raises if the library is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libdgx.so not built ({p}); run __graft_entry__.build() first")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_double_p)


def i32ptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_int32_p)


def i8ptr(a):
    assert a.dtype == np.int8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_int8_p)


def check(status, ctx=None, lib=None):
    if status != DGX_OK:
        msg = ""
        if ctx is not None and lib is not None:
            m = lib.dgx_last_error(ctx)
            msg = m.decode() if m else ""
        raise DgxError(status, msg)
