"""dpgsolver_amd: MI355X-native explicit DG residual path (volume + face residual, inverse mass,
SSP-RK3 stage with positivity) of DPGSolver, behind the C-ABI in include/dgx.h.

The compute path is libdgx.so (hand-written gfx950 kernels); this package only holds ctypes bindings,
the host-side set-up helpers, the solver mirror and the matrix-free implicit solver over the
linearized-residual kernels.
"""
from . import _ffi
from .problem import (Geometry, Mesh, dt_cfl, geometry, geometry_gpu, ic_isentropic_vortex, ic_taylor_green,
                      ic_uniform_flow, randn, read_gmsh, read_restart, solution_nodes_xyz, structured_mesh,
                      tgv_3d, vortex_2d, write_restart, write_vtu)
from .solver import Solver, residual_function
from .implicit import BlockJacobi, gmres, solve_implicit, solve_linear

__all__ = ["Mesh", "Geometry", "Solver", "structured_mesh", "geometry", "ic_isentropic_vortex",
           "ic_taylor_green", "dt_cfl", "randn", "solution_nodes_xyz", "vortex_2d", "tgv_3d",
           "residual_function", "read_gmsh", "ic_uniform_flow", "geometry_gpu", "solve_implicit", "solve_linear",
           "gmres", "BlockJacobi", "write_restart", "read_restart", "write_vtu"]
