"""Restart through the reference's file format with the state coming off and going back onto the GPU.

A run of 2 + 3 SSP-RK3 steps on the GPU is interrupted after 2 steps: the device state is downloaded
(dgx_get_state), written as a restart file (dgx_restart_write: Gmsh 2.2 mesh + Bezier solution
coefficients, restart_writers.c:57-420), read back (dgx_restart_read) into a NEW context, and continued
for 3 steps.  The result must equal the uninterrupted 5-step run up to the restart file's 16 significant
digits (Bezier coefficients printed % .15e, then mapped back to nodal values): relative L2 <= 1e-13.
The read-back state itself equals the written one to the same bound, and the VTU file of the GPU state
is written.  Curved HEX p=3 (the group kernel) and QUAD p=2 (the generic kernel).
"""
import math
import os

import numpy as np
import pytest

import dpgsolver_amd as dg

pytestmark = pytest.mark.gpu


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("kind", ["hex3", "quad2"])
def test_restart_continues_gpu_run(kind, tmp_path):
    _require_gpu()
    print(f"case: restart {kind}", flush=True)
    if kind == "hex3":
        p = 3
        mesh = dg.structured_mesh(3, (4, 4, 4), (0, 0, 0), (2 * math.pi,) * 3, p_g=3, amp=0.05, seed=4)
        geo = dg.geometry(mesh, p)
        u0 = dg.ic_taylor_green(mesh, p)
    else:
        p = 2
        mesh, geo, u0 = dg.vortex_2d(8, p)
    dt = dg.dt_cfl(mesh, p, u0)
    full = dg.Solver(mesh, geo, p, num_flux="roe")
    full.set_state(u0)
    full.step(dt, 5)
    u_full = full.get_state()
    first = dg.Solver(mesh, geo, p, num_flux="roe")
    first.set_state(u0)
    first.step(dt, 2)
    u2 = first.get_state()
    path = os.path.join(str(tmp_path), "restart.msh")
    dg.write_restart(path, mesh, p, u2)
    dg.write_vtu(os.path.join(str(tmp_path), "sol.vtu"), mesh, p, u2)
    assert os.path.getsize(os.path.join(str(tmp_path), "sol.vtu")) > 0
    first.close()
    u2r = dg.read_restart(path, mesh, p)
    assert _rel(u2r, u2) <= 1e-13
    second = dg.Solver(mesh, geo, p, num_flux="roe")
    second.set_state(u2r)
    second.step(dt, 3)
    e = _rel(second.get_state(), u_full)
    print(f"restart {kind}: continued vs uninterrupted rel L2 {e:.2e}", flush=True)
    assert e <= 1e-13
    full.close()
    second.close()
