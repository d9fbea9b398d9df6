"""GPU path on meshes with non-aligned faces (neigh_info[1].ind_ord != 0).

The meshes are periodic structured meshes with the local numbering of some volumes turned by 180
degrees (tests/meshtools.py): the same physical mesh, but faces between turned and unturned volumes see
each other's nodes in reverse order (2-D LINE faces: case 1; 3-D QUAD faces: cases 1 and 3).  The kernels
gather the neighbour's trace through the reference's node correspondence (nc_fc, compute_face_rlhs_T.c:
133-146,312-318); HEX p=3 runs the generic stage kernel here (the group kernel needs aligned faces).

- residual (pre- and post-M^-1) and 3 SSP-RK3 steps equal the oracle on the turned mesh (relative L2
  <= 1e-12, max-norm <= 1e-11 as the other parity tests);
- mapped back to the original numbering they equal the GPU results on the original mesh;
- the JVP equals the oracle's complex step and the VJP passes the dot-product test against it;
- the Jacobian blocks equal the blocks probed with the JVP.
"""
import math

import numpy as np
import pytest

import dpgsolver_amd as dg
from dpgsolver_amd import implicit
from oracle.oracle import Oracle
from tests.meshtools import rotate_state, rotate_volumes, unrotate_state

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL_MAX = 1e-11


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    emax = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    assert emax < TOL_MAX, f"max-norm relative error {emax:.3e}"
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


CASES = [(2, 2, 4, "lf"), (2, 3, 4, "roe"), (3, 2, 3, "roe"), (3, 3, 4, "roe"), (3, 4, 2, "lf")]


def _case(d, p, n):
    L = 2 * math.pi
    mesh = dg.structured_mesh(d, (n,) * d, (0,) * d, (L,) * d, p_g=1)
    turned = [1, 2, 5] if d == 2 else [0, 3, 5, 6]
    rot = rotate_volumes(mesh, turned)
    u = dg.ic_isentropic_vortex(mesh, p) if d == 2 else dg.ic_taylor_green(mesh, p, 0.3)
    # 1e-3 nodal noise: at p = 3, 1e-2 is amplified by the nodal -> Bezier map into a genuine positivity
    # failure (the oracle fails the same way on the unturned mesh)
    u = u + 1e-3 * dg.randn(7, u.shape)
    return mesh, rot, turned, u


@pytest.mark.parametrize("d,p,n,flux", CASES)
def test_turned_mesh_vs_oracle_and_original(d, p, n, flux):
    _require_gpu()
    print(f"case: turned volumes d={d} p={p} n={n} {flux}", flush=True)
    mesh, rot, turned, u = _case(d, p, n)
    assert (rot.face_ord != 0).any()
    geo_r, geo_0 = dg.geometry(rot, p), dg.geometry(mesh, p)
    ur = rotate_state(u, turned, d, p)
    orc = Oracle(rot, p, num_flux=0 if flux == "lf" else 1, geometry=geo_r)
    s = dg.Solver(rot, geo_r, p, num_flux=flux)
    s0 = dg.Solver(mesh, geo_0, p, num_flux=flux)
    s.set_state(ur)
    s0.set_state(u)
    for minv in (False, True):
        r, _ = s.compute_rhs(apply_minv=minv)
        r_ref, _ = orc.residual(ur, apply_minv=minv)
        assert rel_err(r, r_ref) < TOL, minv
        r0, _ = s0.compute_rhs(apply_minv=minv)
        assert rel_err(unrotate_state(r, turned, d, p), r0) < TOL, minv
    dt = dg.dt_cfl(mesh, p, u)
    s.step(dt, 3)
    s0.step(dt, 3)
    a, ap = ur.copy(), np.zeros_like(ur)
    orc.step(dt, a, ap, 3)
    us = s.get_state()
    assert rel_err(us, a) < TOL
    assert rel_err(unrotate_state(us, turned, d, p), s0.get_state()) < TOL
    s.close()
    s0.close()


@pytest.mark.parametrize("d,p,n,flux", [(2, 2, 4, "roe"), (3, 3, 3, "lf")])
def test_turned_mesh_linearization(d, p, n, flux):
    _require_gpu()
    import torch
    print(f"case: turned volumes linearization d={d} p={p}", flush=True)
    mesh, rot, turned, u = _case(d, p, n)
    geo = dg.geometry(rot, p)
    ur = rotate_state(u, turned, d, p)
    orc = Oracle(rot, p, num_flux=0 if flux == "lf" else 1, positivity=False, geometry=geo)
    s = dg.Solver(rot, geo, p, num_flux=flux)
    s.set_state(ur)
    rng = np.random.default_rng(4)
    for minv in (False, True):
        w = rng.standard_normal(ur.shape)
        v = rng.standard_normal(ur.shape)
        jw = s.jvp(w, apply_minv=minv)
        jw_ref = orc.jvp_cs(ur, w, apply_minv=minv)
        assert rel_err(jw, jw_ref) < TOL
        jtv = s.vjp(v, apply_minv=minv)
        lhs, rhs = float(np.sum(v * jw_ref)), float(np.sum(jtv * w))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), np.linalg.norm(v) * np.linalg.norm(jw_ref))
    E, nv, ns = s.shape
    BT = torch.empty((E, nv * ns, nv * ns), dtype=torch.float64, device="cuda")
    s.jacobian_blocks(BT, apply_minv=True)
    col, nc = implicit.element_colors(rot)

    class _Minv:
        shape = s.shape

        def jvp(self, w, apply_minv=False, out=None):
            return s.jvp(w, apply_minv=True, out=out)

    B = implicit.block_diagonal(_Minv(), col, nc, "cuda")
    assert (BT.transpose(1, 2) - B).abs().max().item() <= 1e-12 * B.abs().max().item()
    s.close()
