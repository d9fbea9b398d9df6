"""Geometry set-up on the GPU (dgx_geometry_build_gpu) against the host build (dgx_geometry_build,
itself pinned to the oracle's geometry and the reference's metric definitions).

- straight and curved QUAD / HEX meshes, p = 1..4 with p_g = 1 and p_g = p, and the reference's periodic
  Gmsh QUAD mesh: metric terms, det J, unit normals and face det J equal the host's to 1e-12 of each
  array's largest magnitude (the two builds sum in different orders: direct sums here, sum
  factorization on the host); host and device output pointers agree bit for bit;
- the residual of a solver built on the GPU geometry equals the one built on the host geometry;
- an inverted volume is reported (DGX_ERR_GEOMETRY), p = 5 is refused;
- against the oracle's own geometry (the reference's construction from absolute coordinates) directly,
  to 1e-12 of each array's largest magnitude (the oracle's absolute-coordinate round-off, see
  tests/test_host_setup.py);
- more volumes than one launch takes (2^20 workgroups): every volume is written.
"""
import math
import os

import numpy as np
import pytest

import dpgsolver_amd as dg

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _close(a, b, tol=1e-12):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * scale


def _mesh(d, n, p_g, amp):
    L = 2 * math.pi
    return dg.structured_mesh(d, (n,) * d, (0,) * d, (L,) * d, p_g=p_g, amp=amp, seed=3)


@pytest.mark.parametrize("d,n,p,p_g,amp", [(2, 5, 1, 1, 0.0), (2, 4, 2, 2, 0.08), (2, 3, 3, 3, 0.08),
                                           (2, 3, 4, 4, 0.08), (2, 4, 3, 1, 0.0), (3, 3, 1, 1, 0.0),
                                           (3, 3, 2, 2, 0.06), (3, 3, 3, 3, 0.06), (3, 2, 4, 4, 0.06),
                                           (3, 3, 4, 1, 0.0), (3, 2, 2, 4, 0.06)])
def test_matches_host_geometry(d, n, p, p_g, amp):
    _require_gpu()
    mesh = _mesh(d, n, p_g, amp)
    print(f"geometry d={d} n={n} p={p} p_g={p_g}", flush=True)
    h = dg.geometry(mesh, p)
    g = dg.geometry_gpu(mesh, p)
    for a, b in ((g.metrics_vc, h.metrics_vc), (g.jdet_vc, h.jdet_vc), (g.normals_fc, h.normals_fc),
                 (g.jdet_fc, h.jdet_fc)):
        assert a.shape == b.shape
        assert _close(a, b), np.abs(a - b).max()
    gd = dg.geometry_gpu(mesh, p, device=True)
    for a, b in zip((gd.metrics_vc, gd.jdet_vc, gd.normals_fc, gd.jdet_fc),
                    (g.metrics_vc, g.jdet_vc, g.normals_fc, g.jdet_fc)):
        assert np.array_equal(a.cpu().numpy(), b)


def test_gmsh_mesh():
    _require_gpu()
    mesh, _ = dg.read_gmsh(os.path.join(HERE, "golden", "meshes", "straight_2d_quad_periodic.msh"), 2)
    for p in (1, 3):
        h = dg.geometry(mesh, p)
        g = dg.geometry_gpu(mesh, p)
        for a, b in ((g.metrics_vc, h.metrics_vc), (g.jdet_vc, h.jdet_vc), (g.normals_fc, h.normals_fc),
                     (g.jdet_fc, h.jdet_fc)):
            assert _close(a, b)


def test_residual_on_gpu_geometry():
    _require_gpu()
    mesh, geo, u = dg.tgv_3d(4, 3)
    g = dg.geometry_gpu(mesh, 3)
    s1 = dg.Solver(mesh, geo, 3, num_flux="roe")
    s1.set_state(u)
    r1, _ = s1.compute_rhs()
    s2 = dg.Solver(mesh, g, 3, num_flux="roe")
    s2.set_state(u)
    r2, _ = s2.compute_rhs()
    assert np.abs(r1 - r2).max() <= 1e-10 * np.abs(r1).max()


def test_errors():
    _require_gpu()
    mesh = _mesh(2, 3, 1, 0.0)
    with pytest.raises(dg._ffi.DgxError):
        dg.geometry_gpu(mesh, 5)
    bad = _mesh(2, 3, 1, 0.0)
    xg = bad.xyz_g.copy()
    xg[4, 0] = xg[4, 0][::-1].copy()       # mirror volume 4 in x: det J < 0
    bad.xyz_g = xg
    with pytest.raises(dg._ffi.DgxError) as ei:
        dg.geometry_gpu(bad, 2)
    assert ei.value.status == 5


@pytest.mark.parametrize("d,p,n", [(2, 2, (4, 3)), (2, 3, (3, 3)), (3, 2, (2, 3, 2)), (3, 3, (3, 3, 3)),
                                   (3, 4, (2, 2, 3))])
def test_matches_oracle_geometry(d, p, n):
    _require_gpu()
    from oracle.oracle import Oracle
    L = 2 * math.pi
    mesh = dg.structured_mesh(d, n, (0,) * d, (L,) * d, p_g=p, amp=0.05, seed=3)
    print(f"geometry vs oracle d={d} p={p} n={n}", flush=True)
    g = dg.geometry_gpu(mesh, p)
    m, j, nrm, jf = Oracle(mesh, p).geometry()
    assert np.abs(g.metrics_vc - m).max() <= 1e-12 * np.abs(m).max()
    assert np.abs(g.jdet_vc - j).max() <= 1e-12 * np.abs(j).max()
    assert np.abs(g.normals_fc - nrm).max() <= 1e-12
    assert np.abs(g.jdet_fc - jf).max() <= 1e-12 * np.abs(jf).max()


def test_launch_split_over_2p20_volumes():
    """1025 x 1025 QUAD p=1 volumes (more than 2^20 workgroups of one launch) against the host build."""
    _require_gpu()
    n = 1025
    mesh = dg.structured_mesh(2, (n, n), (0, 0), (1.0, 1.0), p_g=1, amp=0.0, seed=0)
    print(f"geometry launch split E={mesh.n_elem}", flush=True)
    h = dg.geometry(mesh, 1)
    g = dg.geometry_gpu(mesh, 1)
    for a, b in ((g.metrics_vc, h.metrics_vc), (g.jdet_vc, h.jdet_vc), (g.normals_fc, h.normals_fc),
                 (g.jdet_fc, h.jdet_fc)):
        assert _close(a, b), np.abs(a - b).max()
