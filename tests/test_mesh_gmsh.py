"""Mesh ingest in the reference's format (dgx_mesh_read_gmsh): Gmsh 2.2 files with periodic boundaries.

The two periodic QUAD meshes of the reference's mesh integration test
(input/testing/integration/mesh/*.msh, committed under tests/golden/meshes/) are read, and the
connectivity is compared with the expected values the reference's test holds for them (the *.msh.data
files: node_nums, v_to_v, v_to_lf; test_integration_mesh.c:133-220).  The periodic straight mesh is
also checked against the structured 2x2 mesh of the same box through the oracle's residual.
"""
import os

import numpy as np
import pytest

import dpgsolver_amd as dg
from oracle.oracle import Oracle

MESHES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def _read_data_block(path, name):
    """Rows of the 'container Multiarray_Vector_i' block `name:` of a reference .msh.data file."""
    lines = open(path).read().splitlines()
    i = lines.index(name + ":")
    while not lines[i].startswith("ext_0/data"):
        i += 1
    rows = []
    for ln in lines[i + 1:]:
        t = ln.split()
        if not t or not t[0].lstrip("-").isdigit():
            break
        rows.append([int(x) for x in t[1:1 + int(t[0])]])
    return rows


@pytest.mark.parametrize("name", ["straight_2d_quad_periodic.msh", "parametric_2d_quad_periodic_reflected.msh"])
def test_connectivity_matches_reference_data(name):
    mesh, conn = dg.read_gmsh(os.path.join(MESHES, name), 2)
    data = os.path.join(MESHES, name + ".data")
    node_nums = [r for r in _read_data_block(data, "node_nums") if len(r) == 4]   # the volumes
    assert conn["node_nums"].tolist() == node_nums
    assert conn["v_to_v"].tolist() == _read_data_block(data, "v_to_v")
    assert conn["v_to_lf"].tolist() == _read_data_block(data, "v_to_lf")


def test_reflected_mesh_faces_are_not_aligned():
    """The reflected mesh joins volumes through faces whose vertex order differs (ind_ord != 0); the
    path is built for aligned faces, so these carry face_ord != 0 (dgx_create rejects them)."""
    mesh, conn = dg.read_gmsh(os.path.join(MESHES, "parametric_2d_quad_periodic_reflected.msh"), 2)
    assert (mesh.face_ord != 0).any()
    assert (conn["v_to_v"] < 0).any()          # and it has non-periodic boundaries


def test_periodic_mesh_equals_structured_mesh_through_the_oracle():
    """straight_2d_quad_periodic.msh is the periodic 2x2 QUAD mesh of [-1,1]^2: the oracle's residual on
    it equals the residual on the structured mesh of the same box, element by element."""
    p = 3
    fmesh, _ = dg.read_gmsh(os.path.join(MESHES, "straight_2d_quad_periodic.msh"), 2)
    smesh = dg.structured_mesh(2, (2, 2), (-1.0, -1.0), (1.0, 1.0), p_g=1)
    assert fmesh.n_elem == smesh.n_elem == 4 and fmesh.n_face == smesh.n_face == 8
    assert (fmesh.face_ord == 0).all()
    # element correspondence by centroid
    fc = fmesh.xyz_g.mean(axis=2)
    sc = smesh.xyz_g.mean(axis=2)
    perm = [int(np.argmin(np.linalg.norm(sc - fc[e], axis=1))) for e in range(4)]
    assert sorted(perm) == [0, 1, 2, 3]
    fgeo, sgeo = dg.geometry(fmesh, p), dg.geometry(smesh, p)
    us = dg.ic_isentropic_vortex(smesh, p, beta=1.0) + 1e-3 * dg.randn(5, (4, 4, (p + 1) ** 2))
    uf = us[perm]
    for flux in (0, 1):
        rs, _ = Oracle(smesh, p, num_flux=flux, geometry=sgeo).residual(us)
        rf, _ = Oracle(fmesh, p, num_flux=flux, geometry=fgeo).residual(uf)
        # the file's interior vertex sits at -2.75e-12 instead of 0
        assert np.abs(rf - rs[perm]).max() <= 1e-9 * np.abs(rs).max()
