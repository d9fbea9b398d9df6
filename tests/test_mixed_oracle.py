"""The mixed-mesh oracle (oracle/mixed_oracle.py) and the host tables of include/dgx_mixed.h, on the CPU.

The oracle is pinned by what the mathematics fixes, on seeded mixed HEX/WEDGE/TET/PYR meshes:
- free-stream preservation: a constant state has a zero residual (to round-off);
- conservation: on a periodic mesh the face fluxes cancel, so the integral of the pre-M^-1 residual
  (the mean modes) sums to zero for every variable;
- consistency: for the projected entropy wave (an exact Euler solution) the residual converges to the
  projected exact time derivative at rate >= p - 0.3 under refinement (measured 1.11, 2.18, 3.24 for
  p = 1, 2, 3);
- a planted mistake (the right side's face node correspondence dropped) breaks the convergence.
The host tables (dpgsolver_amd.mixed_solver.build_tables) are checked by evaluating the residual from
the tables alone, in the kernels' access pattern (group order, transposed-free operator rows, element
face lists, flux buffer), against the oracle.
"""
import numpy as np
import pytest

from dpgsolver_amd import mixed as M
from dpgsolver_amd import mixed_mesh as MM
from dpgsolver_amd.mixed_solver import build_tables
from oracle.mixed_oracle import MixedOracle
from oracle.oracle import flux_batch, numflux_batch

REF_VOL = {M.HEX: 8.0, M.WEDGE: 4.0, M.TET: 4.0 / 3.0, M.PYR: 8.0 / 3.0}


def _l2(mesh, x):
    off, ns = MM.dof_offsets(mesh)
    return np.sqrt(np.sum(np.repeat(mesh.jdet, 5 * ns) * x ** 2))


@pytest.mark.parametrize("flux", [0, 1])
def test_free_stream_and_conservation(flux):
    mesh = MM.build_mixed_mesh((4, 3, 3), seed=2)
    o = MixedOracle(mesh, num_flux=flux)
    q = np.array([1.1, 0.3, -0.2, 0.5, 3.0])
    uc = MM.project(mesh, lambda x: np.tile(q, (x.shape[0], 1)))
    r = o.residual(uc, apply_minv=False)
    assert np.abs(r).max() < 1e-13
    u = MM.project(mesh, lambda x: MM.entropy_wave(x)) + 1e-3 * np.random.default_rng(1).standard_normal(uc.shape)
    r = o.residual(u, apply_minv=False)
    off, ns = MM.dof_offsets(mesh)
    tot = np.zeros(5)
    for e in range(mesh.n_elem):
        tot += r[off[e] + np.arange(5) * ns[e]] * np.sqrt(REF_VOL[int(mesh.etype[e])])
    assert np.abs(tot).max() < 1e-13 * np.abs(r).max() * mesh.n_elem


def _wave_error(mesh, oracle_cls=MixedOracle):
    o = oracle_cls(mesh, num_flux=0)
    u = MM.project(mesh, lambda x: MM.entropy_wave(x))
    ex = MM.project(mesh, lambda x: MM.entropy_wave_dt(x))
    return _l2(mesh, o.residual(u) - ex) / _l2(mesh, ex)


@pytest.mark.parametrize("p", [1, 2])
def test_entropy_wave_consistency(p):
    e0 = _wave_error(MM.build_mixed_mesh((4, 3, 3), seed=1, p_fixed=p))
    e1 = _wave_error(MM.build_mixed_mesh((8, 6, 6), seed=1, p_fixed=p))
    rate = np.log2(e0 / e1)
    assert rate >= p - 0.3, (e0, e1, rate)


class _NoCorrespondence(MixedOracle):
    """A planted mistake: every right side evaluated as if it listed the face's vertices like the left."""

    def __init__(self, mesh, num_flux=0):
        super().__init__(mesh, num_flux)
        fg = {}
        for key, faces in self.fgroups.items():
            k = list(key)
            k[6] = 0
            fg.setdefault(tuple(k), []).extend(faces)
        self.fgroups = fg


def test_planted_correspondence_mistake_is_caught():
    m0 = MM.build_mixed_mesh((4, 3, 3), seed=1, p_fixed=2)
    m1 = MM.build_mixed_mesh((8, 6, 6), seed=1, p_fixed=2)
    assert (m0.face_case != 0).any()
    e0, e1 = _wave_error(m0, _NoCorrespondence), _wave_error(m1, _NoCorrespondence)
    assert np.log2(e0 / e1) < 1.0 and e1 > 10 * _wave_error(m1)


def _emulate(tab, mesh, ug, num_flux, apply_minv):
    """The residual from the C-ABI tables alone, in the kernels' order of work."""
    a = tab.arrays
    E = mesh.n_elem
    ns_of = np.zeros(E, dtype=np.int64)
    for g in range(tab.n_group):
        ns_of[a["group_start"][g]:a["group_start"][g + 1]] = a["group_ns"][g]
    dof = np.concatenate([[0], np.cumsum(5 * ns_of)])
    fb = []
    for f in range(mesh.n_face):
        oL, oR = a["face_op"][f]
        nq = a["fop_nq"][oL]
        tr = []
        for o, e in ((oL, a["face_elem"][f][0]), (oR, a["face_elem"][f][1])):
            ns = a["fop_ns"][o]
            assert ns == ns_of[e] and a["fop_nq"][o] == nq
            Bf = a["fop_data"][a["fop_off"][o]:a["fop_off"][o] + nq * ns].reshape(nq, ns)
            tr.append(Bf @ ug[dof[e]:dof[e + 1]].reshape(5, ns).T)
        nf = numflux_batch(3, num_flux, tr[0], tr[1], np.tile(a["face_normal"][f], (nq, 1)))
        w = a["fw_data"][a["fw_off"][a["face_wt"][f]]:][:nq]
        assert a["fw_n"][a["face_wt"][f]] == nq
        fb.append(nf * (w * a["face_jac"][f])[:, None])
    ef = [[] for _ in range(E)]
    for f in range(mesh.n_face):
        for s in range(2):
            ef[a["face_elem"][f][s]].append((f, s))
    r = np.zeros_like(ug)
    for g in range(tab.n_group):
        ns, nc, off = a["group_ns"][g], a["group_nc"][g], a["vol_off"][g]
        B = a["vol_ops"][off:off + nc * ns].reshape(nc, ns)
        G = a["vol_ops"][off + nc * ns:off + 4 * nc * ns].reshape(3, nc, ns)
        for e in range(a["group_start"][g], a["group_start"][g + 1]):
            assert len(ef[e]) == len(M.FACES[a["group_type"][g]])
            f = flux_batch(3, B @ ug[dof[e]:dof[e + 1]].reshape(5, ns).T)
            fr = np.einsum("ij,qiv->qjv", a["metrics"][e].reshape(3, 3), f)
            R = np.einsum("dqj,qdv->vj", G, fr)
            for fc, s in ef[e]:
                o = a["face_op"][fc][s]
                nq = a["fop_nq"][o]
                Bf = a["fop_data"][a["fop_off"][o]:a["fop_off"][o] + nq * ns].reshape(nq, ns)
                S = (Bf.T @ fb[fc]).T
                R = R + S if s else R - S
            r[dof[e]:dof[e + 1]] = (R / a["jdet"][e] if apply_minv else R).ravel()
    return r


@pytest.mark.parametrize("flux", [0, 1])
def test_tables_reproduce_oracle(flux):
    mesh = MM.build_mixed_mesh((4, 3, 3), seed=3)
    tab = build_tables(mesh)
    assert sorted(tab.group_keys) == tab.group_keys and tab.n_group >= 12
    u = MM.project(mesh, lambda x: MM.entropy_wave(x)) + 1e-3 * np.random.default_rng(0).standard_normal(
        int(MM.dof_offsets(mesh)[0][-1]))
    o = MixedOracle(mesh, num_flux=flux)
    for minv in (False, True):
        ro = o.residual(u, apply_minv=minv)
        rg = _emulate(tab, mesh, u[tab.gidx], flux, minv)
        re = np.empty_like(rg)
        re[tab.gidx] = rg
        assert np.linalg.norm(re - ro) / np.linalg.norm(ro) < 1e-12
