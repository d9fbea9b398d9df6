"""Batching form (include/dgx_mixed.h) on the GPU against the mixed-mesh oracle.

Small seeded mixed meshes (HEX, WEDGE, TET, PYR, per-element p in 1..5) with a perturbed entropy-wave
state:
- the residual, before and after M^-1, and 3 SSP-RK3 steps equal the oracle to 1e-12 relative (L2) and
  1e-11 (max-norm), for Lax-Friedrichs and Roe-Pike;
- the persistent launch, the grouped launch, the per-group launches and every tile shape the autotuner
  picks give bit-identical results (each lane's sums do not depend on the launch shape);
- repeated graph replays equal single steps;
- the sum-factorized volume paths (HEX and WEDGE tensor products with face-modal HEX lifts and traces,
  TET and PYR collapsed coordinates) equal the dense path (and the oracle) to 1e-12;
- the quad face maps in tensor form (every HEX face and WEDGE quad face operator) equal the dense
  face-modal rows (and the oracle) to 1e-12.
"""
import numpy as np
import pytest

from dpgsolver_amd import mixed_mesh as MM
from dpgsolver_amd.mixed_solver import MixedSolver, build_tables
from oracle.mixed_oracle import MixedOracle

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL_MAX = 1e-11


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel_err(a, b):
    emax = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    assert emax < TOL_MAX, f"max-norm relative error {emax:.3e}"
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _state(mesh, seed):
    u = MM.project(mesh, lambda x: MM.entropy_wave(x))
    return u + 1e-3 * np.random.default_rng(seed).standard_normal(u.shape)


CASES = [((4, 3, 3), 1, "lf", None), ((5, 3, 4), 2, "roe", None), ((3, 4, 3), 3, "roe", 5), ((6, 3, 3), 4, "lf", 1)]


@pytest.mark.parametrize("n,seed,flux,pf", CASES)
def test_mixed_vs_oracle(n, seed, flux, pf):
    _require_gpu()
    print(f"case: mixed mesh n={n} seed={seed} flux={flux} p={pf or '1..5'}", flush=True)
    mesh = MM.build_mixed_mesh(n, seed=seed, p_fixed=pf)
    u = _state(mesh, seed)
    orc = MixedOracle(mesh, num_flux=0 if flux == "lf" else 1)
    s = MixedSolver(mesh, num_flux=flux)
    s.set_state(u)
    assert np.array_equal(s.get_state(), u)
    for minv in (False, True):
        r = s.rhs(apply_minv=minv)
        assert rel_err(r, orc.residual(u, apply_minv=minv)) < TOL, minv
    dt = MM.dt_cfl(mesh)
    s.step(dt, 3)
    ref = orc.step(dt, u, 3)
    assert rel_err(s.get_state(), ref) < TOL
    s.close()


def test_launch_forms_and_tiles_bit_identical():
    _require_gpu()
    print("case: launch forms / tiles", flush=True)
    mesh = MM.build_mixed_mesh((5, 4, 3), seed=7)
    u = _state(mesh, 7)
    tab = build_tables(mesh)
    s = MixedSolver(mesh, num_flux="roe", tables=tab)
    s.set_state(u)
    r0 = s.rhs()
    dt = MM.dt_cfl(mesh)
    s.step(dt, 2)
    u0 = s.get_state()
    _, tiles0, _ = s.get_launch()
    for mode in ("per_group", "persistent", "grouped"):
        for scale in (1, 2, 4):
            tiles = np.maximum(tiles0 // scale, 1)
            for bpc in (1, 3):
                s.set_launch(mode, tiles, bpc)
                s.set_state(u)
                assert np.array_equal(s.rhs(), r0), (mode, scale, bpc)
                s.step(dt, 2)
                assert np.array_equal(s.get_state(), u0), (mode, scale, bpc)
    s.set_state(u)
    best = s.autotune(reps=2)
    assert best.shape == (tab.n_group,) and np.all(best > 0)
    s.set_state(u)
    assert np.array_equal(s.rhs(), r0)
    # graph replays: 1 + 1 steps == 2 steps
    s.step(dt, 1)
    s.step(dt, 1)
    assert np.array_equal(s.get_state(), u0)
    us = s.time_stage(dt, 6)
    assert np.all(us > 0)
    s.set_state(u)
    assert np.array_equal(s.rhs(), r0)
    s.close()


def test_sum_factorized_vs_dense():
    _require_gpu()
    print("case: sum factorization vs dense", flush=True)
    mesh = MM.build_mixed_mesh((6, 3, 4), seed=11)
    u = _state(mesh, 11)
    a = MixedSolver(mesh, num_flux="roe", tables=build_tables(mesh, sum_factorized=True))
    b = MixedSolver(mesh, num_flux="roe", tables=build_tables(mesh, sum_factorized=False))
    orc = MixedOracle(mesh, num_flux=1)
    for s in (a, b):
        s.set_state(u)
    for minv in (False, True):
        ra, rb = a.rhs(apply_minv=minv), b.rhs(apply_minv=minv)
        assert rel_err(ra, rb) < TOL
        assert rel_err(ra, orc.residual(u, apply_minv=minv)) < TOL
    dt = MM.dt_cfl(mesh)
    a.step(dt, 2)
    b.step(dt, 2)
    assert rel_err(a.get_state(), b.get_state()) < TOL
    a.close()
    b.close()


def test_tensor_face_maps_vs_dense_rows():
    _require_gpu()
    print("case: tensor-form quad face maps vs dense face-modal rows", flush=True)
    mesh = MM.build_mixed_mesh((5, 4, 4), seed=13)
    u = _state(mesh, 13)
    tab = build_tables(mesh)
    s = MixedSolver(mesh, num_flux="roe", tables=tab)
    # every face operator with square face-modal rows (HEX faces, WEDGE quad faces) factors
    widths = [int(w) for off, w in zip(tab.arrays["fop_red_off"], tab.arrays["fop_red_n"]) if off >= 0]
    n_quad = sum(1 for w in widths if int(round(w ** 0.5)) ** 2 == w)
    assert n_quad > 0 and s.tensor_face_ops() == n_quad, (s.tensor_face_ops(), n_quad)
    orc = MixedOracle(mesh, num_flux=1)
    dt = MM.dt_cfl(mesh)
    res = {}
    for tensor in (True, False):
        s.set_face_maps(tensor)
        s.set_state(u)
        r = [s.rhs(apply_minv=m) for m in (False, True)]
        s.step(dt, 2)
        res[tensor] = (r, s.get_state())
    for m in range(2):
        assert rel_err(res[True][0][m], res[False][0][m]) < TOL
        assert rel_err(res[True][0][m], orc.residual(u, apply_minv=bool(m))) < TOL
    assert rel_err(res[True][1], res[False][1]) < TOL
    # the two forms sum in different orders: equal results would mean the switch changed nothing
    assert not np.array_equal(res[True][0][0], res[False][0][0])
    s.close()
