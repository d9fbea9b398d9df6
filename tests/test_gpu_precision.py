"""Precision form (DGX_PREC_MIXED) of the HEX p=3 stage kernel, through the C-ABI.

The face traces and the weak derivative contract, along every 4-node line, the FP64 differences of the
data from the line's first node, as FP32 hi/lo parts with FP32 operator entries and FP32 accumulation
(the line's constant part goes through FP64: trace rows sum to 1, the weighted derivative rows to their
FP64 column sums); the face lift and everything else (fluxes, Roe eigen-decomposition and entropy fix,
positivity, metric terms, residual accumulation, RK update) stays FP64.  These tests check that the form
runs and how far it is from the FP64 path:
- its residual error against FP64 is the FP32 rounding of the differenced contractions relative to the
  residual (measured on these 8^3 states: 4.4e-6 at M = 0.1, 4.0e-6 at M = 0.3; 9.4e-5 on the 64^3
  benchmark state; FP32 sums of the undifferenced values gave 3e-2 there);
- it never touches the FP64 path: FP64 results are bit-identical whether or not a mixed context ran;
- the Navier-Stokes HEX p=4 form (own face traces, volume gradient and weak divergence differenced the
  same way) stays close to its FP64 run on the Re = 1600 Taylor-Green state;
- it is refused where it is not built (other orders, Navier-Stokes at p != 4).
"""
import math

import numpy as np
import pytest

import dpgsolver_amd as dg

pytestmark = pytest.mark.gpu


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_mixed_residual_error():
    """Residual error of the precision form against FP64 on 8^3 Taylor-Green states at M = 0.1 and 0.3,
    bounded by 2x the measured values (4.4e-6 and 4.0e-6): the line differencing takes the size of the
    pressure (1/(gamma M^2)) out of the FP32 sums, so the error no longer grows like 1/M^2."""
    _require_gpu()
    errs = {}
    for mach in (0.1, 0.3):
        print(f"case: precision form 8^3 mach={mach}", flush=True)
        mesh, geo, u0 = dg.tgv_3d(8, 3, seed=0, mach=mach)
        u = u0 + 1e-3 * dg.randn(5, u0.shape)
        a = dg.Solver(mesh, geo, 3, num_flux="roe")
        b = dg.Solver(mesh, geo, 3, num_flux="roe", precision="mixed")
        a.set_state(u)
        b.set_state(u)
        r64, _ = a.compute_rhs()
        rmx, _ = b.compute_rhs(precision="mixed")
        errs[mach] = float(np.linalg.norm(rmx - r64) / np.linalg.norm(r64))
        print(f"mixed vs fp64 residual rel L2 {errs[mach]:.3e}", flush=True)
        # the FP64 path is untouched by the mixed context
        r64b, _ = a.compute_rhs()
        assert np.array_equal(r64, r64b)
        a.close()
        b.close()
    assert 0 < errs[0.1] <= 8.8e-6
    assert 0 < errs[0.3] <= 8.0e-6


def test_mixed_steps_stay_close():
    """10 SSP-RK3 steps: the state error of the precision form (measured 4.9e-10) stays below 1e-9."""
    _require_gpu()
    mesh, geo, u0 = dg.tgv_3d(8, 3, seed=0, mach=0.1)
    dt = dg.dt_cfl(mesh, 3, u0)
    a = dg.Solver(mesh, geo, 3, num_flux="roe")
    b = dg.Solver(mesh, geo, 3, num_flux="roe", precision="mixed")
    a.set_state(u0)
    b.set_state(u0)
    a.step(dt, 10)
    b.step(dt, 10)
    ua, ub = a.get_state(), b.get_state()
    err = np.linalg.norm(ub - ua) / np.linalg.norm(ua)
    print(f"state rel L2 after 10 steps {err:.3e}", flush=True)
    assert 0 < err < 1e-9
    a.close()
    b.close()


def test_ns_p4_mixed_close_to_fp64():
    """Navier-Stokes HEX p=4 precision form on the 6^3 Taylor-Green state at Re = 1600, M = 0.1: residual
    error against FP64 and the state error after 10 SSP-RK3 steps."""
    _require_gpu()
    print("case: ns p=4 precision form 6^3", flush=True)
    visc = (1.0 / 1600.0, 0.71)
    L = 2 * math.pi
    mesh = dg.structured_mesh(3, (6, 6, 6), (0, 0, 0), (L, L, L), p_g=1)
    geo = dg.geometry(mesh, 4)
    u = dg.ic_taylor_green(mesh, 4, 0.1)
    a = dg.Solver(mesh, geo, 4, num_flux="roe", viscous=visc)
    b = dg.Solver(mesh, geo, 4, num_flux="roe", viscous=visc, precision="mixed")
    a.set_state(u)
    b.set_state(u)
    r64, _ = a.compute_rhs()
    rmx, _ = b.compute_rhs(precision="mixed")
    err = float(np.linalg.norm(rmx - r64) / np.linalg.norm(r64))
    print(f"ns p=4 mixed vs fp64 residual rel L2 {err:.3e}", flush=True)
    dt = dg.dt_cfl(mesh, 4, u)
    a.step(dt, 10)
    b.step(dt, 10)
    ua, ub = a.get_state(), b.get_state()
    es = float(np.linalg.norm(ub - ua) / np.linalg.norm(ua))
    print(f"ns p=4 mixed vs fp64 state rel L2 after 10 steps {es:.3e}", flush=True)
    assert 0 < err < 1e-3
    assert 0 < es < 1e-8
    a.close()
    b.close()


def test_mixed_refused_where_not_built():
    _require_gpu()
    mesh, geo, u0 = dg.vortex_2d(4, 2)
    with pytest.raises(dg._ffi.DgxError):
        dg.Solver(mesh, geo, 2, num_flux="lf", precision="mixed")
    s = dg.Solver(mesh, geo, 2, num_flux="lf")
    s.set_state(u0)
    with pytest.raises(dg._ffi.DgxError):
        s.compute_rhs(precision="mixed")
    s.close()
    L = 2 * math.pi
    m3 = dg.structured_mesh(3, (2, 2, 2), (0, 0, 0), (L, L, L), p_g=3)
    with pytest.raises(dg._ffi.DgxError):
        dg.Solver(m3, dg.geometry(m3, 3), 3, num_flux="roe", precision="mixed", viscous=(0.01, 0.71))
