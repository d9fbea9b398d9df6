"""GPU parity of the explicit time steppers and of the stop after a positivity failure.

Time steppers (test_case->solver_type_e, solve_explicit.c:150-170): forward Euler (time_step_euler
:226-250) and the low-storage RK 5-stage 4th-order scheme (time_step_ls_rk_54 :303-350), next to the
SSP-RK3 default, through dgx_step (captured graphs) and dgx_rk_stage, against the oracle's stages on the
same mesh, geometry and state.  Cases cover every stage kernel: the generic Euler kernel (2-D p=2), the
HEX p=3 group kernel (bricks and padding slots; it updates the state in place for these steppers
instead of rotating buffers), the generic Navier-Stokes kernels (3-D p=2) and the HEX p=4 Navier-Stokes
kernel.  Bar: relative L2 <= 1e-12, max-norm <= 1e-11 (as the other parity tests).

Positivity failure: the reference aborts in the failing stage (solve.c:204-208).  Here the failing stage
completes on every element (the failing element keeps its un-limited stage value, as the oracle leaves
it) and every later stage, also inside one multi-step dgx_step call, leaves the state untouched: the
state returned equals the failing stage's output.
"""
import ctypes
import math

import numpy as np
import pytest

import dpgsolver_amd as dg
from oracle.oracle import Oracle

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


def _case(kind):
    """(mesh, geo, p, u0, viscous, dt_scale) of one stage-kernel family."""
    L = 2 * math.pi
    if kind == "quad2":
        mesh, geo, u0 = dg.vortex_2d(6, 2)
        return mesh, geo, 2, u0, None, 0.5
    if kind == "hex3":
        # 5 x 4 x 4 cells: bricks plus packed groups with padding slots
        mesh = dg.structured_mesh(3, (5, 4, 4), (0, 0, 0), (L, L, L), p_g=3, amp=0.05, seed=7)
        u0 = dg.ic_taylor_green(mesh, 3) + 1e-3 * dg.randn(11, (mesh.n_elem, 5, 64))
        return mesh, dg.geometry(mesh, 3), 3, u0, None, 0.5
    p = 2 if kind == "ns2" else 4
    mesh = dg.structured_mesh(3, (2, 3, 2), (0, 0, 0), (L, L, L), p_g=p, amp=0.05, seed=3)
    u0 = dg.ic_taylor_green(mesh, p, 0.3)
    u0 = u0 + 1e-5 * dg.randn(103, u0.shape)
    return mesh, dg.geometry(mesh, p), p, u0, (0.05, 0.71), 0.1


@pytest.mark.parametrize("kind", ["quad2", "hex3", "ns2", "ns4"])
@pytest.mark.parametrize("scheme", ["euler", "ls_rk_54"])
def test_time_stepper_vs_oracle(kind, scheme):
    """4 steps through dgx_step vs the oracle's stages; then one step stage by stage through
    dgx_rk_stage, bit-identical to a captured step from the same state."""
    _require_gpu()
    print(f"case: time stepper {scheme} on {kind}", flush=True)
    mesh, geo, p, u0, visc, cfl = _case(kind)
    flux = 1 if kind != "quad2" else 0
    dt = cfl * dg.dt_cfl(mesh, p, u0)
    if scheme == "euler":
        dt *= 0.25            # forward Euler is stable only for a small fraction of the RK3 step
    orc = Oracle(mesh, p, num_flux=flux, geometry=geo, viscous=visc)
    orc.set_time_stepper(scheme)
    u_ref, up = u0.copy(), np.zeros_like(u0)
    mx_ref = orc.step(dt, u_ref, up, nsteps=4)
    s = dg.Solver(mesh, geo, p, num_flux=flux, viscous=visc, time_stepper=scheme)
    assert s.n_stages == orc.n_stages
    s.set_state(u0)
    mx = s.step(dt, 4)
    e = rel_err(s.get_state(), u_ref)
    print(f"{scheme} {kind}: 4 steps rel L2 {e:.3e}", flush=True)
    assert e < TOL
    assert abs(mx - mx_ref) <= 1e-10 * mx_ref
    # stage by stage == captured step (same kernels, same order)
    t = dg.Solver(mesh, geo, p, num_flux=flux, viscous=visc, time_stepper=scheme)
    u4 = s.get_state()
    t.set_state(u4)
    s.set_state(u4)
    for k in range(t.n_stages):
        t.rk_stage(k, dt)
    s.step(dt, 1)
    assert np.array_equal(t.get_state(), s.get_state())
    with pytest.raises(dg._ffi.DgxError):
        t.rk_stage(t.n_stages, dt)       # no such stage
    s.close()
    t.close()


def _failing_state(kind):
    mesh, geo, p, u0, visc, _ = _case(kind)
    u = u0.copy()
    e = 5
    d = mesh.d
    u[e, 0] = -0.5                       # negative average density, zero momentum, negative energy
    u[e, 1:1 + d] = 0.0
    u[e, d + 1] = -1.0
    return mesh, geo, p, u, visc, e


@pytest.mark.parametrize("kind", ["quad2", "hex3", "ns4"])
@pytest.mark.parametrize("scheme", ["ssp_rk_33", "ls_rk_54"])
def test_failure_stops_multi_step_call(kind, scheme):
    """A positivity failure in the first stage of a 3-step dgx_step: DGX_ERR_POSITIVITY with the
    element, and the state equals the failing stage's output (dgx_rk_stage alone, and the oracle's
    stage, in which the failing element keeps its un-limited value); later calls leave it unchanged."""
    _require_gpu()
    print(f"case: failure stop {scheme} on {kind}", flush=True)
    mesh, geo, p, u, visc, e_bad = _failing_state(kind)
    flux = 0    # Lax-Friedrichs: Roe-Pike's sqrt(rho_R / rho_L) of a negative density is NaN, which no
                # positivity check reports (NaN compares false), so the failure would go unnoticed
    dt = 1e-6
    orc = Oracle(mesh, p, num_flux=flux, geometry=geo, viscous=visc)
    orc.set_time_stepper(scheme)
    u_ref, up = u.copy(), np.zeros_like(u)
    _, bad = orc.rk_stage(0, dt, u_ref, up)
    assert bad == e_bad
    a = dg.Solver(mesh, geo, p, num_flux=flux, viscous=visc, time_stepper=scheme)
    b = dg.Solver(mesh, geo, p, num_flux=flux, viscous=visc, time_stepper=scheme)
    a.set_state(u)
    b.set_state(u)
    with pytest.raises(dg._ffi.DgxError) as ei:
        a.rk_stage(0, dt)
    assert ei.value.status == 3
    with pytest.raises(dg._ffi.DgxError) as ei:
        b.step(dt, 3)
    assert ei.value.status == 3
    elem = ctypes.c_int64(-1)
    assert b._lib.dgx_sync(b._ctx, ctypes.byref(elem)) == 3 and elem.value == e_bad
    ua, ub = a.get_state(), b.get_state()
    assert np.array_equal(ua, ub)
    assert rel_err(ua, u_ref) < TOL
    # the rest of a's step and another multi-step call change nothing
    for k in range(1, a.n_stages):
        with pytest.raises(dg._ffi.DgxError):
            a.rk_stage(k, dt)
    with pytest.raises(dg._ffi.DgxError):
        a.step(dt, 2)
    assert np.array_equal(a.get_state(), ua)
    # a new state clears the failure
    a.set_state(u_ref)
    a2 = u_ref.copy()
    a2[e_bad] = u[(e_bad + 1) % mesh.n_elem]
    a.set_state(a2)
    a.step(dt, 1)
    a.close()
    b.close()


def test_solve_explicit_loop_vs_oracle():
    """Solver.solve_explicit (the reference's loop: clamp of the last step by time + dt > time_final -
    1e3 EPS, exit on time == time_final) equals the same loop over the oracle's steps: the same step
    sizes (the last one a fraction of dt) and the same final state; and an exit tolerance above the first
    step's max_rhs stops after one step on both."""
    _require_gpu()
    print("case: solve_explicit loop vs oracle loop", flush=True)
    from dpgsolver_amd.solver import explicit_time_loop
    mesh, geo, p, u0, visc, cfl = _case("quad2")
    dt = cfl * dg.dt_cfl(mesh, p, u0)
    tf = 3.4 * dt
    orc = Oracle(mesh, p, num_flux=0, geometry=geo)
    u_ref, up = u0.copy(), np.zeros_like(u0)
    dts = []

    def orc_step(h):
        dts.append(h)
        return orc.step(h, u_ref, up, 1)

    t_ref, mx_ref, n_ref = explicit_time_loop(orc_step, dt, tf)
    assert n_ref == 4 and t_ref == tf and 0 < dts[-1] < dt
    s = dg.Solver(mesh, geo, p, num_flux="lf")
    s.set_state(u0)
    t, mx, n = s.solve_explicit(dt, tf)
    assert (t, n) == (t_ref, n_ref)
    assert abs(mx - mx_ref) <= 1e-11 * abs(mx_ref)
    assert rel_err(s.get_state(), u_ref) < TOL
    s.set_state(u0)
    t, mx, n = s.solve_explicit(dt, tf, exit_tol=10 * mx_ref + 1e3)
    assert n == 1 and t == dt
    s.close()
