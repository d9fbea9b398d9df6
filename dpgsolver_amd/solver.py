"""Host-side mirror of the reference's explicit solver entry points, over the C-ABI of include/dgx.h.

Reference entry points and their counterparts here:
  compute_rhs(sim)                  src/simulation/solvers/solve.c:133-149          -> Solver.compute_rhs
  time_step_ssp_rk_33 (one stage)   src/simulation/solvers/solve_explicit.c:252-301 -> Solver.rk_stage
  solve_explicit(sim)               src/simulation/solvers/solve_explicit.c:78-118  -> Solver.solve_explicit
  enforce_positivity_highorder      src/simulation/solvers/solve.c:187-233          (inside every stage)
The reference aborts on errors (EXIT_ERROR); here every failure raises DgxError with the C status.
All device work is done by libdgx.so; there is no other compute path.
"""
import ctypes
import math

import numpy as np

from . import _ffi

NUM_FLUX = {"lax_friedrichs": _ffi.FLUX_LAX_FRIEDRICHS, "lf": _ffi.FLUX_LAX_FRIEDRICHS,
            "roe_pike": _ffi.FLUX_ROE_PIKE, "roe": _ffi.FLUX_ROE_PIKE}
PRECISION = {"fp64": _ffi.PREC_FP64, "mixed": _ffi.PREC_MIXED}
# test_case->solver_type_e (solve_explicit.c:150-170)
SOLVER_TYPE_E = {"ssp_rk_33": _ffi.SOLVER_E_SSP_RK_33, "euler": _ffi.SOLVER_E_EULER,
                 "ls_rk_54": _ffi.SOLVER_E_LS_RK_54}


EPS = 1.0e-15   # the reference's EPS (definitions_tol.h)


def explicit_time_loop(step, dt, time_final, exit_tol=0.0, solver_proc="E", exit_ratio=0.0):
    """The loop of solve_explicit (solve_explicit.c:92-116) and check_exit (:192-222): before each step,
    if time + dt > time_final - 1e3 EPS then dt = time_final - time (and stays so); time += dt; one
    step of the stepper (step(dt) returns its max_rhs); the first step's max_rhs is max_rhs0; exit when
    max_rhs < exit_tol, and for solver_proc "E" when time == time_final, for "EI" when
    max_rhs / max_rhs0 < exit_ratio.  Returns (time, max_rhs, steps)."""
    if time_final < 0.0:
        raise ValueError("time_final must be >= 0")
    if solver_proc not in ("E", "EI"):
        raise ValueError("solver_proc must be 'E' or 'EI'")
    time = 0.0
    max_rhs0 = 0.0
    t_step = 0
    while True:
        if time + dt > time_final - 1e3 * EPS:
            dt = time_final - time
        time += dt
        max_rhs = step(dt)
        if t_step == 0:
            max_rhs0 = max_rhs
        t_step += 1
        exit_now = max_rhs < exit_tol
        if solver_proc == "E":
            exit_now = exit_now or time == time_final
        elif max_rhs0 > 0.0 and max_rhs / max_rhs0 < exit_ratio:
            exit_now = True
        if exit_now:
            return time, max_rhs, t_step


class Solver:
    """Explicit DG (Euler, QUAD/HEX, GL/GL collocated) on one GPU.

    mesh, geo: dpgsolver_amd.problem.Mesh / Geometry; p: solution order; num_flux: "lf" | "roe";
    positivity: apply enforce_positivity_highorder after every stage (the reference does for Euler and
    Navier-Stokes, test_case.c:63-79); viscous: None (Euler) or (mu, Pr) for Navier-Stokes with the
    BR2-stable viscous flux; time_stepper: "ssp_rk_33" | "euler" | "ls_rk_54" (solver_type_e);
    operators: None (built for p) or a dpgsolver_amd._ffi.DgxOperators holding the GL/GL 1-D factors.
    """

    def __init__(self, mesh, geo, p, num_flux="roe", positivity=True, precision="fp64", viscous=None,
                 time_stepper="ssp_rk_33", operators=None):
        self._lib = _ffi.load()
        self.d, self.p = mesh.d, p
        self.nv = mesh.d + 2
        self.ns = (p + 1) ** mesh.d
        self.n_elem = mesh.n_elem
        cfg = _ffi.DgxConfig()
        cfg.d, cfg.p, cfg.pde = mesh.d, p, 0
        cfg.num_flux = NUM_FLUX[num_flux] if isinstance(num_flux, str) else int(num_flux)
        cfg.positivity = 1 if positivity else 0
        cfg.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
        if viscous is not None:
            cfg.pde = _ffi.PDE_NAVIER_STOKES
            cfg.mu, cfg.Pr = float(viscous[0]), float(viscous[1])
        cfg.solver_type_e = SOLVER_TYPE_E[time_stepper] if isinstance(time_stepper, str) else int(time_stepper)
        self._mesh_c = mesh.c_struct()
        self._geo_c = geo.c_struct()
        self._keep = (mesh, geo)
        ctx = ctypes.c_void_p()
        ops = ctypes.byref(operators) if operators is not None else None
        st = self._lib.dgx_create(ctypes.byref(cfg), ctypes.byref(self._mesh_c), ctypes.byref(self._geo_c), ops,
                                  ctypes.byref(ctx))
        _ffi.check(st)
        self._ctx = ctx
        self.n_stages = self._lib.dgx_stages_per_step(ctx)

    # -- lifetime ------------------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.dgx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        _ffi.check(st, self._ctx, self._lib)

    @property
    def shape(self):
        return (self.n_elem, self.nv, self.ns)

    @property
    def stream(self):
        return self._lib.dgx_get_stream(self._ctx)

    # -- state ---------------------------------------------------------------------------------------
    def set_state(self, u):
        """u: numpy (host) or torch tensor (host or cuda) in the reference layout [E][nv][Ns]."""
        ptr, kind, keep = _buffer(u, self.shape)
        self._check(self._lib.dgx_set_state(self._ctx, _ffi.LAYOUT_REF, kind, ptr))
        del keep

    def get_state(self, out=None):
        out = np.empty(self.shape) if out is None else out
        ptr, kind, keep = _buffer(out, self.shape, writable=True)
        self._check(self._lib.dgx_get_state(self._ctx, _ffi.LAYOUT_REF, kind, ptr))
        return out

    # -- residual and stepping -----------------------------------------------------------------------
    def compute_rhs(self, apply_minv=True, precision="fp64", out=None):
        """Residual of the current state; returns (rhs [E][nv][Ns], max_rhs)."""
        out = np.empty(self.shape) if out is None else out
        ptr, kind, keep = _buffer(out, self.shape, writable=True)
        mx = ctypes.c_double(0.0)
        self._check(self._lib.dgx_rhs(self._ctx, PRECISION[precision], int(bool(apply_minv)), _ffi.LAYOUT_REF,
                                      kind, ptr, ctypes.byref(mx)))
        return out, mx.value

    def max_rhs(self, apply_minv=True):
        """max over volumes of ||rhs||_inf of the current state (compute_max_rhs_dg_like), without
        downloading the residual; apply_minv=False: of the residual before M^-1."""
        mx = ctypes.c_double(0.0)
        self._check(self._lib.dgx_rhs(self._ctx, _ffi.PREC_FP64, int(bool(apply_minv)), _ffi.LAYOUT_REF,
                                      _ffi.PTR_HOST, None, ctypes.byref(mx)))
        return mx.value

    def rk_stage(self, stage, dt):
        mx = ctypes.c_double(0.0)
        self._check(self._lib.dgx_rk_stage(self._ctx, int(stage), float(dt), ctypes.byref(mx)))
        return mx.value

    def step(self, dt, nsteps=1):
        """nsteps steps of the time stepper as hipGraph launches; returns max_rhs of the last stage."""
        mx = ctypes.c_double(0.0)
        self._check(self._lib.dgx_step(self._ctx, float(dt), int(nsteps), ctypes.byref(mx)))
        return mx.value

    def solve_explicit(self, dt, time_final, exit_tol=0.0, solver_proc="E", exit_ratio=0.0):
        """Time loop of solve_explicit (solve_explicit.c:78-118) on this solver's steps; see
        explicit_time_loop.  Returns (time, max_rhs of the last step, steps taken)."""
        return explicit_time_loop(lambda h: self.step(h, 1), dt, time_final, exit_tol, solver_proc, exit_ratio)

    # -- linearization -----------------------------------------------------------------------------
    def jvp(self, w, apply_minv=False, out=None):
        """J w, J = dR/du at the current state (R before M^-1, or M^-1 R with apply_minv); [E][nv][Ns]."""
        return self._lin(self._lib.dgx_jvp, w, apply_minv, out)

    def vjp(self, v, apply_minv=False, out=None):
        """J^T v at the current state (same J as jvp); [E][nv][Ns]."""
        return self._lin(self._lib.dgx_vjp, v, apply_minv, out)

    def _lin(self, fn, x, apply_minv, out):
        if out is None:
            out = _like(x, self.shape)
        pin, kin, keep_in = _buffer(x, self.shape)
        pout, kout, keep_out = _buffer(out, self.shape, writable=True)
        if kin != kout:
            raise ValueError("input and output must both be host or both be device buffers")
        self._check(fn(self._ctx, int(bool(apply_minv)), _ffi.LAYOUT_REF, kin, pin, pout))
        del keep_in, keep_out
        return out

    def jacobian_blocks(self, out, apply_minv=True):
        """Transposed volume-diagonal blocks of J (of M^-1 R with apply_minv) at the current state into
        out, a float64 or float32 cuda tensor [E][nv*Ns][nv*Ns]: out[e][j][i] = dR_i/du_j
        (dgx_jacobian_blocks / dgx_jacobian_blocks_f32)."""
        n = self.nv * self.ns
        import torch
        if not (isinstance(out, torch.Tensor) and out.is_cuda and out.dtype in (torch.float64, torch.float32)
                and out.is_contiguous() and tuple(out.shape) == (self.n_elem, n, n)):
            raise ValueError(f"out must be a contiguous float64/float32 cuda tensor of shape {(self.n_elem, n, n)}")
        torch.cuda.synchronize()
        fn = self._lib.dgx_jacobian_blocks if out.dtype == torch.float64 else self._lib.dgx_jacobian_blocks_f32
        self._check(fn(self._ctx, int(bool(apply_minv)), ctypes.c_void_p(out.data_ptr())))
        return out

    def sync(self):
        elem = ctypes.c_int64(-1)
        st = self._lib.dgx_sync(self._ctx, ctypes.byref(elem))
        self._check(st)

    @property
    def last_elapsed_ms(self):
        return self._lib.dgx_last_elapsed_ms(self._ctx)


def _like(x, shape):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return torch.empty(shape, dtype=torch.float64, device=x.device)
    except ImportError:  # pragma: no cover
        pass
    return np.empty(shape)


def residual_function():
    """torch.autograd.Function R(u) of a Solver: forward = dgx_rhs at u, backward = dgx_vjp.

    Usage: R = residual_function(); r = R.apply(u, solver, apply_minv) with u a float64 cuda tensor
    [E][nv][Ns] requiring grad; r.backward(g) gives u.grad = J^T g (J at u).  The solver's state is set
    to u by the forward call.
    """
    import torch

    class _Residual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u, solver, apply_minv=False):
            u = u.detach().contiguous()
            solver.set_state(u)
            r = torch.empty_like(u)
            solver.compute_rhs(apply_minv=apply_minv, out=r)
            ctx.solver, ctx.apply_minv = solver, apply_minv
            ctx.save_for_backward(u)
            return r

        @staticmethod
        def backward(ctx, g):
            (u,) = ctx.saved_tensors
            s = ctx.solver
            s.set_state(u)
            return s.vjp(g.contiguous(), apply_minv=ctx.apply_minv), None, None

        @staticmethod
        def jvp(ctx, w, *_):
            return ctx.solver.jvp(w.contiguous(), apply_minv=ctx.apply_minv)

    return _Residual


def _buffer(a, shape, writable=False):
    """(pointer, DGX_PTR_*, keep-alive) for a numpy array or torch tensor of `shape`, float64, contiguous."""
    try:
        import torch
    except ImportError:  # pragma: no cover
        torch = None
    if torch is not None and isinstance(a, torch.Tensor):
        if a.dtype != torch.float64 or not a.is_contiguous() or tuple(a.shape) != tuple(shape):
            raise ValueError(f"expected a contiguous float64 tensor of shape {shape}")
        kind = _ffi.PTR_DEVICE if a.is_cuda else _ffi.PTR_HOST
        return ctypes.c_void_p(a.data_ptr()), kind, a
    a_np = a
    if not isinstance(a_np, np.ndarray) or a_np.dtype != np.float64 or not a_np.flags["C_CONTIGUOUS"]:
        if writable:
            raise ValueError("output must be a contiguous float64 numpy array")
        a_np = np.ascontiguousarray(a, dtype=np.float64)
    if tuple(a_np.shape) != tuple(shape):
        raise ValueError(f"expected shape {shape}, got {a_np.shape}")
    return ctypes.c_void_p(a_np.ctypes.data), _ffi.PTR_HOST, a_np
