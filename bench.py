"""Benchmark of the explicit DG residual + SSP-RK3 stage on one MI355X.

Workload (default): 3-D Euler, DG p=3, 64^3 curved periodic HEX mesh on [0,2pi]^3 (sinusoidal node
perturbation 0.05h, seed 0), Taylor-Green initial state at Mach 0.1, Roe-Pike flux, SSP-RK3 with the
positivity limiter, fixed dt from CFL 0.5.  One bench step = one SSP-RK3 time step = 3 stages, launched
as one captured hipGraph through the C-ABI (dgx_step); inputs are resident in HBM before timing.

Prints one JSON line:
  value    = DOF-stages/s (DOF = E * (p+1)^3 * 5), whole job
  roofline = the stage kernel against HBM: algorithmic bytes per stage launch / measured stage time,
             against the 8 TB/s peak (frac) and against a measured 2 GiB device copy (frac_of_measured_copy)
  cpu_baseline = the oracle (CPU restatement of the reference algorithm) on a bounded sample of the
             same workload (a 16^3 mesh of the same construction), in DOF-stages/s: sum-factorized element
             operators on all host threads (value, cores) and on one thread; dense per-element operators
             (the reference's form) on all host threads (dense) and on one thread (single_thread; the
             reference is single-threaded)
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_element(d, p):
    """B_elem = 3 S + G_v + G_f + C_f (state read + write + RK register, metrics + det J, face normals
    + face det J with each face counted once, face connectivity)."""
    ns = (p + 1) ** d
    nf = (p + 1) ** (d - 1)
    nv = d + 2
    faces_per_elem = d          # periodic structured mesh: F = d E
    S = ns * nv * 8
    Gv = ns * (d * d + 1) * 8
    Gf = faces_per_elem * nf * (d + 1) * 8
    Cf = faces_per_elem * 12
    return 3 * S + Gv + Gf + Cf


def read_traffic(path):
    """HBM bytes per stage launch from a committed rocprofv3 PMC summary (line 'traffic_bytes: N')."""
    try:
        with open(path) as f:
            for ln in f:
                if ln.strip().startswith("traffic_bytes:"):
                    return float(ln.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def host_threads():
    """Host threads for the all-cores baseline: the CPUs this process may run on, at most
    OMP_NUM_THREADS and at most 16 (the host share of one GPU on the measurement machines)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, min(n, 16))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(args, d, p, num_flux):
    """Time the oracle's rk stages on a 16^3 (3-D) / 32^2 (2-D) sample in its two operator forms: the
    dense per-element operators (the reference's form; one thread, as the reference runs, and all host
    threads) and the sum-factorized form (Kronecker factors applied one direction at a time, diagonal
    M^-1; all host threads and one thread).  `value` is the fastest of them, sum-factorized on all host
    threads (OpenMP over volumes, faces in two phases)."""
    import numpy as np
    import dpgsolver_amd as dg
    from oracle.oracle import Oracle

    if d == 3:
        n = args.cpu_sample_n
        mesh, geo, u0 = dg.tgv_3d(n, p, seed=args.seed)
        sample = f"{n}^3 curved HEX p={p}, TGV M=0.1, Roe-Pike, SSP-RK3 + positivity"
    else:
        mesh, geo, u0 = dg.vortex_2d(32, p)
        sample = f"32^2 QUAD p={p}, isentropic vortex, LF, SSP-RK3 + positivity"
    orc = Oracle(mesh, p, num_flux=num_flux, geometry=geo)
    dt = dg.dt_cfl(mesh, p, u0)
    dof = mesh.n_elem * (p + 1) ** d * (d + 2)

    def rate(threads, seconds, sf):
        orc.set_threads(threads)
        orc.set_sum_factorized(sf)
        u, up = u0.copy(), np.zeros_like(u0)
        stages = 0
        t0 = time.perf_counter()
        while True:
            orc.rk_stage(stages % 3, dt, u, up)
            stages += 1
            el = time.perf_counter() - t0
            if el > seconds and stages % 3 == 0:
                return dof * stages / el, stages, el

    nt = host_threads()
    part = args.cpu_seconds / 4
    d1 = rate(1, part, False)
    dn = rate(nt, part, False) if nt > 1 else d1
    f1 = rate(1, part, True)
    fn = rate(nt, part, True) if nt > 1 else f1
    orc.set_sum_factorized(False)
    host = f"{cpu_model()}, {os.cpu_count()} logical CPUs, {len(os.sched_getaffinity(0))} available"

    def sub(r, threads, form):
        return {"value": r[0], "cores": threads, "sample": f"{r[1]} stages in {r[2]:.1f} s ({form})"}

    return {"value": fn[0], "unit": "DOF-stages/s", "cores": nt, "kind": "port",
            "sample": f"{sample}; {fn[1]} stages in {fn[2]:.1f} s on {nt} threads (oracle, sum-factorized "
                      f"element operators; {host})",
            "sum_factorized_single_thread": sub(f1, 1, "sum-factorized"),
            "dense": sub(dn, nt, "dense per-element operators"),
            "single_thread": sub(d1, 1, "dense per-element operators, the reference's form")}


PROTOCOL_STEPS = 50   # SSP-RK3 steps of the state comparisons of the 64^3 workload


def precision_form(args, dg, torch, mesh, geo, u0, dt, fp64_solver, fp64_stage_ms):
    """The precision form (line-differenced FP32 hi/lo contractions, FP32 accumulation) against the FP64
    run just timed: stage time ratio over the same timed steps, residual error at the initial state,
    state error after PROTOCOL_STEPS steps from u0 (both forms), and FP64's own round-off floor (residual
    change under a 1e-16 relative input perturbation)."""
    dev = torch.device("cuda:0")
    ut = torch.from_numpy(u0).to(dev)
    mixed = dg.Solver(mesh, geo, 3, num_flux="roe", precision="mixed")
    mixed.set_state(ut)
    if args.warmup > 0:
        mixed.step(dt, args.warmup)
    mixed.step(dt, args.steps)
    mixed_stage_ms = mixed.last_elapsed_ms / (3 * args.steps)
    mixed.set_state(ut)
    mixed.step(dt, PROTOCOL_STEPS)
    u_mixed = mixed.get_state(torch.empty_like(ut))
    mixed.set_state(ut)
    r_mixed = torch.empty_like(ut)
    mixed.compute_rhs(precision="mixed", out=r_mixed)
    mixed.close()
    fp64_solver.set_state(ut)
    fp64_solver.step(dt, PROTOCOL_STEPS)
    u_fp64 = fp64_solver.get_state(torch.empty_like(ut))
    fp64_solver.set_state(ut)
    r64 = torch.empty_like(ut)
    fp64_solver.compute_rhs(out=r64)
    gen = torch.Generator(device="cpu").manual_seed(0)
    fp64_solver.set_state(ut * (1 + 1.1e-16 * torch.randn(ut.shape, generator=gen, dtype=torch.float64).to(dev)))
    r_eps = torch.empty_like(ut)
    fp64_solver.compute_rhs(out=r_eps)

    def rel(a, b):
        return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))

    return {"kernel": "dgx::hex3_stage_kernel<1,true>", "speedup_vs_fp64": fp64_stage_ms / mixed_stage_ms,
            "stage_us": mixed_stage_ms * 1e3, "residual_rel_l2": rel(r_mixed, r64),
            "state_rel_l2": rel(u_mixed, u_fp64), "state_steps": PROTOCOL_STEPS,
            "fp64_roundoff_floor_rel_l2": rel(r_eps, r64), "target_rel_err": 1e-9,
            "meets_target": rel(r_mixed, r64) <= 1e-9}


def ns_precision_form(args, dg, torch, mesh, geo, u0, dt, viscous, fp64_solver, fp64_stage_ms, fp64_ke):
    """The Navier-Stokes HEX p=4 precision form (line-differenced FP32 hi/lo contractions of the own face
    traces, the volume gradient and the weak divergence) against the FP64 run: stage time ratio, residual
    error at the initial state, and over the KE_STEPS-step protocol the kinetic-energy history's largest
    relative deviation and the final state error, next to FP64's own residual round-off floor."""
    dev = torch.device("cuda:0")
    ut = torch.from_numpy(u0).to(dev)
    mixed = dg.Solver(mesh, geo, 4, num_flux="roe", viscous=viscous, precision="mixed")
    mixed.set_state(ut)
    if args.warmup > 0:
        mixed.step(dt, args.warmup)
    mixed.step(dt, args.steps)
    mixed_stage_ms = mixed.last_elapsed_ms / (3 * args.steps)
    mixed.set_state(ut)
    r_mixed = torch.empty_like(ut)
    mixed.compute_rhs(precision="mixed", out=r_mixed)
    ke_m = ke_history(dg, torch, mixed, u0, dt, KE_STEPS, 10)
    u_mixed = mixed.get_state(torch.empty_like(ut))
    mixed.close()
    fp64_solver.set_state(ut)
    fp64_solver.step(dt, KE_STEPS)
    u_fp64 = fp64_solver.get_state(torch.empty_like(ut))
    fp64_solver.set_state(ut)
    r64 = torch.empty_like(ut)
    fp64_solver.compute_rhs(out=r64)
    gen = torch.Generator(device="cpu").manual_seed(0)
    fp64_solver.set_state(ut * (1 + 1.1e-16 * torch.randn(ut.shape, generator=gen, dtype=torch.float64).to(dev)))
    r_eps = torch.empty_like(ut)
    fp64_solver.compute_rhs(out=r_eps)

    def rel(a, b):
        return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))

    ke_dev = max(abs(a - b) / abs(b) for a, b in zip(ke_m["history"], fp64_ke["history"]))
    return {"kernel": "dgx::ns4_stage_kernel<1,true>", "speedup_vs_fp64": fp64_stage_ms / mixed_stage_ms,
            "stage_us": mixed_stage_ms * 1e3, "residual_rel_l2": rel(r_mixed, r64),
            "state_rel_l2": rel(u_mixed, u_fp64), "state_steps": KE_STEPS,
            "kinetic_energy_max_rel_dev": ke_dev, "kinetic_energy_history": ke_m["history"],
            "fp64_roundoff_floor_rel_l2": rel(r_eps, r64), "target_rel_err": 1e-9}


def fp64_vs_oracle(args, dg, torch, mesh, geo, u0, dt, p, num_flux, solver, viscous=None, steps=0, minv=False,
                   pre_minv=True):
    """GPU FP64 against the oracle at the benchmark's full size: the pre-M^-1 residual (pre_minv) and the
    M^-1 residual (minv) at the initial state, and (steps > 0) the state after that many SSP-RK3 steps, as
    relative L2 errors, next to the GPU's own round-off floor (the residual change under a 1e-16 relative
    input perturbation).  The oracle runs on all host threads in its sum-factorized form (the same
    operators as its dense form, to round-off; M^-1 diagonal), which makes the full sizes take seconds."""
    import numpy as np
    from oracle.oracle import Oracle
    t0 = time.perf_counter()
    orc = Oracle(mesh, p, num_flux=num_flux, geometry=geo, viscous=viscous)
    orc.set_threads(host_threads())
    orc.set_sum_factorized(True)

    def rel(a, b):
        return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))

    out = {}
    solver.set_state(u0)
    kinds = ([False] if pre_minv else []) + ([True] if minv else [])
    for mv in kinds:
        r_gpu, _ = solver.compute_rhs(apply_minv=mv)
        r_orc, _ = orc.residual(u0, apply_minv=mv)
        out["residual_rel_l2" if mv else "residual_pre_minv_rel_l2"] = rel(r_gpu, r_orc)
    rng = np.random.default_rng(0)
    solver.set_state(u0 * (1 + 1.1e-16 * rng.standard_normal(u0.shape)))
    r_eps, _ = solver.compute_rhs(apply_minv=kinds[-1])
    out["gpu_roundoff_floor_rel_l2"] = rel(r_eps, r_gpu)
    if steps > 0:
        u, up = u0.copy(), np.zeros_like(u0)
        orc.step(dt, u, up, steps)
        solver.set_state(u0)
        solver.step(dt, steps)
        out["state_rel_l2"] = rel(solver.get_state(), u)
        out["state_steps"] = steps
    out["oracle_threads"] = host_threads()
    out["oracle_form"] = "sum-factorized"
    out["seconds"] = time.perf_counter() - t0
    return out


def measured_copy_gbs(torch, nbytes=1 << 31, reps=10):
    """Device-to-device copy rate (read + write bytes / time) of torch's vectorized elementwise kernel
    (dst = 1.0 * src, 16-byte accesses) on nbytes: the achievable HBM bandwidth the roofline fraction is also
    quoted against.  (dst.copy_(src) goes through the runtime's blit kernel, ~13 % slower on this part.)"""
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def kinetic_energy(u):
    return float((0.5 * (u[:, 1] ** 2 + u[:, 2] ** 2 + u[:, 3] ** 2) / u[:, 0]).sum())


KE_STEPS = 200   # SSP-RK3 steps of the Navier-Stokes kinetic-energy history


def ke_history(dg, torch, solver, u0, dt, steps, every):
    """Kinetic energy sum_nodes rho |u|^2 / 2 (nodal sum over the whole mesh) every `every` SSP-RK3 steps
    over `steps` steps from u0, computed from the device state."""
    ut = torch.from_numpy(u0).cuda()
    solver.set_state(ut)
    buf = torch.empty_like(ut)

    def ke():
        solver.get_state(buf)
        return float((0.5 * (buf[:, 1] ** 2 + buf[:, 2] ** 2 + buf[:, 3] ** 2) / buf[:, 0]).sum())

    hist = [ke()]
    for _ in range(steps // every):
        solver.step(dt, every)
        hist.append(ke())
    return {"every_steps": every, "steps": steps, "dt": dt, "history": hist}


def ns_small_vs_oracle(dg, steps):
    """The Navier-Stokes protocol (TGV Re=1600, M=0.1, HEX p=4, Roe + BR2, SSP-RK3) at the size the host
    oracle (sum-factorized form, all host threads) finishes in about a minute (10^3): state and
    kinetic-energy error after `steps` steps, next to the GPU round-off floor of the state (the same steps
    from a 1e-16 relative perturbation of the start)."""
    import numpy as np
    from oracle.oracle import Oracle
    n, p, visc = 10, 4, (1.0 / 1600.0, 0.71)
    mesh = dg.structured_mesh(3, (n, n, n), (0, 0, 0), (2 * math.pi,) * 3, p_g=1)
    geo = dg.geometry(mesh, p)
    u0 = dg.ic_taylor_green(mesh, p, 0.1)
    dt = dg.dt_cfl(mesh, p, u0, 0.5)
    orc = Oracle(mesh, p, num_flux=1, geometry=geo, viscous=visc)
    orc.set_threads(host_threads())
    orc.set_sum_factorized(True)
    u, up = u0.copy(), np.zeros_like(u0)
    orc.step(dt, u, up, steps)
    s = dg.Solver(mesh, geo, p, num_flux="roe", viscous=visc)
    s.set_state(u0)
    s.step(dt, steps)
    ug = s.get_state()
    rng = np.random.default_rng(1)
    s.set_state(u0 * (1 + 1.1e-16 * rng.standard_normal(u0.shape)))
    s.step(dt, steps)
    ue = s.get_state()
    s.close()

    def rel(a, b):
        return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))

    return {"mesh": f"{n}^3", "steps": steps, "state_rel_l2": rel(ug, u), "gpu_roundoff_floor_rel_l2": rel(ue, ug),
            "kinetic_energy_rel": abs(kinetic_energy(ug) - kinetic_energy(u)) / kinetic_energy(u)}


def bench_vjp(args, dg, torch):
    """J^T v of the pre-M^-1 residual at 32^3 HEX p=3 (Taylor-Green M=0.1 + 1e-3 N(0,1) noise), device
    buffers; one unit = one J^T v over the whole mesh."""
    mesh, geo, u0 = dg.tgv_3d(32, 3, seed=args.seed)
    u = u0 + 1e-3 * dg.randn(1234, u0.shape)
    s = dg.Solver(mesh, geo, 3, num_flux="roe")
    s.set_state(torch.from_numpy(u).cuda())
    v = torch.from_numpy(dg.randn(10_000, u.shape)).cuda()
    out = torch.empty_like(v)
    for _ in range(max(args.warmup, 1)):
        s.vjp(v, out=out)
    ms = []
    for _ in range(args.steps):
        s.vjp(v, out=out)
        ms.append(s.last_elapsed_ms)       # HIP events around upload, traces, adjoint kernel
    t = sorted(ms)[len(ms) // 2] * 1e-3
    b_elem = algorithmic_bytes_per_element(3, 3)
    achieved = b_elem * mesh.n_elem / t / 1e9
    print(json.dumps({"metric": "J^T v of the residual, 32^3 HEX p=3 Euler", "value": 1.0 / t, "unit": "VJP/s",
                      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                      "data": "synthetic",
                      "config": {"workload": "residual adjoint 32^3 curved HEX p=3, TGV M=0.1 + 1e-3 noise, Roe-Pike",
                                 "elements": mesh.n_elem},
                      "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                                   "kernel": "dgx::lin_kernel<3,3,1,true> (+ layout and trace kernels)"}}),
          flush=True)


def bench_implicit(args, dg, torch):
    """Pseudo-transient Newton-Krylov (dpgsolver_amd.implicit) on a 3-D Euler HEX p=3 n^3 periodic mesh,
    perturbed M=2 uniform flow, Roe-Pike: `--newton-steps` Newton steps from the same initial state, each
    = the volume-diagonal blocks (dgx_jacobian_blocks_f32) + their inversion in place
    (dgx_block_inverse_f32) + GMRES(60) to rtol 1e-4 (at most 600 iterations) preconditioned by block
    Jacobi (`--nk-sweep gauss-seidel`: one multicolour block Gauss-Seidel sweep).  One unit = one GMRES
    iteration (one JVP, one preconditioner application, CGS2 over the basis)."""
    from dpgsolver_amd import implicit
    n, p = args.nk_n, 3
    mesh = dg.structured_mesh(3, (n, n, n), (0, 0, 0), (2 * math.pi,) * 3, p_g=1)
    geo = dg.geometry(mesh, p)
    u0 = dg.ic_uniform_flow(mesh, p, amp=1e-2)
    s = dg.Solver(mesh, geo, p, num_flux="roe")

    class Timed:
        """The solver, with the JVP's HIP-event time (layout copy, traces, lin kernel) accumulated."""

        def __init__(self, inner):
            self.inner, self.shape = inner, inner.shape
            self.jvp_ms, self.jvp_calls = 0.0, 0
            self.blocks_ms, self.blocks_calls = 0.0, 0

        def __getattr__(self, name):
            return getattr(self.inner, name)

        def jvp(self, w, apply_minv=False, out=None):
            r = self.inner.jvp(w, apply_minv=apply_minv, out=out)
            self.jvp_ms += self.inner.last_elapsed_ms
            self.jvp_calls += 1
            return r

        def jacobian_blocks(self, out, apply_minv=True):
            r = self.inner.jacobian_blocks(out, apply_minv=apply_minv)
            self.blocks_ms += self.inner.last_elapsed_ms
            self.blocks_calls += 1
            return r

    class TimedPrec(implicit.BlockJacobi):
        def __init__(self, *a, **k):
            super().__init__(*a, **k)
            self.update_s, self.apply_s, self.applies = 0.0, 0.0, 0

        def update(self, inv_dt):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            super().update(inv_dt)
            torch.cuda.synchronize()
            self.update_s += time.perf_counter() - t0

        def apply(self, y, transpose=False):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = super().apply(y, transpose)
            torch.cuda.synchronize()
            self.apply_s += time.perf_counter() - t0
            self.applies += 1
            return r

    def run(steps, timed):
        s.set_state(torch.from_numpy(u0).cuda())
        pc = TimedPrec(timed, mesh, "cuda", sweep=args.nk_sweep)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, hist = implicit.solve_implicit(timed, mesh, p, exit_tol=0.0, max_steps=steps, precond=pc)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, hist, pc

    run(1, Timed(s))                                   # warm-up: allocations, library kernels
    timed = Timed(s)
    wall, hist, pc = run(args.newton_steps, timed)
    its = sum(h[2] for h in hist)
    jvp_ms = timed.jvp_ms / max(timed.jvp_calls, 1)
    b_elem = algorithmic_bytes_per_element(3, p)
    achieved = b_elem * mesh.n_elem / (jvp_ms * 1e-3) / 1e9
    gmres_s = wall - pc.update_s
    dof = mesh.n_elem * s.ns * s.nv
    blk_bytes = mesh.n_elem * (s.nv * s.ns) ** 2 * 4
    apply_ms = pc.apply_s / max(pc.applies, 1) * 1e3
    out = {"metric": "GMRES iterations/s of the matrix-free Newton-Krylov solve, HEX p=3 Euler",
           "value": its / gmres_s, "unit": "GMRES iterations/s", "n_gpus": 1, "steps": args.newton_steps,
           "warmup": 1, "ms_per_step": wall / args.newton_steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64 (f32 preconditioner blocks)",
           "data": "synthetic",
           "config": {"workload": f"implicit Newton-GMRES {n}^3 HEX p=3 periodic [0,2pi]^3, M=2 uniform flow + "
                                  f"1e-2 perturbation, Roe-Pike, block {args.nk_sweep} (f32), GMRES(60) rtol 1e-4",
                      "elements": mesh.n_elem, "dof": dof, "newton_steps": args.newton_steps,
                      "gmres_iterations": its, "max_rhs": [h[0] for h in hist]},
           "breakdown_ms": {"newton_step": wall / args.newton_steps * 1e3,
                            "precond_update_per_step": pc.update_s / args.newton_steps * 1e3,
                            "gmres_per_iteration": gmres_s / max(its, 1) * 1e3,
                            "jacobian_blocks_per_step": timed.blocks_ms / max(timed.blocks_calls, 1),
                            "jvp_per_call": jvp_ms, "precond_apply_per_call": apply_ms,
                            "jvp_calls": timed.jvp_calls},
           "precond_apply": {"bytes": blk_bytes, "achieved": blk_bytes / (apply_ms * 1e-3) / 1e9,
                             "unit": "GB/s", "kernel": "dgx::block_apply_kernel (FP32 blocks, FP64 vectors)"},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "dgx::lin_kernel<3,3,1,false> (+ layout and trace kernels)"}}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_jvp(args, dg, mesh, geo, u0, p, timed)
    print(json.dumps(out), flush=True)


def bench_geometry(args, dg, torch):
    """Geometry set-up (metric terms, det J, normals, face det J) of the n^3 curved HEX p=3 mesh (p_g = 3)
    on the GPU, device outputs; one unit = one build of the whole mesh.  The CPU baseline is the host
    build (dgx_geometry_build, OpenMP over volumes) on a bounded sample of the same construction."""
    n, p = args.geo_n, 3
    L = 2 * math.pi
    mesh = dg.structured_mesh(3, (n, n, n), (0, 0, 0), (L, L, L), p_g=p, amp=0.05, seed=args.seed)
    dg.geometry_gpu(dg.structured_mesh(3, (2, 2, 2), (0, 0, 0), (L, L, L), p_g=p, amp=0.05), p, device=True)
    kms, walls = [], []
    for _ in range(max(args.steps, 1)):
        tm = {}
        t0 = time.perf_counter()
        g = dg.geometry_gpu(mesh, p, device=True, timing=tm)
        walls.append(time.perf_counter() - t0)
        kms.append(tm["kernel_ms"])
        del g
    k_ms = sorted(kms)[len(kms) // 2]
    wall = sorted(walls)[len(walls) // 2]
    E, F = mesh.n_elem, mesh.n_face
    Ng, Nc, Nf = (p + 1) ** 3, (p + 1) ** 3, (p + 1) ** 2
    bytes_alg = 8 * (E * 3 * Ng + E * 9 * Nc + E * Nc + F * Nf * 4) + 8 * F + F * 2
    achieved = bytes_alg / (k_ms * 1e-3) / 1e9
    out = {"metric": "geometry set-up of a curved HEX p=3 mesh on the GPU", "value": E / (wall), "unit": "volumes/s",
           "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": wall * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"geometry of the {n}^3 curved HEX mesh (p_g = 3, 0.05h sinusoidal perturbation) "
                                  f"for p = 3: metric terms (3-D curl form), det J, normals, face det J",
                      "elements": E, "faces": F, "kernel_ms": k_ms,
                      "wall_includes": "host face grouping, upload of the geometry nodes, kernel"},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "dgx::geometry_kernel<3>",
                        "algorithmic_bytes_per_launch": bytes_alg}}
    if not args.no_cpu_baseline:
        ns = args.cpu_sample_n
        sm = dg.structured_mesh(3, (ns, ns, ns), (0, 0, 0), (L, L, L), p_g=p, amp=0.05, seed=args.seed)
        builds = 0
        t0 = time.perf_counter()
        while True:
            dg.geometry(sm, p)
            builds += 1
            el = time.perf_counter() - t0
            if el > args.cpu_seconds:
                break
        omp = os.environ.get("OMP_NUM_THREADS", "")
        cores = int(omp) if omp.isdigit() and int(omp) > 0 else os.cpu_count()
        out["cpu_baseline"] = {"value": sm.n_elem * builds / el, "unit": "volumes/s", "cores": cores,
                               "kind": "port",
                               "sample": f"host dgx_geometry_build (C, OpenMP, {cores} threads) on the {ns}^3 mesh "
                                         f"of the same construction, {builds} builds in {el:.1f} s ({cpu_model()})"}
    print(json.dumps(out), flush=True)


def cpu_baseline_jvp(args, dg, mesh, geo, u0, p, timed):
    """The oracle's complex-step JVP of M^-1 R (the CPU restatement's linearization) on the same mesh
    and state, on all host threads, for about --cpu-seconds; unit JVP/s."""
    from oracle.oracle import Oracle
    orc = Oracle(mesh, p, num_flux=1, positivity=False, geometry=geo)
    nt = host_threads()
    orc.set_threads(nt)
    w = dg.randn(5, u0.shape)
    calls = 0
    t0 = time.perf_counter()
    while True:
        orc.jvp_cs(u0, w, apply_minv=True)
        calls += 1
        el = time.perf_counter() - t0
        if el > args.cpu_seconds:
            break
    gpu = timed.jvp_calls / (timed.jvp_ms * 1e-3) if timed.jvp_ms > 0 else None
    return {"value": calls / el, "unit": "JVP/s", "cores": nt, "kind": "port",
            "sample": f"{calls} complex-step JVPs of M^-1 R in {el:.1f} s on {nt} threads, same {mesh.n_elem}-element "
                      f"mesh and state ({cpu_model()})",
            "gpu_jvp_per_s": gpu}


FP64_PEAK_TFLOPS = 78.6
MIXED_N = (60, 45, 45)


def mixed_stage_work(mesh, tab):
    """Algorithmic HBM bytes and a flop scale of one stage of the batching form.
    Bytes: per element the state read, written and the RK register read (3 x 5 Ns x 8) + metrics and
    det J (10 x 8); per face the unit normal, face Jacobian and connectivity (5 x 8).
    Flops, for scale only (the dense per-element operator form; the sum-factorized kernels execute fewer):
    volume 2 x 5 Nc Ns (interpolation) + 2 x 15 Nc Ns (weak divergence) + 3 x 22 Nc (fluxes); per face
    and side 2 x 5 Nq Ns (trace) + 2 x 5 Nq Ns (lift); per face node ~140 (Roe-Pike)."""
    a = tab.arrays
    byts = 0.0
    flops = 0.0
    for g in range(tab.n_group):
        cnt = int(a["group_start"][g + 1] - a["group_start"][g])
        ns, nc = int(a["group_ns"][g]), int(a["group_nc"][g])
        byts += cnt * (3 * 5 * ns * 8 + 80)
        flops += cnt * (2 * 5 * nc * ns + 2 * 15 * nc * ns + 3 * 22 * nc)
    byts += mesh.n_face * 40
    nq = a["fop_nq"][a["face_op"][:, 0]].astype(float)
    nsl = a["fop_ns"][a["face_op"][:, 0]].astype(float)
    nsr = a["fop_ns"][a["face_op"][:, 1]].astype(float)
    flops += float((nq * (20 * (nsl + nsr) + 140)).sum())
    return byts, flops


def bench_mixed(args, dg, torch):
    """Batching form: 3-D Euler on a seeded conforming mixed HEX/WEDGE/TET/PYR mesh of the unit cube
    (about 200k elements, 40/30/20/10 % by type), per-element p uniform in 1..5, entropy wave, Roe-Pike,
    SSP-RK3; one step = 3 stages replayed as one captured hipGraph.  Measures the three autotuned launch
    forms (persistent work-counter grid, one grouped launch, per-(type, p) launches) and times the steps of
    value with the fastest."""
    import numpy as np
    from dpgsolver_amd import mixed_mesh as MM
    from dpgsolver_amd.mixed_solver import MixedSolver, build_tables
    t0 = time.perf_counter()
    n = tuple(args.mixed_n)
    mesh = MM.build_mixed_mesh(n, seed=args.seed)
    tab = build_tables(mesh)
    u0 = MM.project(mesh, lambda x: MM.entropy_wave(x))
    dt = MM.dt_cfl(mesh)
    s = MixedSolver(mesh, num_flux="roe", tables=tab)
    s.set_state(u0)
    print(f"[bench] mixed setup {time.perf_counter() - t0:.1f} s, E={mesh.n_elem}, dof={s.n_dof}, "
          f"groups={tab.n_group}, dt={dt:.3e}", file=sys.stderr, flush=True)
    default_tiles = s.get_launch()[1].tolist()
    group_us = s.autotune(reps=3)
    mode, tiles, bpc = s.get_launch()
    s.set_state(u0)
    forms = {}
    for form in ("per_group", "grouped", "persistent"):
        s.set_launch(form)
        us = s.time_stage(dt, 30)
        forms[form] = {"stage_us": float(us[0]), "face_kernel_us": float(us[1]), "volume_kernel_us": float(us[2])}
    # the timed steps run the fastest form (every form is one hipGraph per step)
    single = min(forms, key=lambda f: forms[f]["stage_us"])
    s.set_launch(single)
    if args.warmup > 0:
        s.step(dt, args.warmup)
    s.set_state(u0)
    torch.cuda.synchronize()
    tw = time.perf_counter()
    s.step(dt, args.steps)
    wall = time.perf_counter() - tw
    stages = 3 * args.steps
    # value: the timed steps (graph replays, the step call synchronizes its stream); roofline: that
    # form's stage time measured with HIP events on the solver's stream
    t_stage = forms[single]["stage_us"] * 1e-6
    value = s.n_dof * stages / wall
    byts, flops = mixed_stage_work(mesh, tab)
    ach_bw = byts / t_stage / 1e9
    ach_fl = flops / t_stage / 1e12
    u = s.get_state()
    out = {
        "metric": "DOF-stages/s Euler mixed HEX/WEDGE/TET/PYR p=1..5 (batching form); % roofline",
        "value": value,
        "unit": "DOF-stages/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"3D Euler mixed mesh {n[0]}x{n[1]}x{n[2]} boxes of the periodic unit cube, "
                                f"seed {args.seed}: " + ", ".join(f"{k} {v}" for k, v in mesh.counts().items())
                                + "; p uniform 1..5; entropy wave; Roe-Pike; SSP-RK3 (hipGraph)"),
                   "elements": mesh.n_elem, "faces": mesh.n_face, "dof": s.n_dof, "groups": tab.n_group, "dt": dt},
        "roofline": {"bound": "hbm", "achieved": ach_bw, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_bw / HBM_PEAK_GBS,
                     "traffic": read_traffic(os.path.join(HERE, "profiles", "stage_mixed_200k_traffic.md")),
                     "algorithmic_bytes_per_stage": byts, "stage_us": t_stage * 1e6,
                     "kernel": {"persistent": "vol_persistent + face_persistent (+ reset_counters)",
                                "grouped": "vol_grouped + face_range (+ reset_counters)",
                                "per_group": "vol_range per (type, p) group + face_range per face size class"}[single],
                     "launch_form": single,
                     "fp64_dense_equivalent_flops_per_stage": flops,
                     "fp64_dense_equivalent_TFLOPs": ach_fl, "fp64_peak_TFLOPs": FP64_PEAK_TFLOPS},
        "persistent_waves_per_simd": s.persistent_waves(),
        "launch_forms": forms,
        "persistent_vs_per_group": forms["per_group"]["stage_us"] / forms["persistent"]["stage_us"],
        "grouped_vs_per_group": forms["per_group"]["stage_us"] / forms["grouped"]["stage_us"],
        "autotune": {"default_tiles": default_tiles, "tiles": tiles.tolist(), "blocks_per_cu": bpc,
                     "groups": [{"type": ["HEX", "WEDGE", "TET", "PYR"][t], "p": p_,
                                 "elements": int(tab.arrays["group_start"][g + 1] - tab.arrays["group_start"][g]),
                                 "volume_us": float(group_us[g])}
                                for g, (t, p_) in enumerate(tab.group_keys)]},
        "state_finite": bool(np.isfinite(u).all()),
    }
    s.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_mixed(args)
    print(json.dumps(out), flush=True)


def cpu_baseline_mixed(args):
    """The mixed-mesh oracle (numpy over the oracle library's flux functions, one thread) on a bounded
    sample of the same construction: residual evaluations (the stage's work) for about cpu_seconds."""
    import numpy as np
    from dpgsolver_amd import mixed_mesh as MM
    from oracle.mixed_oracle import MixedOracle
    n = (8, 4, 4)
    mesh = MM.build_mixed_mesh(n, seed=args.seed)
    o = MixedOracle(mesh, num_flux=1)
    u = MM.project(mesh, lambda x: MM.entropy_wave(x))
    o.residual(u)
    t0 = time.perf_counter()
    k = 0
    while True:
        o.residual(u)
        k += 1
        el = time.perf_counter() - t0
        if el > args.cpu_seconds:
            break
    return {"value": o.n_dof * k / el, "unit": "DOF-stages/s", "cores": 1, "kind": "port",
            "sample": (f"mixed mesh {n[0]}x{n[1]}x{n[2]} boxes ({mesh.n_elem} elements, {o.n_dof} DOF, "
                       f"p 1..5), {k} residual evaluations in {el:.1f} s"),
            "cpu": cpu_model()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="hex64",
                    choices=["hex64", "hex32", "vortex2d", "ns48", "vjp32", "nk", "geo", "mixed"])
    ap.add_argument("--mixed-n", type=int, nargs=3, default=list(MIXED_N),
                    help="boxes per direction of the mixed mesh (x layers, y, z)")
    ap.add_argument("--geo-n", type=int, default=64)
    ap.add_argument("--nk-n", type=int, default=16)
    ap.add_argument("--newton-steps", type=int, default=4)
    ap.add_argument("--nk-sweep", default="jacobi", choices=["jacobi", "gauss-seidel"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-precision-form", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true",
                    help="skip the full-size FP64 residual comparison with the oracle")
    ap.add_argument("--oracle-check", action="store_true",
                    help="also compare the state after the protocol's steps with the oracle (minutes of host time)")
    ap.add_argument("--cpu-sample-n", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--traffic-file", default=os.path.join(HERE, "profiles", "stage_hex_p3_64_traffic.md"))
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the explicit residual path is built and measured on a single "
              "MI355X (multi-GPU sharding is design only). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import dpgsolver_amd as dg

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.config == "nk":
        return bench_implicit(args, dg, torch)
    if args.config == "geo":
        return bench_geometry(args, dg, torch)
    if args.config == "mixed":
        return bench_mixed(args, dg, torch)

    t_setup = time.perf_counter()
    viscous = None
    if args.config == "vortex2d":
        d, p, num_flux = 2, 2, 0
        mesh, geo, u0 = dg.vortex_2d(32, p)
        workload = "2D Euler QUAD p=2 32^2 periodic, isentropic vortex, Lax-Friedrichs, SSP-RK3"
    elif args.config == "ns48":
        d, p, num_flux = 3, 4, 1
        viscous = (1.0 / 1600.0, 0.71)
        mesh = dg.structured_mesh(3, (48, 48, 48), (0, 0, 0), (2 * math.pi,) * 3, p_g=1)
        geo = dg.geometry(mesh, p)
        u0 = dg.ic_taylor_green(mesh, p, 0.1)
        workload = ("3D Navier-Stokes HEX p=4 48^3 periodic [0,2pi]^3, Taylor-Green Re=1600 M=0.1 Pr=0.71, "
                    "Roe-Pike + BR2, SSP-RK3 + positivity")
    else:
        d, p, num_flux = 3, 3, 1
        n = 64 if args.config == "hex64" else 32
        if args.config == "vjp32":
            return bench_vjp(args, dg, torch)
        mesh, geo, u0 = dg.tgv_3d(n, p, seed=args.seed)
        workload = (f"3D Euler HEX p=3 {n}^3 curved periodic [0,2pi]^3, Taylor-Green M=0.1, Roe-Pike, "
                    f"SSP-RK3 + positivity")
    dt = dg.dt_cfl(mesh, p, u0, 0.5)
    solver = dg.Solver(mesh, geo, p, num_flux=num_flux, precision=args.precision, viscous=viscous)
    solver.set_state(u0)
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] setup {t_setup:.1f} s, E={mesh.n_elem}, dt={dt:.3e}", file=sys.stderr, flush=True)

    if args.warmup > 0:
        solver.step(dt, args.warmup)
    torch.cuda.synchronize()
    solver.sync()
    t0 = time.perf_counter()
    mx = solver.step(dt, args.steps)
    solver.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = solver.last_elapsed_ms        # HIP events on the solver stream around the graph launches
    ms = ev_ms if ev_ms > 0 else wall * 1e3
    stages = 3 * args.steps
    ns = (p + 1) ** d
    dof = mesh.n_elem * ns * (d + 2)
    value = dof * stages / (ms * 1e-3)
    t_stage_s = ms * 1e-3 / stages
    b_elem = algorithmic_bytes_per_element(d, p)
    bytes_alg = b_elem * mesh.n_elem
    achieved = bytes_alg / t_stage_s / 1e9
    traffic = (read_traffic(args.traffic_file) if args.config == "hex64" else
               read_traffic(os.path.join(HERE, "profiles", "stage_ns_hex_p4_48_traffic.md"))
               if args.config == "ns48" else None)
    print(f"[bench] {args.steps} steps in {ms:.2f} ms (wall {wall * 1e3:.2f} ms), stage {t_stage_s * 1e6:.1f} us, "
          f"max_rhs {mx:.4e}", file=sys.stderr, flush=True)

    out = {
        "metric": ("DOF-stages/s Euler HEX p=3: mixed vs FP64 at <=1e-9 err; % HBM roofline" if not viscous
                   else "DOF-stages/s Navier-Stokes HEX p=4 (BR2); % HBM roofline"),
        "value": value,
        "unit": "DOF-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": workload, "elements": mesh.n_elem, "dof": dof, "p": p,
                   "stages_per_step": 3, "precision": args.precision, "dt": dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ((f"dgx::ns4_stage_kernel<{num_flux}>" if (d, p) == (3, 4) else
                                 f"dgx::ns_gtrace_kernel<{d},{p}> + dgx::ns_stage_kernel<{d},{p},{num_flux}>")
                                if viscous else f"dgx::hex3_stage_kernel<{num_flux},{str(args.precision == 'mixed').lower()}>"
                                if (d, p) == (3, 3)
                                else f"dgx::stage_kernel<{d},{p},{num_flux}>"),
                     "algorithmic_bytes_per_launch": bytes_alg, "bytes_per_element": b_elem,
                     "stage_us": t_stage_s * 1e6},
    }
    copy_gbs = measured_copy_gbs(torch)
    out["roofline"]["measured_copy_gbs"] = copy_gbs
    out["roofline"]["frac_of_measured_copy"] = achieved / copy_gbs
    if args.config in ("hex64", "hex32") and args.precision == "fp64" and not args.no_precision_form:
        out["precision_form"] = precision_form(args, dg, torch, mesh, geo, u0, dt, solver, ms / stages)
    if args.config in ("hex64", "hex32") and args.precision == "fp64" and not args.no_oracle_check:
        # full-size FP64 parity of the residual before and after M^-1 (seconds of host time); the state
        # after PROTOCOL_STEPS steps (the oracle's stages at full size: a minute or more) with --oracle-check
        out["fp64_vs_oracle"] = fp64_vs_oracle(args, dg, torch, mesh, geo, u0, dt, p, num_flux, solver,
                                               steps=PROTOCOL_STEPS if args.oracle_check else 0, minv=True)
    if args.config == "ns48" and not args.no_oracle_check:
        # the full-size Navier-Stokes residual before and after M^-1 through the oracle (tens of seconds)
        out["fp64_vs_oracle"] = fp64_vs_oracle(args, dg, torch, mesh, geo, u0, dt, p, num_flux, solver,
                                               viscous=viscous, minv=True)
    if viscous:
        out["kinetic_energy"] = ke_history(dg, torch, solver, u0, dt, KE_STEPS, 10)
        if args.precision == "fp64" and not args.no_precision_form:
            out["precision_form"] = ns_precision_form(args, dg, torch, mesh, geo, u0, dt, viscous, solver,
                                                      ms / stages, out["kinetic_energy"])
        if args.oracle_check:
            out["small_mesh_vs_oracle"] = ns_small_vs_oracle(dg, KE_STEPS)
    if not args.no_cpu_baseline and not viscous:
        out["cpu_baseline"] = cpu_baseline(args, d, p, num_flux)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
