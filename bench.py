"""Benchmark of the explicit DG residual + SSP-RK3 stage on one MI355X.

Workload (default): 3-D Euler, DG p=3, 64^3 curved periodic HEX mesh on [0,2pi]^3 (sinusoidal node
perturbation 0.05h, seed 0), Taylor-Green initial state at Mach 0.1, Roe-Pike flux, SSP-RK3 with the
positivity limiter, fixed dt from CFL 0.5.  One bench step = one SSP-RK3 time step = 3 stages, launched
as one captured hipGraph through the C-ABI (dgx_step); inputs are resident in HBM before timing.

Prints one JSON line:
  value    = DOF-stages/s (DOF = E * (p+1)^3 * 5), whole job
  roofline = the stage kernel against HBM: algorithmic bytes per stage launch / measured stage time
  cpu_baseline = the oracle (dense CPU restatement of the reference algorithm, 1 thread) on a bounded
             sample of the same workload (a 16^3 mesh of the same construction), in DOF-stages/s
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


def cpu_baseline(args, d, p, num_flux):
    """Time the oracle's rk stages on a 16^3 (3-D) / 32^2 (2-D) sample, single thread."""
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
    u, up = u0.copy(), np.zeros_like(u0)
    dof = mesh.n_elem * (p + 1) ** d * (d + 2)
    stages = 0
    t0 = time.perf_counter()
    while True:
        orc.rk_stage(stages % 3, dt, u, up)
        stages += 1
        el = time.perf_counter() - t0
        if el > args.cpu_seconds and stages % 3 == 0:
            break
    return {"value": dof * stages / el, "unit": "DOF-stages/s", "cores": 1, "kind": "port",
            "sample": f"{sample}; {stages} stages in {el:.1f} s (oracle, dense per-element operators)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="hex64", choices=["hex64", "hex32", "vortex2d"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-n", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
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

    t_setup = time.perf_counter()
    if args.config == "vortex2d":
        d, p, num_flux = 2, 2, 0
        mesh, geo, u0 = dg.vortex_2d(32, p)
        workload = "2D Euler QUAD p=2 32^2 periodic, isentropic vortex, Lax-Friedrichs, SSP-RK3"
    else:
        d, p, num_flux = 3, 3, 1
        n = 64 if args.config == "hex64" else 32
        mesh, geo, u0 = dg.tgv_3d(n, p, seed=args.seed)
        workload = (f"3D Euler HEX p=3 {n}^3 curved periodic [0,2pi]^3, Taylor-Green M=0.1, Roe-Pike, "
                    f"SSP-RK3 + positivity")
    dt = dg.dt_cfl(mesh, p, u0, 0.5)
    solver = dg.Solver(mesh, geo, p, num_flux=num_flux, precision=args.precision)
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
    traffic = read_traffic(args.traffic_file) if args.config == "hex64" else None
    print(f"[bench] {args.steps} steps in {ms:.2f} ms (wall {wall * 1e3:.2f} ms), stage {t_stage_s * 1e6:.1f} us, "
          f"max_rhs {mx:.4e}", file=sys.stderr, flush=True)

    out = {
        "metric": "DOF-stages/s Euler HEX p=3: mixed vs FP64 at <=1e-9 err; % HBM roofline",
        "value": value,
        "unit": "DOF-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32-mfma",
        "data": "synthetic",
        "config": {"workload": workload, "elements": mesh.n_elem, "dof": dof, "p": p,
                   "stages_per_step": 3, "precision": args.precision, "dt": dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": (f"dgx::hex3_stage_kernel<{num_flux}>" if (d, p) == (3, 3)
                                else f"dgx::stage_kernel<{d},{p},{num_flux}>"),
                     "algorithmic_bytes_per_launch": bytes_alg, "bytes_per_element": b_elem,
                     "stage_us": t_stage_s * 1e6},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args, d, p, num_flux)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
