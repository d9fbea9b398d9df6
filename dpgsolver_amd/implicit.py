"""Matrix-free implicit solver: pseudo-transient Newton-Krylov on the explicit path's residual.

Mirrors the reference's implicit DG loop (solve_implicit.c:95-110; implicit_step :232-256; check_exit
:258-284; solve_and_update / update_coefs :393-468) with its CFL-ramped pseudo-time term
(compute_CFL_ramping and compute_dt_cfl_constrained, dg/solve_dg.c:222-318), and replaces the assembled
PETSc matrix and KSP by restarted GMRES on Jacobian-vector products of the residual (dgx_jvp, one kernel
launch per Krylov iteration; the state never leaves HBM).  Each step solves

    (M/dt - dR/du) du = R(u)   <=>   (I/dt - M^-1 dR/du) du = M^-1 R(u),    u <- u + du,

with dt per volume = cfl * h_v / max(|V| + c) over the volume's Bezier coefficients and
cfl = cfl_initial * max(1, max_rhs_0 / max_rhs).  The reference preconditions with ILU(1)
(external/petsc/options_files/petsc_options_gmres_default.txt); a matrix-free solver has no assembled
matrix to factor, so the preconditioner here is block Jacobi on the volume-diagonal blocks, which are
probed with JVPs on a colouring of the volumes and inverted in one batched call.

Transposed systems ((I/dt - M^-1 dR/du)^T psi = g, for output adjoints) use dgx_vjp and the transposed
blocks.  Euler only: dgx_jvp / dgx_vjp refuse Navier-Stokes contexts.
"""
import math

import numpy as np

GAMMA = 1.4


def gmres(matvec, b, rtol=1e-4, restart=60, max_iter=600, precond=None):
    """Restarted GMRES (Saad & Schultz 1986) with Givens rotations; Arnoldi by classical Gram-Schmidt
    applied twice (as stable as the modified form, and two matrix-vector products over the basis
    instead of k dot products, so one host synchronisation per iteration).  With `precond`, the
    flexible right-preconditioned form (Saad 1993): the preconditioned vectors z_k are kept and the
    update is x += Z y, so a preconditioner applied in lower precision (not exactly linear) does not
    limit the attainable residual.

    matvec maps a flat float64 tensor to one of the same shape (torch, CPU or GPU); b is the right-hand
    side.  Returns (x, relative residual ||b - A x|| / ||b||, iterations)."""
    import torch
    x = torch.zeros_like(b)
    bnorm = float(torch.linalg.norm(b))
    if bnorm == 0.0:
        return x, 0.0, 0
    V = torch.empty((restart + 1, b.numel()), dtype=b.dtype, device=b.device)
    Z = torch.empty((restart, b.numel()), dtype=b.dtype, device=b.device) if precond is not None else V
    r = b
    beta = bnorm
    it = 0
    while it < max_iter:
        V[0] = r / beta
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k_used = 0
        for k in range(restart):
            if precond is not None:
                Z[k] = precond(V[k])
            w = matvec(Z[k])
            Vk = V[:k + 1]
            h = torch.mv(Vk, w)
            w = w - torch.mv(Vk.t(), h)
            h2 = torch.mv(Vk, w)
            w = w - torch.mv(Vk.t(), h2)
            hk1 = torch.linalg.norm(w)
            col = torch.cat([h + h2, hk1.view(1)]).cpu().numpy()
            H[:k + 2, k] = col
            hk1 = float(col[-1])
            for j in range(k):
                t = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                H[j, k] = t
            den = math.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = (H[k, k] / den, H[k + 1, k] / den) if den > 0.0 else (1.0, 0.0)
            H[k, k], H[k + 1, k] = den, 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            it += 1
            k_used = k + 1
            if abs(g[k + 1]) <= rtol * bnorm or it >= max_iter or hk1 == 0.0:
                break
            V[k + 1] = w / hk1
        y = np.zeros(k_used)
        for i in range(k_used - 1, -1, -1):
            if H[i, i] != 0.0:
                y[i] = (g[i] - H[i, i + 1:k_used] @ y[i + 1:k_used]) / H[i, i]
        x = x + torch.mv(Z[:k_used].t(), torch.from_numpy(y).to(b.device))
        r = b - matvec(x)
        beta = float(torch.linalg.norm(r))
        if beta <= rtol * bnorm or beta == 0.0:
            break
    return x, beta / bnorm, it


def element_colors(mesh):
    """Greedy distance-1 colouring of the volumes over the face list: volumes that share a face get
    different colours.  Returns (colour per volume, number of colours)."""
    E = mesh.n_elem
    nbrs = [set() for _ in range(E)]
    for l, r in np.asarray(mesh.face_elem).reshape(-1, 2):
        if l == r:
            raise ValueError(f"volume {l} is its own face neighbour; no block colouring exists")
        nbrs[l].add(int(r))
        nbrs[r].add(int(l))
    col = np.full(E, -1, dtype=np.int64)
    for e in range(E):
        used = {col[n] for n in nbrs[e]}
        c = 0
        while c in used:
            c += 1
        col[e] = c
    return col, int(col.max()) + 1


def block_diagonal(solver, colors, n_colors, device):
    """Volume-diagonal blocks of J = d(M^-1 R)/du at the solver's current state, [E][nv*Ns][nv*Ns]
    (row and column index v*Ns + node), by probing: one JVP per (colour, variable, node) with unit
    entries on all volumes of that colour, no two of which share a face, so every volume's block column
    arrives free of its neighbours' contributions."""
    import torch
    E, nv, ns = solver.shape
    n = nv * ns
    B = torch.empty((E, n, n), dtype=torch.float64, device=device)
    w = torch.zeros(solver.shape, dtype=torch.float64, device=device)
    wf = w.view(E, n)
    jw = torch.empty_like(w)
    sync = _sync_fn(device)
    for c in range(n_colors):
        idx = torch.from_numpy(np.nonzero(np.asarray(colors) == c)[0]).to(device)
        for j in range(n):
            wf[idx, j] = 1.0
            sync()
            solver.jvp(w, apply_minv=True, out=jw)
            wf[idx, j] = 0.0
            B[idx, :, j] = jw.view(E, n).index_select(0, idx)
    return B


class BlockJacobi:
    """Block-Jacobi preconditioner of A = diag(inv_dt) - M^-1 dR/du on the volume-diagonal blocks, or,
    with sweep="gauss-seidel", one multicolour block Gauss-Seidel sweep over the volume colouring (each
    colour's blocks solved with the coupling to the colours already swept, obtained by one JVP of the
    partial solution: the matrix-free counterpart of an incomplete factorisation's lower triangle).

    The inverted blocks are kept in `dtype` (float32 by default): applying them is a batched GEMV that
    reads every block once per Krylov iteration, the largest HBM stream of the iteration, and the
    preconditioner's precision does not limit the solution's (GMRES measures its residual on A itself,
    with the FP64 JVP)."""

    def __init__(self, solver, mesh, device="cuda", dtype=None, sweep="jacobi"):
        import torch
        if sweep not in ("jacobi", "gauss-seidel"):
            raise ValueError(f"unknown sweep {sweep!r}")
        self.solver, self.device, self.sweep = solver, device, sweep
        self.dtype = torch.float32 if dtype is None else dtype
        self.colors, self.n_colors = element_colors(mesh)
        self.sel = [torch.from_numpy(np.nonzero(self.colors == c)[0].astype(np.int32)).to(device)
                    for c in range(self.n_colors)]
        self.Pinv = self.PinvT = None

    def update(self, inv_dt):
        """The blocks B at the solver's current state (one dgx_jacobian_blocks launch on the GPU; JVP
        probes otherwise) and the inverses of diag(inv_dt) - B per volume."""
        import torch
        E, nv, ns = self.solver.shape
        n = nv * ns
        idt = torch.as_tensor(inv_dt, dtype=torch.float64, device=self.device).reshape(-1)
        direct = torch.device(self.device).type == "cuda" and hasattr(self.solver, "jacobian_blocks")
        if direct and n <= 320 and self.dtype == torch.float32:
            # one FP32 buffer: the transposed blocks, inverted in place.  (shift I - B^T)^-1 =
            # ((shift I - B)^-1)^T is the column-major form the forward application reads; the row-major
            # inverse for transposed solves is made on first use.
            self.Pinv = self.PinvT = None
            BT = torch.empty((E, n, n), dtype=torch.float32, device=self.device)
            self.solver.jacobian_blocks(BT, apply_minv=True)
            self.PinvT = block_inverse(BT, idt.expand(E).contiguous(), out=BT)
            return
        if direct:
            B = torch.empty((E, n, n), dtype=torch.float64, device=self.device)
            self.solver.jacobian_blocks(B, apply_minv=True)
            B = B.transpose(1, 2).contiguous()
        else:
            B = block_diagonal(self.solver, self.colors, self.n_colors, self.device)
        # larger blocks (HEX p=4), other precisions and the CPU: the solver library, in batches of 64
        # blocks (its batched LU fails to allocate its workspace from 128 blocks of 320 x 320 up on
        # MI355X with ROCm 7)
        B.neg_()
        B.diagonal(dim1=1, dim2=2).add_(idt.view(-1, 1) if idt.numel() > 1 else idt)
        self.Pinv = torch.empty(B.shape, dtype=self.dtype, device=self.device)
        for i in range(0, E, 64):
            self.Pinv[i:i + 64] = torch.linalg.inv(B[i:i + 64])
        self.PinvT = None

    def apply(self, y, transpose=False):
        import torch
        if self.sweep == "gauss-seidel":
            return self._gauss_seidel(y, transpose)
        if self.PinvT is None:
            self.PinvT = self.Pinv.transpose(1, 2).contiguous()
        if transpose and self.Pinv is None:
            self.Pinv = self.PinvT.transpose(1, 2).contiguous()
        E, n = self.PinvT.shape[0], self.PinvT.shape[1]
        if self.PinvT.is_cuda and self.dtype == torch.float32 and n <= 320 and y.dtype == torch.float64:
            # dgx_block_apply reads column-major blocks: Pinv^T row-major for Pinv, Pinv itself for Pinv^T
            return block_apply(self.Pinv if transpose else self.PinvT, y)
        if self.Pinv is None:
            self.Pinv = self.PinvT.transpose(1, 2).contiguous()
        P = self.PinvT if transpose else self.Pinv
        return torch.bmm(P, y.reshape(E, n, 1).to(self.dtype)).reshape(-1).to(y.dtype)

    def _gauss_seidel(self, y, transpose):
        """x_c = A_cc^-1 (y_c - sum over swept colours c' of A_cc' x_c'), colours in order (A^T: reverse
        order, transposed blocks, couplings by VJP).  A_cc' = -(M^-1 J)_cc' off the diagonal blocks, so the
        coupling of the partial solution x (zero on the colours not yet swept) is + (M^-1 J x)_c."""
        import torch
        if self.PinvT is None:
            self.PinvT = self.Pinv.transpose(1, 2).contiguous()
        if transpose and self.Pinv is None:
            self.Pinv = self.PinvT.transpose(1, 2).contiguous()
        shape = self.solver.shape
        A_col = self.Pinv if transpose else self.PinvT
        lin = self.solver.vjp if transpose else self.solver.jvp
        order = range(self.n_colors - 1, -1, -1) if transpose else range(self.n_colors)
        y = y.reshape(-1)
        x = torch.zeros_like(y)
        jx = torch.empty(shape, dtype=torch.float64, device=y.device)
        sync = _sync_fn(y.device)
        for k, c in enumerate(order):
            if k == 0:
                rhs = y
            else:
                sync()
                lin(x.view(shape), apply_minv=True, out=jx)
                rhs = y + jx.reshape(-1)
            block_apply_sel(A_col, self.sel[c], rhs, x)
        return x


def block_apply_sel(A_colmajor, sel, x, y):
    """y_b = A_b x_b for the blocks b in sel (dgx_block_apply_sel); the other entries of y are kept."""
    import ctypes
    import torch
    from . import _ffi
    nb, n = A_colmajor.shape[0], A_colmajor.shape[1]
    if not (A_colmajor.is_cuda and A_colmajor.dtype == torch.float32 and A_colmajor.is_contiguous()):
        raise ValueError("blocks must be a contiguous float32 cuda tensor [nb][n][n]")
    for v in (x, y):
        if not (v.is_cuda and v.dtype == torch.float64 and v.is_contiguous() and v.numel() == nb * n):
            raise ValueError("x and y must be contiguous float64 cuda tensors of nb * n values")
    if not (sel.is_cuda and sel.dtype == torch.int32 and sel.is_contiguous()):
        raise ValueError("sel must be a contiguous int32 cuda tensor")
    lib = _ffi.load()
    _ffi.check(lib.dgx_block_apply_sel(sel.numel(), ctypes.c_void_p(sel.data_ptr()), n,
                                       ctypes.c_void_p(A_colmajor.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                       ctypes.c_void_p(y.data_ptr()),
                                       ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)))
    return y


def block_inverse(B, shift, out=None):
    """(shift_b I - B_b)^-1 for every block on the GPU (dgx_block_inverse / dgx_block_inverse_f32: FP32
    Gauss-Jordan with partial pivoting, one workgroup per block, n <= 320).  B: float64 or float32 cuda
    tensor [nb][n][n]; shift: float64 cuda tensor [nb]; out: float32 [nb][n][n] (may be B when B is
    float32).  Returns out; raises if a block is singular."""
    import ctypes
    import torch
    from . import _ffi
    nb, n = B.shape[0], B.shape[1]
    if not (B.is_cuda and B.dtype in (torch.float64, torch.float32) and B.is_contiguous() and B.shape == (nb, n, n)):
        raise ValueError("B must be a contiguous float64/float32 cuda tensor [nb][n][n]")
    if not (shift.is_cuda and shift.dtype == torch.float64 and shift.is_contiguous() and shift.numel() == nb):
        raise ValueError("shift must be a contiguous float64 cuda tensor of nb values")
    if out is None:
        out = torch.empty((nb, n, n), dtype=torch.float32, device=B.device)
    elif not (out.is_cuda and out.dtype == torch.float32 and out.is_contiguous() and out.shape == (nb, n, n)):
        raise ValueError("out must be a contiguous float32 cuda tensor [nb][n][n]")
    sing = torch.empty(nb, dtype=torch.int32, device=B.device)
    lib = _ffi.load()
    fn = lib.dgx_block_inverse if B.dtype == torch.float64 else lib.dgx_block_inverse_f32
    st = fn(nb, n, ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(shift.data_ptr()),
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(sing.data_ptr()),
            ctypes.c_void_p(torch.cuda.current_stream(B.device).cuda_stream))
    _ffi.check(st)
    bad = torch.nonzero(sing)
    if bad.numel():
        b = int(bad[0, 0])
        raise FloatingPointError(f"block {b} is singular (no pivot in column {int(sing[b]) - 1})")
    return out


def block_apply(A_colmajor, x):
    """y_b = A_b x_b on the GPU (dgx_block_apply): A_colmajor float32 cuda [nb][n][n] holding each A_b
    column-major (i.e. A_b^T row-major), x float64 cuda with nb * n values.  Returns float64, flat."""
    import ctypes
    import torch
    from . import _ffi
    nb, n = A_colmajor.shape[0], A_colmajor.shape[1]
    if not (A_colmajor.is_cuda and A_colmajor.dtype == torch.float32 and A_colmajor.is_contiguous()):
        raise ValueError("blocks must be a contiguous float32 cuda tensor [nb][n][n]")
    x = x.reshape(-1)
    if not (x.is_cuda and x.dtype == torch.float64 and x.is_contiguous() and x.numel() == nb * n):
        raise ValueError("x must be a contiguous float64 cuda tensor of nb * n values")
    y = torch.empty_like(x)
    lib = _ffi.load()
    _ffi.check(lib.dgx_block_apply(nb, n, ctypes.c_void_p(A_colmajor.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                   ctypes.c_void_p(y.data_ptr()),
                                   ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)))
    return y


def solve_linear(solver, b, inv_dt=0.0, transpose=False, precond=None, rtol=1e-4, restart=60, max_iter=600):
    """Solve (diag(inv_dt) - M^-1 dR/du) x = b, or its transpose, at the solver's current state by
    flexible right-preconditioned GMRES.  b: torch float64 tensor [E][nv][Ns] on the solver's device; inv_dt: a
    scalar or one value per volume.  Returns (x, relative residual, iterations)."""
    import torch
    shape = solver.shape
    sync = _sync_fn(b.device)
    out = torch.empty(shape, dtype=torch.float64, device=b.device)
    idt = torch.as_tensor(inv_dt, dtype=torch.float64, device=b.device)
    idt = idt.view(-1, 1, 1) if idt.dim() else idt
    op = solver.vjp if transpose else solver.jvp

    def apply_a(w):
        wv = w.view(shape)
        sync()
        op(wv, apply_minv=True, out=out)
        return (wv * idt - out).reshape(-1)

    prec = (lambda y: precond.apply(y, transpose)) if precond is not None else None
    x, rel, its = gmres(apply_a, b.reshape(-1).clone(), rtol=rtol, restart=restart, max_iter=max_iter,
                        precond=prec)
    return x.view(shape), rel, its


def volume_lengths(mesh):
    """Smallest edge length of every volume from its corner geometry nodes (compute_min_length_measure,
    dg/solve_dg.c:320-380: the minimum over the volume's edges for d >= 2)."""
    d, ng = mesh.d, mesh.p_g + 1
    xg = mesh.xyz_g.reshape(mesh.n_elem, d, *([ng] * d))      # [E][d][(t)][s][r]: r runs fastest
    h = np.full(mesh.n_elem, np.inf)
    for c in range(1 << d):
        a = [(c >> m) & 1 for m in range(d)]                 # corner bits per reference direction r, s, t
        for k in range(d):
            if a[k]:
                continue
            b = list(a)
            b[k] = 1
            ia = tuple((ng - 1) * a[d - 1 - m] for m in range(d))
            ib = tuple((ng - 1) * b[d - 1 - m] for m in range(d))
            ea = xg[(slice(None), slice(None)) + ia]
            eb = xg[(slice(None), slice(None)) + ib]
            h = np.minimum(h, np.linalg.norm(ea - eb, axis=1))
    return h


def max_wave_speed(u, sb, d):
    """Per volume, the maximum of |V| + c over the Bezier coefficients of the conserved variables
    (compute_dt_cfl_constrained, dg/solve_dg.c:283-297; compute_max_wavespeed,
    test_case/solution/euler/solution_euler.c:143-170).  u: [E][nv][Ns]; sb: nodal -> Bezier (n1, n1)."""
    import torch
    E, nv, ns = u.shape
    n1 = sb.shape[0]
    x = u.reshape(E, nv, *([n1] * d))
    sbt = sb.t().contiguous()
    for ax in range(d):
        x = torch.movedim(torch.movedim(x, -1 - ax, -1) @ sbt, -1, -1 - ax)
    x = x.reshape(E, nv, ns)
    rho = x[:, 0]
    vel2 = (x[:, 1:d + 1] ** 2).sum(dim=1) / rho ** 2
    pres = (GAMMA - 1.0) * (x[:, d + 1] - 0.5 * rho * vel2)
    return (vel2.sqrt() + (GAMMA * pres / rho).sqrt()).amax(dim=1)


def solve_implicit(solver, mesh, p, cfl_initial=1.0, exit_tol=1e-10, exit_ratio=0.0, max_steps=50,
                   rtol=1e-4, restart=60, max_iter=600, precondition=True, verbose=False, device="cuda",
                   precond=None, sweep="jacobi"):
    """Pseudo-transient Newton-Krylov from the solver's current state, as the reference's implicit loop
    (solve_implicit.c:95-108, implicit_step :232-256, check_exit :258-284): every step computes max_rhs,
    the maximum of |R| over the volumes before M^-1 (compute_rlhs_dg returns compute_max_rhs_dg_like of
    the un-scaled residual, dg/solve_dg.c:138-145), ramps the CFL number with it
    (cfl = cfl_initial * max(1, max_rhs_0 / max_rhs), compute_CFL_ramping :223-257), solves and applies
    the update, and only then stops when max_rhs < exit_tol or max_rhs / max_rhs_0 < exit_ratio (so the
    update of the step that passes the test is applied), or after max_steps steps.

    Defaults follow the reference's GMRES options (restart 60, rtol 1e-4).  Returns (final state
    [E][nv][Ns] as a torch tensor on `device`, history): one entry (max_rhs, cfl, GMRES iterations, GMRES
    relative residual) per step, max_rhs being the one of the state the step started from, and a last
    entry (max_rhs of the final state, 0, 0, 0).  The solver is left at the final state.  `precond`: a
    BlockJacobi (or an object with its update/apply) to use instead of a new one (`sweep`: "jacobi" or
    "gauss-seidel", the latter on the GPU only); precondition=False runs plain GMRES."""
    import torch
    from .problem import ops1d
    dev = torch.device(device)
    sync = _sync_fn(dev)
    shape = solver.shape
    u = solver.get_state(torch.empty(shape, dtype=torch.float64, device=dev))
    h = torch.from_numpy(volume_lengths(mesh)).to(dev)
    sb = torch.from_numpy(ops1d(p)["sb"]).to(dev)
    if not precondition:
        precond = None
    elif precond is None:
        precond = BlockJacobi(solver, mesh, dev, sweep=sweep)
    r = torch.empty_like(u)
    hist = []
    max_rhs0 = None
    for step in range(max_steps):
        sync()
        solver.set_state(u)
        mx = solver.max_rhs(apply_minv=False)
        if not math.isfinite(mx):
            raise FloatingPointError(f"non-finite residual at implicit step {step}")
        if max_rhs0 is None:
            max_rhs0 = mx
        solver.compute_rhs(apply_minv=True, out=r)
        cfl = cfl_initial * max(1.0, max_rhs0 / mx)
        inv_dt = max_wave_speed(u, sb, mesh.d) / (cfl * h)
        bad = torch.nonzero(~torch.isfinite(inv_dt))
        if bad.numel():
            raise FloatingPointError(f"implicit step {step}: no wave speed in volume {int(bad[0, 0])} (a Bezier "
                                     "coefficient with non-positive density or pressure)")
        if precond is not None:
            precond.update(inv_dt)
        du, rel, its = solve_linear(solver, r, inv_dt, precond=precond, rtol=rtol, restart=restart,
                                    max_iter=max_iter)
        hist.append((mx, cfl, its, rel))
        if verbose:
            print(f"implicit step {step}: max_rhs {mx:.3e} cfl {cfl:.3g} gmres {its} ({rel:.1e})", flush=True)
        u = u + du
        if mx < exit_tol or mx < exit_ratio * max_rhs0:
            break
    sync()
    solver.set_state(u)
    hist.append((solver.max_rhs(apply_minv=False), 0.0, 0, 0.0))
    return u, hist


def _sync_fn(device):
    import torch
    return torch.cuda.synchronize if torch.device(device).type == "cuda" else (lambda: None)
