"""Host logic of the matrix-free implicit solver (dpgsolver_amd/implicit.py), on the CPU.

- GMRES against numpy's dense solve on known-answer systems (restarted and full), and the finite
  termination of full GMRES (n iterations for an n x n system);
- the volume colouring is a proper distance-1 colouring (periodic odd counts need 3 colours, a volume
  that is its own neighbour is refused);
- the length measure equals the grid spacing on structured meshes (the reference's minimum edge length);
- the wave speed is taken over the Bezier coefficients (a linear density reaches its end values there,
  which the nodal values never do);
- the probed volume-diagonal blocks equal the diagonal blocks of the oracle's dense complex-step
  Jacobian, and the pseudo-transient loop drives a perturbed uniform flow to a steady state, with the
  oracle's residual and complex-step JVP standing in for the GPU entry points.
"""
import math
import os

import numpy as np
import pytest
import torch

import dpgsolver_amd as dg
from dpgsolver_amd import implicit
from oracle.oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _system(n, seed, shift):
    rng = np.random.default_rng(seed)
    A = shift * np.eye(n) + rng.standard_normal((n, n)) / math.sqrt(n)
    x = rng.standard_normal(n)
    return A, x, A @ x


@pytest.mark.parametrize("restart", [10, 200])
def test_gmres_known_answer(restart):
    A, x, b = _system(120, 1, 2.5)
    At = torch.from_numpy(A)
    sol, rel, its = implicit.gmres(lambda v: At @ v, torch.from_numpy(b), rtol=1e-12, restart=restart,
                                   max_iter=2000)
    assert rel <= 1e-12
    assert np.linalg.norm(b - A @ sol.numpy()) <= 1.01e-12 * np.linalg.norm(b)
    assert np.allclose(sol.numpy(), np.linalg.solve(A, b), rtol=0, atol=1e-9)


def test_gmres_full_terminates_in_n_iterations():
    """Full GMRES minimises over the whole Krylov space after n steps: exact up to round-off."""
    A, x, b = _system(40, 2, 0.3)          # poorly conditioned, indefinite spectrum
    At = torch.from_numpy(A)
    sol, rel, its = implicit.gmres(lambda v: At @ v, torch.from_numpy(b), rtol=1e-14, restart=40, max_iter=40)
    assert its <= 40
    assert np.linalg.norm(sol.numpy() - x) <= 1e-8 * np.linalg.norm(x)


def test_gmres_zero_rhs():
    sol, rel, its = implicit.gmres(lambda v: v, torch.zeros(7, dtype=torch.float64))
    assert its == 0 and rel == 0.0 and not sol.any()


def _check_coloring(mesh):
    col, nc = implicit.element_colors(mesh)
    fe = np.asarray(mesh.face_elem).reshape(-1, 2)
    assert np.all(col[fe[:, 0]] != col[fe[:, 1]])
    assert set(col.tolist()) == set(range(nc))
    return nc


def test_element_colors():
    assert _check_coloring(dg.structured_mesh(2, (4, 6), (0, 0), (1, 1))) == 2
    assert 3 <= _check_coloring(dg.structured_mesh(2, (3, 4), (0, 0), (1, 1))) <= 4   # odd cycles need 3
    assert _check_coloring(dg.structured_mesh(3, (3, 3, 3), (0, 0, 0), (1, 1, 1))) <= 4
    mesh, _ = dg.read_gmsh(os.path.join(HERE, "golden", "meshes", "straight_2d_quad_periodic.msh"), 2)
    _check_coloring(mesh)
    with pytest.raises(ValueError):
        implicit.element_colors(dg.structured_mesh(2, (1, 2), (0, 0), (1, 1)))


def test_volume_lengths_structured():
    h = implicit.volume_lengths(dg.structured_mesh(2, (4, 5), (0, 0), (2, 1)))
    assert np.allclose(h, 0.2, rtol=1e-14)
    h = implicit.volume_lengths(dg.structured_mesh(3, (2, 3, 4), (0, 0, 0), (1, 0.9, 2)))
    assert np.allclose(h, 0.3, rtol=1e-14)
    h = implicit.volume_lengths(dg.structured_mesh(2, (4, 4), (0, 0), (1, 1), p_g=2, amp=0.05, seed=1))
    assert np.all(h > 0.15) and np.all(h < 0.35)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_max_wave_speed_over_bezier_coefficients(p):
    """rho = 1 + x/2 along r, V = 0, pressure p0: the Bezier coefficients of a linear function are its
    values at equally spaced points, so the smallest is rho(-1) = 1/2 and the speed is sqrt(gamma p0/0.5);
    the GL nodal values stay inside (-1, 1) and would give a smaller one."""
    g, p0 = 1.4, 0.8
    x = dg.problem.ops1d(p)["x"]
    n1 = p + 1
    rho = np.array([1.0 + 0.5 * x[n % n1] for n in range(n1 * n1)])
    u = np.zeros((1, 4, n1 * n1))
    u[0, 0] = rho
    u[0, 3] = p0 / (g - 1.0)
    sb = torch.from_numpy(dg.problem.ops1d(p)["sb"])
    s = float(implicit.max_wave_speed(torch.from_numpy(u), sb, 2)[0])
    assert abs(s - math.sqrt(g * p0 / 0.5)) <= 1e-13
    assert s > math.sqrt(g * p0 / rho.min()) + 1e-3
    uni = np.zeros((2, 5, n1 ** 3))
    uni[:, 0], uni[:, 1], uni[:, 2], uni[:, 3] = 1.2, 0.6, -0.3, 0.24
    uni[:, 4] = 1.0 / (g - 1.0) + 0.5 * (0.36 + 0.09 + 0.0576) / 1.2
    s = implicit.max_wave_speed(torch.from_numpy(uni), sb, 3).numpy()
    vel = math.sqrt(0.36 + 0.09 + 0.0576) / 1.2
    assert np.allclose(s, vel + math.sqrt(g * 1.0 / 1.2), rtol=1e-13)


class _OracleStandIn:
    """The Solver methods solve_implicit calls, on the oracle (residual and complex-step JVP)."""

    def __init__(self, mesh, geo, p, flux):
        self.o = Oracle(mesh, p, num_flux=flux, positivity=False, geometry=geo)
        self.shape = (mesh.n_elem, mesh.d + 2, (p + 1) ** mesh.d)

    def set_state(self, u):
        self.u = u.numpy().copy()

    def get_state(self, out):
        out.copy_(torch.from_numpy(self.u))
        return out

    def compute_rhs(self, apply_minv=True, out=None):
        r, mx = self.o.residual(self.u, apply_minv)
        out.copy_(torch.from_numpy(r))
        return out, mx

    def max_rhs(self, apply_minv=True):
        return self.o.residual(self.u, apply_minv)[1]

    def jvp(self, w, apply_minv=False, out=None):
        out.copy_(torch.from_numpy(self.o.jvp_cs(self.u, w.numpy(), apply_minv=apply_minv)))
        return out


def _uniform_case(n, p, d=2, amp=1e-2):
    L = 2 * math.pi
    mesh = dg.structured_mesh(d, (n,) * d, (0,) * d, (L,) * d, p_g=1)
    geo = dg.geometry(mesh, p)
    return mesh, geo, dg.ic_uniform_flow(mesh, p, amp=amp)


def test_block_diagonal_matches_dense_jacobian():
    mesh, geo, u = _uniform_case(3, 1)
    s = _OracleStandIn(mesh, geo, 1, 1)
    s.set_state(torch.from_numpy(u))
    col, nc = implicit.element_colors(mesh)
    B = implicit.block_diagonal(s, col, nc, "cpu").numpy()
    N, m = u.size, u[0].size
    J = np.empty((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        J[:, j] = s.o.jvp_cs(u, e.reshape(u.shape), apply_minv=True).ravel()
    for e in range(mesh.n_elem):
        blk = J[e * m:(e + 1) * m, e * m:(e + 1) * m]
        assert np.abs(B[e] - blk).max() <= 1e-13 * np.abs(blk).max()


def test_pseudo_transient_loop_converges_on_standin():
    mesh, geo, u = _uniform_case(4, 1)
    s = _OracleStandIn(mesh, geo, 1, 1)
    s.set_state(torch.from_numpy(u))
    uf, hist = implicit.solve_implicit(s, mesh, 1, cfl_initial=1.0, exit_tol=1e-10, max_steps=25, device="cpu")
    mx = [h[0] for h in hist]
    assert mx[-1] < 1e-10, mx
    assert hist[0][1] == 1.0 and hist[1][1] > 1.0          # the CFL ramps with max_rhs_0 / max_rhs
    r, mxo = s.o.residual(uf.numpy(), apply_minv=False)
    assert mxo < 1e-10 and mxo == mx[-1]
    # the steady state of a periodic domain near a uniform flow is that uniform flow
    dev = np.abs(uf.numpy() - uf.numpy().mean(axis=(0, 2), keepdims=True)).max()
    assert dev < 1e-6, dev


def test_implicit_loop_order_and_max_rhs_as_reference():
    """The loop of solve_implicit.c:95-108: max_rhs is the un-scaled residual's (before M^-1, the value
    compute_rlhs_dg returns), the CFL ramps with max_rhs_0 / max_rhs of it, and the step whose max_rhs
    passes the exit test still solves and applies its update before the loop stops."""
    mesh, geo, u = _uniform_case(4, 1)
    s = _OracleStandIn(mesh, geo, 1, 1)
    s.set_state(torch.from_numpy(u))
    _, h3 = implicit.solve_implicit(s, mesh, 1, cfl_initial=1.0, exit_tol=0.0, max_steps=3, device="cpu")
    assert len(h3) == 4 and all(h[2] > 0 for h in h3[:3])
    assert h3[0][0] == s.o.residual(u, apply_minv=False)[1]          # not the M^-1 residual's
    assert h3[0][0] != s.o.residual(u, apply_minv=True)[1]
    for k in range(3):
        assert h3[k][1] == max(1.0, h3[0][0] / h3[k][0])
    # an exit tolerance between the first and second step's max_rhs: the second step passes, and its
    # update is applied (two updates), so the final max_rhs is the third value of the 3-step run
    tol = 0.5 * (h3[0][0] + h3[1][0])
    assert h3[1][0] < tol < h3[0][0]
    s.set_state(torch.from_numpy(u))
    uf, h2 = implicit.solve_implicit(s, mesh, 1, cfl_initial=1.0, exit_tol=tol, max_steps=10, device="cpu")
    assert len(h2) == 3 and h2[0][:2] == h3[0][:2] and h2[1][:2] == h3[1][:2]
    assert abs(h2[-1][0] - h3[2][0]) <= 1e-12 * h3[2][0]
