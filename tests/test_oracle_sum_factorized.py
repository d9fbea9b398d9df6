"""The oracle's sum-factorized mode (the stronger CPU baseline) applies the same operators as its dense
per-element path: residuals with and without M^-1, the maximum |rhs|, SSP-RK3 stages with the positivity
limiter and the Navier-Stokes (BR2) residual agree to round-off (the summation order differs), on curved 2-D
and 3-D meshes, LF and Roe-Pike.  A transposed 1-D factor or a factor applied in the wrong direction fails
them."""
import math

import numpy as np
import pytest

import dpgsolver_amd as dg
from oracle.oracle import Oracle


def _case(d, p, n):
    if d == 3:
        mesh = dg.structured_mesh(3, (n, n, n), (0, 0, 0), (2 * math.pi,) * 3, p_g=p, amp=0.05, seed=7)
        geo = dg.geometry(mesh, p)
        u = dg.ic_taylor_green(mesh, p, 0.3)
        return mesh, geo, u + 1e-2 * dg.randn(11, u.shape)
    mesh = dg.structured_mesh(2, (n, n + 1), (-5, -5), (5, 5), p_g=p, amp=0.04, seed=3)
    geo = dg.geometry(mesh, p)
    u = dg.ic_isentropic_vortex(mesh, p)
    return mesh, geo, u + 1e-2 * dg.randn(5, u.shape)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("d,p,n,num_flux", [(3, 3, 3, 1), (3, 2, 3, 0), (2, 3, 4, 1), (3, 4, 2, 1)])
def test_sum_factorized_residual_equals_dense(d, p, n, num_flux):
    mesh, geo, u = _case(d, p, n)
    o = Oracle(mesh, p, num_flux=num_flux, geometry=geo)
    for minv in (True, False):
        o.set_sum_factorized(False)
        r0, m0 = o.residual(u, apply_minv=minv)
        o.set_sum_factorized(True)
        r1, m1 = o.residual(u, apply_minv=minv)
        assert _rel(r1, r0) < 1e-12, (minv, _rel(r1, r0))
        assert abs(m1 - m0) <= 1e-12 * m0


def test_sum_factorized_stages_equal_dense():
    mesh, geo, u0 = _case(3, 3, 3)
    dt = dg.dt_cfl(mesh, 3, u0)
    o = Oracle(mesh, 3, num_flux=1, geometry=geo)
    out = []
    for sf in (False, True):
        o.set_sum_factorized(sf)
        u, up = u0.copy(), np.zeros_like(u0)
        for s in range(3):
            o.rk_stage(s, dt, u, up)
        out.append(u)
    assert _rel(out[1], out[0]) < 1e-12


@pytest.mark.parametrize("d,p,n", [(3, 3, 2), (2, 4, 3)])
def test_sum_factorized_navier_stokes_equals_dense(d, p, n):
    """The BR2 viscous path: volume and face gradients (the derivative factors), their face liftings
    and the viscous fluxes go through the same operator form."""
    mesh, geo, u = _case(d, p, n)
    o = Oracle(mesh, p, num_flux=1, geometry=geo, viscous=(0.05, 0.71))
    o.set_sum_factorized(False)
    r0, _ = o.residual(u)
    o.set_sum_factorized(True)
    r1, _ = o.residual(u)
    assert _rel(r1, r0) < 1e-12, _rel(r1, r0)
