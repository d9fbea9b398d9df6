"""Pins of the oracle's numerical fluxes by routes independent of the reference's algebra (CPU).

The reference's Roe-Pike flux (numerical_flux_euler_T.c:470-603) is written in the "Pike" form
(lc1, lc2, disInter1/2) with the entropy fix of Qu et al. 2015 eq. (35).  These tests compare it with
the textbook matrix form (Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, 3rd ed.,
section 11.2):

    F* = 1/2 n.(F_L + F_R) - 1/2 R |Lambda| R^-1 (W_R - W_L)

where A = R Lambda R^-1 is the Roe matrix (R |Lambda| R^-1 evaluated as sum_k |lambda_k| P_k with the spectral projectors
P_k of its three distinct eigenvalues).  A is NOT built from the reference's formulas here: it is the
central-difference Jacobian of the (closed-form pinned) physical flux n.F(W) at the Roe-averaged
state, checked by the Roe property A (W_R - W_L) = n.(F_R - F_L), and its eigenvalues come from a
numerical eigen-decomposition.  A plausible mistake in the Pike form (lc1 and lc2 exchanged, a 1/c dropped,
VnL and VnR exchanged in the entropy fix) changes the result at O(1) and fails these checks; the
consistency, contact and rotation tests of test_oracle_golden.py make every dissipation term vanish
or are blind to scalar mistakes, so they cannot.

The Lax-Friedrichs wave speed is pinned the same way: its dissipation coefficient must be the largest
spectral radius of the directional flux Jacobian over all directions (|V| + c), independent of n.
"""
import numpy as np
import pytest

from oracle import oracle as O

GAMMA = 1.4
GM1 = GAMMA - 1.0


def _state(d, rho, vel, p):
    u = np.empty(d + 2)
    u[0] = rho
    u[1:1 + d] = rho * np.asarray(vel[:d])
    u[d + 1] = p / GM1 + 0.5 * rho * float(np.dot(vel[:d], vel[:d]))
    return u


def _prim(d, w):
    rho = w[0]
    vel = w[1:1 + d] / rho
    p = GM1 * (w[d + 1] - 0.5 * rho * vel @ vel)
    return rho, vel, p


def _nflux(d, w, n):
    """n . F(W) from the oracle's pointwise flux (pinned by closed form and rotation covariance)."""
    return n @ O.flux_point(d, w)


def _jacobian(d, w, n, h=1e-6):
    """Central-difference Jacobian of n.F(W) at W (relative step; error ~1e-10 relative)."""
    nv = d + 2
    A = np.empty((nv, nv))
    for j in range(nv):
        e = np.zeros(nv)
        e[j] = h * max(1.0, abs(w[j]))
        A[:, j] = (_nflux(d, w + e, n) - _nflux(d, w - e, n)) / (2 * e[j])
    return A


def _roe_state(d, wl, wr):
    """Conservative state at the Roe average (sqrt(rho)-weighted velocity and total enthalpy)."""
    rl, vl, pl = _prim(d, wl)
    rr, vr, pr = _prim(d, wr)
    hl = (wl[d + 1] + pl) / rl
    hr = (wr[d + 1] + pr) / rr
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    rho = sl * sr
    vel = (sl * vl + sr * vr) / (sl + sr)
    H = (sl * hl + sr * hr) / (sl + sr)
    p = rho * (H - 0.5 * vel @ vel) * GM1 / GAMMA
    return _state(d, rho, np.pad(vel, (0, 3 - d)), p), vel, H


class _Roe:
    """Textbook Roe decomposition of one face point."""

    def __init__(self, d, wl, wr, n):
        self.d, self.wl, self.wr, self.n = d, wl, wr, n
        self.w_hat, self.vel, self.H = _roe_state(d, wl, wr)
        self.A = _jacobian(d, self.w_hat, n)
        self.dW = wr - wl
        self.dF = _nflux(d, wr, n) - _nflux(d, wl, n)
        # A is diagonalizable with three distinct eigenvalues (Vn - c, Vn with multiplicity d, Vn + c);
        # its spectral projectors are polynomials in A, so no (ill-conditioned) eigenvectors are needed
        ev = np.sort(np.linalg.eigvals(self.A).real)
        self.lam = np.array([ev[0], ev[1:-1].mean(), ev[-1]])
        nv = d + 2
        I = np.eye(nv)
        self.P = []
        for k in range(3):
            P = I.copy()
            for j in range(3):
                if j != k:
                    P = P @ (self.A - self.lam[j] * I) / (self.lam[k] - self.lam[j])
            self.P.append(P)
        self.c = np.sqrt(GM1 * (self.H - 0.5 * self.vel @ self.vel))
        self.Vn = self.vel @ n
        self.VnL = _prim(d, wl)[1] @ n
        self.VnR = _prim(d, wr)[1] @ n

    def roe_property_error(self):
        return np.linalg.norm(self.A @ self.dW - self.dF) / max(1.0, np.linalg.norm(self.dF))

    def wave(self, target):
        """alpha_k r_k (the projection of dW on the eigenspace) of the eigenvalue closest to target."""
        k = int(np.argmin(np.abs(self.lam - target)))
        return self.P[k] @ self.dW

    def textbook(self, lam_abs=None):
        lam_abs = np.abs(self.lam) if lam_abs is None else lam_abs
        D = sum(lam_abs[k] * (self.P[k] @ self.dW) for k in range(3))
        return 0.5 * (_nflux(self.d, self.wl, self.n) + _nflux(self.d, self.wr, self.n)) - 0.5 * D


def _random_pair(d, rng, speed=1.0):
    n = rng.normal(size=d)
    n /= np.linalg.norm(n)
    wl = _state(d, rng.uniform(0.5, 2.0), speed * rng.normal(size=3), rng.uniform(0.4, 3.0))
    wr = _state(d, rng.uniform(0.5, 2.0), speed * rng.normal(size=3), rng.uniform(0.4, 3.0))
    return wl, wr, n


def _fix_inactive(r):
    return (abs(r.Vn - r.c) <= abs(r.VnL - r.c)) and (abs(r.Vn + r.c) >= abs(r.VnR + r.c))


@pytest.mark.parametrize("d", [2, 3])
def test_roe_matrix_has_roe_property(d):
    """The finite-difference Roe matrix at the Roe average satisfies A dW = dF (the defining
    property), and its eigenvalues are Vn - c, Vn (d times), Vn + c: the checker itself is sound."""
    rng = np.random.default_rng(100 + d)
    for _ in range(20):
        wl, wr, n = _random_pair(d, rng)
        r = _Roe(d, wl, wr, n)
        assert r.roe_property_error() < 1e-8
        ev = np.sort(np.linalg.eigvals(r.A).real)
        assert np.allclose(ev, np.sort([r.Vn - r.c] + [r.Vn] * d + [r.Vn + r.c]), atol=1e-6)
        # the projectors resolve the identity and reproduce A
        assert np.allclose(sum(r.P), np.eye(d + 2), atol=1e-7)
        assert np.allclose(sum(l * P for l, P in zip(r.lam, r.P)), r.A, atol=1e-7 * np.abs(r.A).max())


@pytest.mark.parametrize("d", [2, 3])
def test_roe_pike_equals_textbook_matrix_form(d):
    """Where the entropy fix is inactive, the Roe-Pike flux equals the textbook form
    1/2 n.(F_L+F_R) - 1/2 R|Lambda|R^-1 dW (Toro 11.2) with the numerically decomposed Roe matrix."""
    rng = np.random.default_rng(200 + d)
    checked = 0
    while checked < 40:
        wl, wr, n = _random_pair(d, rng)
        r = _Roe(d, wl, wr, n)
        if not _fix_inactive(r):
            continue
        nf = O.numflux_point(d, 1, wl, wr, n)
        ref = r.textbook()
        assert np.linalg.norm(nf - ref) <= 1e-8 * max(1.0, np.linalg.norm(ref)), (nf, ref)
        checked += 1


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("direction", [1, -1])
def test_roe_pike_supersonic_is_upwind(d, direction):
    """Supersonic flow through the face with the entropy fix inactive (all eigenvalues of one sign,
    VnL >= Vn >= VnR for flow along n): every |lambda| = +-lambda, so the dissipation is +-A dW = +-dF
    and the flux is exactly the upwind side's n.F (left for flow along n, right against it)."""
    rng = np.random.default_rng(300 + d + (direction > 0))
    checked = 0
    while checked < 20:
        n = rng.normal(size=d)
        n /= np.linalg.norm(n)
        pl, pr = rng.uniform(0.5, 1.5, size=2)
        rl, rr = rng.uniform(0.8, 1.2, size=2)
        # normal speeds well above the sound speed (c ~ 1.2), tangential parts random
        vnl, vnr = direction * rng.uniform(2.5, 4.0, size=2)
        tl, tr = rng.normal(size=3) * 0.3, rng.normal(size=3) * 0.3
        n3 = np.pad(n, (0, 3 - d))
        tl -= (tl @ n3) * n3
        tr -= (tr @ n3) * n3
        wl = _state(d, rl, vnl * n3 + tl, pl)
        wr = _state(d, rr, vnr * n3 + tr, pr)
        r = _Roe(d, wl, wr, n)
        lo, hi = sorted([r.VnL, r.VnR])
        if not (lo <= r.Vn <= hi) or abs(r.Vn) < r.c + 0.5:
            continue
        if direction > 0 and not (r.VnL >= r.Vn >= r.VnR):
            continue
        if direction < 0 and not (r.VnL <= r.Vn <= r.VnR):
            continue
        assert _fix_inactive(r)
        nf = O.numflux_point(d, 1, wl, wr, n)
        up = _nflux(d, wl if direction > 0 else wr, n)
        assert np.allclose(nf, up, rtol=1e-12, atol=1e-12 * np.abs(up).max())
        checked += 1


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("wave", [1, 5])
def test_roe_pike_entropy_fix_changes_one_wave(d, wave):
    """With the entropy fix active on one acoustic wave only, F - F_textbook is a multiple of that
    wave's term alone: -1/2 (l_fix - |lambda|) alpha_k r_k, with l_fix = |VnL - c| for wave 1 (left
    state) and |VnR + c| for wave 5 (right state).  This pins the side of the fix: using the other
    side's normal velocity moves the flux off this line."""
    rng = np.random.default_rng(400 + 10 * d + wave)
    checked = 0
    for _ in range(20000):
        if checked >= 15:
            break
        wl, wr, n = _random_pair(d, rng, speed=1.3)
        r = _Roe(d, wl, wr, n)
        fix1 = abs(r.VnL - r.c) < abs(r.Vn - r.c)
        fix5 = abs(r.VnR + r.c) > abs(r.Vn + r.c)
        if wave == 1 and not (fix1 and not fix5):
            continue
        if wave == 5 and not (fix5 and not fix1):
            continue
        # a fix that moves the eigenvalue by a visible amount
        if wave == 1:
            l_fix, lam = abs(r.VnL - r.c), r.Vn - r.c
        else:
            l_fix, lam = abs(r.VnR + r.c), r.Vn + r.c
        if abs(l_fix - abs(lam)) < 0.05:
            continue
        nf = O.numflux_point(d, 1, wl, wr, n)
        pred = r.textbook() - 0.5 * (l_fix - abs(lam)) * r.wave(lam)
        scale = max(1.0, np.linalg.norm(pred))
        assert np.linalg.norm(nf - pred) <= 1e-8 * scale, (wave, nf, pred)
        # and the fix is visible: the textbook flux alone does not match
        assert np.linalg.norm(nf - r.textbook()) > 1e-4 * scale
        checked += 1
    assert checked >= 15


def _spectral_radius(d, w, n):
    return np.abs(np.linalg.eigvals(_jacobian(d, w, n))).max()


@pytest.mark.parametrize("d", [2, 3])
def test_lax_friedrichs_wave_speed_is_direction_independent(d):
    """LF dissipation D = F* - 1/2 n.(F_L+F_R) equals 1/2 lam (W_L - W_R) with one scalar lam for
    every normal, and lam is the largest spectral radius of the directional flux Jacobian over all
    directions and both sides (|V| + c, reached along V), not the face-normal one (|Vn| + c)."""
    rng = np.random.default_rng(500 + d)
    for _ in range(10):
        wl = _state(d, rng.uniform(0.5, 2.0), 1.5 * rng.normal(size=3), rng.uniform(0.4, 3.0))
        wr = _state(d, rng.uniform(0.5, 2.0), 1.5 * rng.normal(size=3), rng.uniform(0.4, 3.0))
        dw = 0.5 * (wl - wr)
        lams = []
        for _k in range(6):
            n = rng.normal(size=d)
            n /= np.linalg.norm(n)
            D = O.numflux_point(d, 0, wl, wr, n) - 0.5 * (_nflux(d, wl, n) + _nflux(d, wr, n))
            lam = (D @ dw) / (dw @ dw)
            assert np.allclose(D, lam * dw, rtol=1e-12, atol=1e-12 * np.abs(D).max())
            lams.append(lam)
        assert np.ptp(lams) <= 1e-12 * max(lams)
        # largest directional spectral radius, attained along each side's velocity
        rad = []
        for w in (wl, wr):
            v = _prim(d, w)[1]
            rad.append(_spectral_radius(d, w, v / np.linalg.norm(v)))
            for _k in range(4):          # no other direction exceeds it
                n = rng.normal(size=d)
                n /= np.linalg.norm(n)
                assert _spectral_radius(d, w, n) <= rad[-1] * (1 + 1e-8)
        assert abs(lams[0] - max(rad)) <= 1e-7 * max(rad)
