"""Closed-form shallow-water solutions used as accuracy pins.

Carrier & Greenspan (1958), "Water waves of finite amplitude on a sloping
beach", J. Fluid Mech. 4: the periodic standing wave on a plane beach, a
non-breaking solution of the nonlinear equations with a moving shoreline.

Non-dimensional form (g = 1, still-water depth = x, x positive offshore):

    d_t + (u d)_x = 0,   u_t + u u_x + eta_x = 0,   d = x + eta.

The hodograph variables sigma = 4 sqrt(d) and lambda turn the system into
the linear equation (sigma phi_sigma)_sigma = sigma phi_lambda_lambda, whose
separable solution phi = A J0(sigma) cos(lambda) maps back through

    u   = -phi_sigma / sigma      = A J1(sigma) cos(lambda) / sigma
    eta = phi_lambda / 4 - u^2/2  = -(A/4) J0(sigma) sin(lambda) - u^2/2
    x   = sigma^2 / 16 - eta
    t   = lambda / 2 + u

(signs for x offshore-positive; tests check the PDE residual of this map).
The shoreline is sigma = 0.  The map is one-to-one for A < 1 (no breaking).

Dimensional scales for a beach of slope alpha and length scale L0:
x = L0 x*, eta = alpha L0 eta*, u = sqrt(g alpha L0) u*,
t = sqrt(L0 / (g alpha)) t*.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import j0, j1, jv

G = 9.81


def _u_eta(sig, lam, A):
    r = np.where(sig > 1e-12, j1(sig) / np.maximum(sig, 1e-300), 0.5)
    u = A * r * np.cos(lam)
    eta = -0.25 * A * j0(sig) * np.sin(lam) - 0.5 * u * u
    return u, eta


def _lambda_at(sig, t, A):
    """Solve t = lambda/2 + u(sigma, lambda) for lambda (vectorised bisection).
    |du/dlambda| <= A/2 < 1/2, so the right side is increasing in lambda and
    |u| <= A/2 brackets the root in [2t - A, 2t + A]."""
    sig = np.asarray(sig, dtype=np.float64)
    lo = np.full(sig.shape, 2 * t - A - 1e-9)
    hi = np.full(sig.shape, 2 * t + A + 1e-9)
    for _ in range(56):
        mid = 0.5 * (lo + hi)
        u, _ = _u_eta(sig, mid, A)
        f = 0.5 * mid + u - t
        lo = np.where(f < 0, mid, lo)
        hi = np.where(f < 0, hi, mid)
    return 0.5 * (lo + hi)


def cg_point(sig, lam, A):
    """(x, t, u, eta) at hodograph point(s) (sigma, lambda)."""
    u, eta = _u_eta(sig, lam, A)
    return sig * sig / 16 - eta, 0.5 * lam + u, u, eta


def cg_shoreline(t, A):
    """Non-dimensional shoreline (x_s, u_s, eta_s) at time t (sigma = 0)."""
    lam = float(_lambda_at(np.zeros(1), t, A)[0])
    u, eta = _u_eta(np.zeros(1), np.array([lam]), A)
    return -float(eta[0]), float(u[0]), float(eta[0])


def cg_runup(A):
    """Maximum shoreline elevation over a period (non-dimensional)."""
    lam = np.linspace(0, 2 * math.pi, 200001)
    _, eta = _u_eta(np.zeros_like(lam), lam, A)
    return float(eta.max())


def cg_field(x, t, A, sig_max=None):
    """(eta, u, wet) at non-dimensional positions x and time t.  Dry points
    (landward of the shoreline) get eta = -x (zero depth) and u = 0."""
    x = np.asarray(x, dtype=np.float64)
    if sig_max is None:
        sig_max = 4 * math.sqrt(max(x.max(), 0.0) + 1.0) + 1.0
    sig = np.linspace(0.0, sig_max, 4001)
    lam = _lambda_at(sig, t, A)
    xs, _, us, es = cg_point(sig, lam, A)
    if np.any(np.diff(xs) <= 0):
        raise ValueError("Carrier-Greenspan map not monotone (breaking)")
    wet = x >= xs[0]
    eta = np.where(wet, np.interp(x, xs, es), -x)
    u = np.where(wet, np.interp(x, xs, us), 0.0)
    return eta, u, wet


def cg_at(x, t, A, iters=30):
    """(eta, u) at one wet position x > 0 for an array of times t, by Newton
    iteration on (sigma, lambda) with the analytic Jacobian
    (d/dsigma of J1(s)/s is -J2(s)/s)."""
    t = np.asarray(t, dtype=np.float64)
    sig = np.full(t.shape, 4 * math.sqrt(x))
    lam = 2 * t.copy()
    for _ in range(iters):
        c, s_ = np.cos(lam), np.sin(lam)
        r1 = j1(sig) / sig
        u = A * r1 * c
        eta = -0.25 * A * j0(sig) * s_ - 0.5 * u * u
        u_s = -A * c * jv(2, sig) / sig
        u_l = -A * r1 * s_
        e_s = 0.25 * A * j1(sig) * s_ - u * u_s
        e_l = -0.25 * A * j0(sig) * c - u * u_l
        f1 = sig * sig / 16 - eta - x
        f2 = 0.5 * lam + u - t
        a, b, cc, d = sig / 8 - e_s, -e_l, u_s, 0.5 + u_l
        det = a * d - b * cc
        sig = sig - (d * f1 - b * f2) / det
        lam = lam - (a * f2 - cc * f1) / det
    u, eta = _u_eta(sig, lam, A)
    return eta, u


class Beach:
    """Dimensional scales of a plane beach of slope alpha, length scale L0."""

    def __init__(self, alpha: float, L0: float):
        self.alpha, self.L0 = alpha, L0
        self.T0 = math.sqrt(L0 / (G * alpha))
        self.U0 = math.sqrt(G * alpha * L0)
        self.Z0 = alpha * L0

    def field(self, x_m, t_s, A):
        eta, u, wet = cg_field(np.asarray(x_m) / self.L0, t_s / self.T0, A)
        return eta * self.Z0, u * self.U0, wet

    def at(self, x_m, t_s, A):
        eta, u = cg_at(x_m / self.L0, np.asarray(t_s) / self.T0, A)
        return eta * self.Z0, u * self.U0

    def depth_below_datum(self, x_m):
        """Still-water depth H(x) = alpha x (negative on land)."""
        return self.alpha * np.asarray(x_m, dtype=np.float64)


def carrier_greenspan_scenario(dx: float, A: float = 0.5, alpha: float = 0.05, L0: float = 1000.0,
                               xl: float = -0.3, xr: float = 1.3, t_series: float = 400.0):
    """Plane beach of slope alpha over x in [xl L0, xr L0] (land on the left,
    a wall there), one row of cells, frictionless plain shallow water, the
    right side driven by the exact solution through the CC inflow series
    (ghost state at x = xr L0 + dx/2, sampled every 0.5 s).  Starts from the
    exact state at t = 0.  Returns (scenario, beach, cell centres)."""
    from tsuner_amd import scenarios as S
    B = Beach(alpha, L0)
    n = int(round((xr - xl) * L0 / dx))
    x = xl * L0 + (np.arange(n) + 0.5) * dx
    H = B.depth_below_datum(x)
    eta0, u0, wet = B.field(x, 0.0, A)
    h = np.where(wet, np.maximum(eta0 + H, 0.0), 0.0)
    p = S.base_params(n, 1, dx, dx, friction=0.0)
    s = S.default_scheme(bedload=0, suspension=0, bc=[S.BC_WALL, S.BC_INFLOW, S.BC_WALL, S.BC_WALL],
                         extrema=1)
    tCC = np.arange(0.0, t_series, 0.5)
    xg = xr * L0 + 0.5 * dx
    CC = np.zeros((len(tCC), 2 * n + 2, 4))
    e, u = B.at(xg, tCC, A)
    CC[:, n, 0], CC[:, n, 1], CC[:, n, 3] = e, u, B.depth_below_datum(xg)
    V = S.pack_state(h[None, :], (h * u0)[None, :], np.zeros((1, n)), H[None, :])
    sc = S.Scenario("carrier_greenspan_dx%g" % dx, p, s, V, np.ones((1, n), np.int32), CC, tCC, steps=0)
    return sc, B, x
