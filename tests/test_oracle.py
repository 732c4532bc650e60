"""Pins of the CPU oracle (oracle/tsu_oracle.c) independent of its own code.

The reference ships no tests, fixtures or golden vectors, and cannot be built
here (it needs nvcc, CUDA headers and libnetcdf), so each oracle function is
pinned by what the mathematics fixes: textbook closed forms evaluated in a
different algebraic form, exact symmetries of the scheme, conservation
identities, special cases with known answers (consistency, lake at rest,
uniform states) and convergence to the Ritter dam-break solution.
"""
import math

import numpy as np
import pytest

from tsuner_amd import scenarios as S

G = 9.81


@pytest.fixture(scope="module")
def sed():
    sc = S.sediment_shelf(n=8, dx=10.0, inflow=False)
    return sc.params, sc.scheme


# ---------------------------------------------------------------- limiters --

def test_vanleer_is_the_harmonic_mean(oracle):
    rng = np.random.default_rng(0)
    for _ in range(2000):
        a, b = rng.uniform(1e-3, 10, 2) * rng.choice([-1, 1])
        got = oracle.limiter(S.LIM_VANLEER, a, b)
        assert got == pytest.approx(2 * a * b / (a + b), rel=1e-14)
        assert abs(got) <= 2 * min(abs(a), abs(b)) * (1 + 1e-15)   # TVD bound
        assert oracle.limiter(S.LIM_VANLEER, b, a) == got
        assert oracle.limiter(S.LIM_VANLEER, -a, -b) == -got
        assert oracle.limiter(S.LIM_VANLEER, a, -b) == 0.0
    assert oracle.limiter(S.LIM_VANLEER, 1e-17, 1e-17) == 0.0   # |a|+|b| <= DBL_EPSILON
    assert oracle.limiter(S.LIM_VANLEER, 0.0, 3.0) == 0.0


def test_minmod(oracle):
    assert oracle.limiter(S.LIM_MINMOD, 1.0, 3.0) == 1.0
    assert oracle.limiter(S.LIM_MINMOD, -4.0, -2.0) == -2.0
    assert oracle.limiter(S.LIM_MINMOD, -1.0, 2.0) == 0.0
    assert oracle.limiter(S.LIM_MINMOD, 0.0, 2.0) == 0.0


# ------------------------------------------------------- velocity / filter --

def test_velocity_desingularisation_and_cap(oracle):
    ux, uy = oracle.velocity(2.0, 3.0, -1.0)
    assert ux == pytest.approx(1.5, rel=1e-15) and uy == pytest.approx(-0.5, rel=1e-15)
    assert oracle.velocity(9.99e-7, 5.0, 5.0) == (0.0, 0.0)      # h < 1e-6: dry
    ux, uy = oracle.velocity(1e-6, 1e-6, 0.0)                     # h == 1e-6 counts as wet
    assert ux == pytest.approx(1.0, rel=1e-9)
    ux, uy = oracle.velocity(1.0, 30.0, 40.0)                     # |u| = 50 -> capped at 10
    assert math.hypot(ux, uy) == pytest.approx(10.0, rel=1e-15)
    assert ux / uy == pytest.approx(0.75, rel=1e-15)


def test_state_filter(oracle):
    h, qx, qy, hc = oracle.state_filter(-0.3, 1.0, 1.0, 0.2)
    assert (h, qx, qy, hc) == (0.0, 0.0, 0.0, 0.0)
    h, qx, qy, hc = oracle.state_filter(2.0, 4.0, -2.0, -0.1)     # negative concentration clipped
    assert h == 2.0 and hc == 0.0
    assert qx == pytest.approx(4.0, rel=1e-15) and qy == pytest.approx(-2.0, rel=1e-15)
    h, qx, qy, hc = oracle.state_filter(1.0, 50.0, 0.0, 0.05)     # momentum of a capped velocity
    assert qx == pytest.approx(10.0, rel=1e-15) and hc == pytest.approx(0.05, rel=1e-15)


# ---------------------------------------------------------- sediment laws --

def zhang_settling(d, R, nu=1.002e-6):
    """Zhang (1961) settling velocity, written as the root of
    w^2 + 2*(13.95 nu/d) w - 1.09 (R-1) g d = 0."""
    a = 13.95 * nu / d
    c = 1.09 * (R - 1) * G * d
    return c / (a + math.sqrt(a * a + c))   # the positive root, cancellation-free form


def test_bedload_ashida_michiue_textbook(oracle, sed):
    p, s = sed
    R, d, poro, thc = p["R"], p["DS"], p["porosity"], p["OC"]
    cf = 0.025 / 8.0                   # bed shear stress tau/rho = cf |u|^2
    for h, un, ut in [(0.5, 1.2, 0.3), (2.0, -0.8, 0.9), (0.01, 2.0, 0.0), (1.0, 0.1, 0.0)]:
        u = math.hypot(un, ut)
        theta = cf * u * u / ((R - 1) * G * d)
        qstar = 17.0 * max(theta - thc, 0) * max(math.sqrt(theta) - math.sqrt(thc), 0)
        qb = qstar * math.sqrt((R - 1) * G * d ** 3) / (1 - poro) * un / u
        got = oracle.bedload_flux(p, s, h, un, ut)
        assert got == pytest.approx(-qb, rel=1e-12, abs=1e-300)
    # depth gates h > 5 d and h > HQB
    assert oracle.bedload_flux(p, s, 5 * d * 0.99, 2.0, 0.0) == 0.0


def test_bedload_grass(oracle, sed):
    p, s = sed
    s = dict(s, bedload_law=S.BEDLOAD_GRASS, grass_ag=1e-3)
    # q_b = A_g / (1 - p) u_n |u|^2 (|u| = 1.5 here, so |u| and |u|^2 differ)
    got = oracle.bedload_flux(p, s, 1.0, 1.2, 0.9)
    assert got == pytest.approx(-1e-3 / (1 - p["porosity"]) * 1.2 * 1.5 ** 2, rel=1e-14)
    assert oracle.bedload_flux(p, s, 5 * p["DS"] * 0.99, 1.2, 0.9) == 0.0


def test_erosion_deposition_garcia_parker(oracle, sed):
    p, s = sed
    R, d, poro, eps_h = p["R"], p["DS"], p["porosity"], p["epsilon_h"]
    nu = 1.002e-6
    vs = zhang_settling(d, R, nu)
    Rp = d * math.sqrt((R - 1) * G * d) / nu
    a1, a2 = (1.0, 0.6) if Rp > 2.36 else (0.586, 1.23)
    r0 = 1.64 + 0.4 * (d / p["Dsg"]) ** 1.64           # near-bed / depth-mean concentration ratio
    for h, qx, qy, hc in [(1.0, 1.5, 0.0, 0.0), (0.3, 0.6, 0.6, 0.003), (2.0, 0.2, 0.0, 0.04), (1.0, 9.0, 0.0, 0.0)]:
        u = math.hypot(qx, qy) / h
        ustar = math.sqrt(p["FR"]) * u
        Zu = a1 * Rp ** a2 * ustar / vs
        A = 1.3e-7
        # Garcia-Parker E = A Z^5 / (1 + (A/0.3) Z^5); the reference rounds
        # A/0.3 = 4.33e-7 to 4.3e-7 (SedFlux.cuh:68)
        Es = A * Zu ** 5 / (1 + 4.3e-7 * Zu ** 5)
        Fe = min(vs * (1 - poro) * Es, 0.1)
        Fd = min(vs * r0 * hc / (h + eps_h), 0.1)
        got = oracle.erosion_rate(p, s, h, qx, qy, hc)
        assert got == pytest.approx(Fe - Fd, rel=1e-9, abs=1e-18)


# ------------------------------------------------------------- edge flux ----

def flux_phys(h, qn, qt, hc):
    u = qn / h if h > 0 else 0.0
    return np.array([qn, qn * u, qt * u, 0.0, hc * u])


def test_edge_consistency_equal_states(oracle, sed):
    p, s = sed
    s = dict(s, bedload=0)
    st = np.array([1.7, 0.9, -0.4, 3.0, 0.02])
    dm, dp = oracle.edge_flux(p, s, st, st, (1, 0), 1)
    F = flux_phys(st[0] + 0.0, st[1], st[2], st[4])
    u = st[1] / (st[0] + np.finfo(float).eps)
    F = np.array([st[0] * u, st[0] * u * u, st[0] * u * (st[2] / (st[0] + np.finfo(float).eps)), 0, st[4] / (st[0] + np.finfo(float).eps) * st[0] * u])
    np.testing.assert_allclose(dm, F, rtol=1e-14, atol=1e-300)
    np.testing.assert_array_equal(dp, -dm)


def test_edge_lake_at_rest_is_zero(oracle, sed):
    p, s = sed
    # dyadic values keep eta = h - H exact on both sides; last: emerged right cell
    for Hl, Hr, eta in [(3.0, 1.0, 0.0), (2.0, 5.0, 0.5), (2.0, -1.0, 0.0), (7.25, 0.5, -0.25)]:
        L = np.array([max(eta + Hl, 0), 0, 0, Hl, 0])
        R = np.array([max(eta + Hr, 0), 0, 0, Hr, 0])
        dm, dp = oracle.edge_flux(p, s, L, R, (1, 0), 1)
        assert np.all(dm == 0) and np.all(dp == 0)


def hll_textbook(hL, uL, vL, hR, uR, vR):
    """HLL flux of the 2-D SWE (x-split) with Einfeldt-type speeds from the
    Roe average; flux includes the pressure g h^2 / 2."""
    cL, cR = math.sqrt(G * hL), math.sqrt(G * hR)
    ut = (math.sqrt(hL) * uL + math.sqrt(hR) * uR) / (math.sqrt(hL) + math.sqrt(hR))
    ct = math.sqrt(G * 0.5 * (hL + hR))
    SL, SR = min(uL - cL, ut - ct), max(uR + cR, ut + ct)
    UL = np.array([hL, hL * uL, hL * vL])
    UR = np.array([hR, hR * uR, hR * vR])
    FL = np.array([hL * uL, hL * uL * uL + 0.5 * G * hL * hL, hL * uL * vL])
    FR = np.array([hR * uR, hR * uR * uR + 0.5 * G * hR * hR, hR * uR * vR])
    if SL >= 0:
        return FL
    if SR <= 0:
        return FR
    return (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL)


@pytest.mark.parametrize("case", [(2.0, 0.5, 0.1, 1.0, -0.3, 0.2), (1.0, 0.0, 0.0, 3.0, 0.0, 0.0),
                                  (0.5, 2.5, 0.0, 0.4, 2.2, 1.0), (3.0, -3.0, 0.3, 2.9, -2.5, 0.0)])
def test_edge_flat_bed_matches_textbook_hll(oracle, sed, case):
    p, s = sed
    s = dict(s, bedload=0)
    hL, uL, vL, hR, uR, vR = case
    L = np.array([hL, hL * uL, hL * vL, 0.0, 0.0])
    R = np.array([hR, hR * uR, hR * vR, 0.0, 0.0])
    dm, dp = oracle.edge_flux(p, s, L, R, (1, 0), 0)
    F = hll_textbook(hL, uL, vL, hR, uR, vR)
    # the path-conservative form carries the pressure of the left state in the cell
    assert dm[0] == pytest.approx(F[0], rel=1e-12, abs=1e-12)
    assert dm[1] + 0.5 * G * hL * hL == pytest.approx(F[1], rel=1e-12)
    assert dm[2] == pytest.approx(F[2], rel=1e-12, abs=1e-12)
    # conservation identities of the fluctuations
    assert dm[0] + dp[0] == pytest.approx(0.0, abs=1e-13 * max(1, abs(dm[0])))
    hm = 0.5 * (hL + hR)
    assert dm[1] + dp[1] == pytest.approx(G * hm * (hR - hL), rel=1e-12, abs=1e-12)


def test_edge_mirror_and_rotation_symmetry(oracle, sed):
    p, s = sed
    rng = np.random.default_rng(3)
    for _ in range(300):
        L = np.array([rng.uniform(0, 3), rng.normal(0, 2), rng.normal(0, 2), rng.uniform(-1, 3), 0.0])
        R = np.array([rng.uniform(0, 3), rng.normal(0, 2), rng.normal(0, 2), rng.uniform(-1, 3), 0.0])
        L[4], R[4] = 0.01 * L[0], 0.02 * R[0]
        fx = int(rng.integers(0, 2))
        dm, dp = oracle.edge_flux(p, s, L, R, (1, 0), fx)
        # mirror x -> -x: swap sides and negate qx; the fluctuation sent to
        # the left cell becomes the one sent to the right cell (normal
        # momentum odd, every other component even)
        M = np.array([1, -1, 1, 1, 1.0])
        dm2, dp2 = oracle.edge_flux(p, s, R * M, L * M, (1, 0), fx)
        np.testing.assert_array_equal(dm2, dp * M)
        np.testing.assert_array_equal(dp2, dm * M)
        # rotation by 90 degrees: the y-edge of (qx, qy) is the x-edge of (qy, -qx)
        dmy, dpy = oracle.edge_flux(p, s, L, R, (0, 1), fx)
        Lr = np.array([L[0], L[2], -L[1], L[3], L[4]])
        Rr = np.array([R[0], R[2], -R[1], R[3], R[4]])
        dmx, dpx = oracle.edge_flux(p, s, Lr, Rr, (1, 0), fx)
        np.testing.assert_array_equal(dmy, [dmx[0], -dmx[2], dmx[1], dmx[3], dmx[4]])
        np.testing.assert_array_equal(dpy, [dpx[0], -dpx[2], dpx[1], dpx[3], dpx[4]])


# ------------------------------------------------------------ dt reduction --

def test_cell_dt_closed_form(oracle, sed):
    p, _ = sed
    h, qx, qy = 2.0, 3.0, -1.0
    c = math.sqrt(G * h)
    want = p["CFL"] * 0.5 * min(p["dx"] / (1.5 + c), p["dy"] / (0.5 + c))
    assert oracle.cell_dt(p, h, qx, qy) == pytest.approx(want, rel=1e-15)
    assert oracle.cell_dt(p, 0.0, 0.0, 0.0) == math.inf


def test_min_reduction_narrowing():
    import pyoracle
    x = 0.1 + 1e-12                        # not representable in float
    v = np.full(5000, 10.0)
    v[1234] = x
    v[10] = np.nan
    assert pyoracle.min_reduction(v, 1) == float(np.float32(x))
    assert pyoracle.min_reduction(v, 1) != x
    assert pyoracle.min_reduction(v[:4096], 1) == x                # small grids: exact double minimum
    assert pyoracle.min_reduction(v, 0) == x
    assert pyoracle.min_reduction(np.full(5000, np.inf), 1) == 1e30   # all dry


# ------------------------------------------------------------ grid level ----

def test_uniform_periodic_state_friction_and_exchange_closed_form(oracle):
    """A uniform state on a periodic grid has zero net flux, so one RK2 step is
    the closed form q1 = q/(1+c), q2 = q1/(1+c), q' = (q + q2)/2 with
    c = dt FR |u|/(h+eps) (friction uses U^n in both stages), and the bed/
    suspension exchange moves dt*phi per stage from H to hc."""
    n = 8
    p = S.base_params(n, n, 10.0, 10.0, friction=0.025)
    s = S.default_scheme(bc=[S.BC_PERIODIC] * 4, bedload=0)
    h, q, H, hc = 1.3, 0.9, 4.0, 0.0
    V = S.pack_state(np.full((n, n), h), np.full((n, n), q), np.zeros((n, n)), np.full((n, n), H),
                     np.full((n, n), hc))
    o = oracle.OracleSim(p, s, V, np.ones((n, n), np.int32))
    dt = o.dt
    eps = np.finfo(float).eps
    u = q / (h + eps)
    c = dt * p["FR"] * u / (h + eps)
    phi = oracle.erosion_rate(p, s, h, q, 0.0, hc)
    o.step(1)
    W = o.state()
    q2 = q / (1 + c) / (1 + c)
    qn = 0.5 * q + 0.5 * q2
    assert np.allclose(W[1], qn * h / (h + eps) * 1.0, rtol=1e-14)
    assert np.all(W[0] == h)
    # H: stage 1 H - dt phi; stage 2 (H - dt phi) - dt phi; RK average
    Hn = 0.5 * H + 0.5 * (H - 2 * dt * phi)
    assert np.allclose(W[3], Hn, rtol=1e-14, atol=1e-18)
    assert np.allclose(W[3] + W[4], H + hc, rtol=1e-15)


def test_uniform_periodic_state_manning_closed_form(oracle):
    """The Manning law of the plain-SWE configuration, semi-implicit like the
    reference's quadratic law: q <- q / (1 + c) with
    c = dt g n^2 |u| / h^(4/3), from U^n in both stages."""
    n = 6
    p = S.base_params(n, n, 10.0, 10.0)
    s = S.default_scheme(bc=[S.BC_PERIODIC] * 4, bedload=0, suspension=0, friction_law=S.FRIC_MANNING,
                         manning_n=0.03)
    h, q = 0.7, 1.1
    V = S.pack_state(np.full((n, n), h), np.full((n, n), q), np.zeros((n, n)), np.full((n, n), 2.0))
    o = oracle.OracleSim(p, s, V, np.ones((n, n), np.int32))
    dt = o.dt
    eps = np.finfo(float).eps
    u = q / (h + eps)
    c = dt * G * 0.03 ** 2 * u / (h + eps) ** (4.0 / 3.0)
    o.step(1)
    qn = 0.5 * q + 0.5 * q / (1 + c) / (1 + c)
    assert np.allclose(o.state()[1], qn * h / (h + eps), rtol=1e-14, atol=0)


def test_lake_at_rest_grid(oracle):
    sc = S.lake_at_rest(n=48)
    o = oracle.from_scenario(sc)
    o.step(20)
    np.testing.assert_array_equal(o.state(), sc.V)


def test_mass_conservation_walls(oracle):
    # bedload only, over wet/dry terrain: water and bed volume are conserved
    sc = S.gaussian_source(n=40, dx=50.0, seed=2, suspension=0)
    o = oracle.from_scenario(sc)
    m0, b0 = sc.V[0].sum(), sc.V[3].sum()
    o.step(40)
    W = o.state()
    assert abs(W[0].sum() - m0) <= 1e-13 * m0
    assert abs(W[3].sum() - b0) <= 1e-13 * abs(b0)
    assert np.abs(W[3] - sc.V[3]).max() > 0   # the bed did move


def test_sediment_conservation_deep_water(oracle):
    # Suspension on: Sum(H + hc) is conserved while no cell is shallow enough
    # for deposition to overdraw hc (the reference clips the negative hc in
    # state_filter, SedFlux.cuh:123-125, which creates sediment there).
    n = 40
    p = S.base_params(n, n, 20.0, 20.0)
    s = S.default_scheme()
    x, y = S.cell_centres(n, n, 20.0, 20.0)
    X, Y = np.meshgrid(x, y)
    H = 6.0 + 0.5 * np.sin(X / 90.0) * np.cos(Y / 70.0)
    h = H + 0.8 * np.exp(-((X - 300) ** 2 + (Y - 500) ** 2) / 2e4)
    V = S.pack_state(h, 1.2 * h, 0.3 * h, H, 1e-3 * h)
    o = oracle.OracleSim(p, s, V, S.nonerodible_mask(n, n, 7))
    m0, s0 = V[0].sum(), (V[3] + V[4]).sum()
    o.step(40)
    W = o.state()
    assert abs(W[0].sum() - m0) <= 1e-13 * m0
    assert abs((W[3] + W[4]).sum() - s0) <= 1e-13 * s0
    assert np.abs(W[4] - V[4]).max() > 1e-6   # sediment was exchanged


def test_mirror_symmetry_plain_swe(oracle):
    sc = S.plain_swe(n=48, dx=200.0)
    o = oracle.from_scenario(sc)
    o.step(30)
    W = o.state()
    np.testing.assert_array_equal(W[0], W[0][::-1])
    np.testing.assert_array_equal(W[1], W[1][::-1])
    np.testing.assert_array_equal(W[2], -W[2][::-1])
    assert np.abs(W[1]).max() > 0


def test_transpose_symmetry(oracle):
    sc = S.gaussian_source(n=32, dx=60.0, seed=4)
    V = sc.V
    Vt = np.ascontiguousarray(np.stack([V[0].T, V[2].T, V[1].T, V[3].T, V[4].T]))
    a = oracle.OracleSim(sc.params, sc.scheme, V, sc.fixH)
    b = oracle.OracleSim(sc.params, sc.scheme, Vt, np.ascontiguousarray(sc.fixH.T))
    a.step(25)
    b.step(25)
    A, B = a.state(), b.state()
    np.testing.assert_array_equal(A, np.stack([B[0].T, B[2].T, B[1].T, B[3].T, B[4].T]))


def ritter(x, t, h0):
    c0 = math.sqrt(G * h0)
    h = np.where(x <= -c0 * t, h0, np.where(x >= 2 * c0 * t, 0.0, (2 * c0 - x / t) ** 2 / (9 * G)))
    return h


def dambreak_error(n, h0=1.0, L=100.0, T=4.0):
    p = S.base_params(n, 1, L / n, L / n, friction=0.0)
    p["CFL"] = 0.45
    s = S.default_scheme(bedload=0, suspension=0, bc=[S.BC_WALL] * 4)
    x = (np.arange(n) + 0.5) * L / n - L / 2
    h = np.where(x < 0, h0, 0.0)[None, :]
    V = S.pack_state(h, np.zeros_like(h), np.zeros_like(h), np.zeros_like(h))
    import pyoracle
    o = pyoracle.OracleSim(p, s, V, np.ones((1, n), np.int32))
    while o.t < T:
        o.step(1)
    W = o.state()[0, 0]
    return np.abs(W - ritter(x, o.t, h0)).mean(), o.t


def test_ritter_dam_break_convergence():
    e1, _ = dambreak_error(200)
    e2, _ = dambreak_error(800)
    assert e1 < 0.02
    assert e2 < 0.6 * e1          # converges to the analytic dry-bed solution


# ------------------------------------------------- Carrier-Greenspan runup --

def test_carrier_greenspan_solution_satisfies_the_equations():
    """The closed form itself: centred differences of (eta, u) at wet points
    satisfy the non-dimensional shallow-water equations to truncation error,
    and at small amplitude it reduces to the linear standing wave
    eta = -(A/4) J0(4 sqrt(x)) sin(2t)."""
    import analytic as AN
    from scipy.special import j0
    A, h = 0.5, 1e-4
    for x0, t0 in [(0.6, 0.4), (0.05, 1.3), (1.2, 2.9), (0.3, -0.7)]:
        E, U = {}, {}
        for a in (-1, 0, 1):
            e, u = AN.cg_at(x0 + a * h, np.array([t0 - h, t0, t0 + h]), A)
            for b in (-1, 0, 1):
                E[a, b], U[a, b] = e[b + 1], u[b + 1]
        d = lambda a, b: x0 + a * h + E[a, b]
        mass = (d(0, 1) - d(0, -1)) / (2 * h) + (d(1, 0) * U[1, 0] - d(-1, 0) * U[-1, 0]) / (2 * h)
        mom = ((U[0, 1] - U[0, -1]) + U[0, 0] * (U[1, 0] - U[-1, 0]) + (E[1, 0] - E[-1, 0])) / (2 * h)
        assert abs(mass) < 1e-6 and abs(mom) < 1e-6, (x0, t0, mass, mom)
    # the field solver (sigma grid) agrees with the point solver
    xs = np.array([0.2, 0.7, 1.1])
    e1, u1, wet = AN.cg_field(xs, 0.9, A)
    e2, u2 = np.array([AN.cg_at(x, np.array([0.9]), A) for x in xs])[:, :, 0].T
    assert wet.all()
    np.testing.assert_allclose(e1, e2, atol=1e-7)
    np.testing.assert_allclose(u1, u2, atol=1e-7)
    # linear limit
    a = 1e-4
    for x0 in (0.1, 0.5, 1.0):
        t = np.array([0.3, 1.1])
        e, _ = AN.cg_at(x0, t, a)
        np.testing.assert_allclose(e, -a / 4 * j0(4 * np.sqrt(x0)) * np.sin(2 * t), rtol=2e-3, atol=1e-12)
    # shoreline: sigma = 0 gives u = (A/2) cos(lambda), eta = -(A/4) sin(lambda) - u^2/2
    assert AN.cg_runup(A) == pytest.approx(A / 4, rel=1e-9)   # max at sin(lambda) = -1


def _cg_run(dx, periods=1.25):
    import analytic as AN
    import pyoracle
    A = 0.5
    sc, B, x = AN.carrier_greenspan_scenario(dx, A=A)
    o = pyoracle.from_scenario(sc)
    t_end = periods * math.pi * B.T0          # the standing wave's period is pi T0
    while o.t < t_end:
        o.step(1)
    W = o.state()[:, 0]
    hmax = o.extrema()[0][0]
    e_an, _, wet_an = B.field(x, o.t, A)
    wet = (W[0] > 1e-3) & wet_an
    amp = A / 4 * B.Z0
    eta_err = np.abs((W[0] - W[3]) - e_an)[wet].mean() / amp
    runup = (-sc.V[3, 0][hmax > 1e-3]).max()          # highest wetted ground
    return eta_err, runup / (AN.cg_runup(A) * B.Z0) - 1.0


def test_carrier_greenspan_runup_convergence():
    """Moving-shoreline accuracy on a 1:20 plane beach (6.25 m runup): the
    free surface after 1.25 periods and the runup converge to the exact
    nonlinear solution as dx halves."""
    e8, r8 = _cg_run(8.0)
    e4, r4 = _cg_run(4.0)
    assert e8 < 5e-3 and e4 < 0.6 * e8, (e8, e4)
    assert abs(r4) < 0.03 and abs(r4) < abs(r8), (r8, r4)


def test_shared_denominator_quotient_is_the_ieee_quotient(oracle):
    """The FP64 GPU path divides several numerators by one denominator with
    one reciprocal and an FMA residual step (Markstein); on seeded random
    pairs over the solver's ranges it equals x / d every time.  A variant
    without the residual step (q = x * y alone) must fail the same check,
    so the check can tell."""
    assert oracle.lib().orc_shared_quotient_mismatches(2_000_000, 12345) == 0
    rng = np.random.default_rng(1)
    d = np.abs(rng.normal(size=100000)) + 1e-3
    x = rng.normal(size=100000)
    assert np.any(x * (1.0 / d) != x / d)


# ------------------------------------------- face reconstruction, one cell --

def _triple(h, u, v, H, c):
    """(minus, cell, plus) states from primitive values at k = -1, 0, 1."""
    return [np.array([h[k], h[k] * u[k], h[k] * v[k], H[k], h[k] * c[k]]) for k in range(3)]


@pytest.mark.parametrize("lim", [S.LIM_VANLEER, S.LIM_MINMOD])
def test_faces_reproduce_linear_data(oracle, lim):
    """Second-order MUSCL on data whose primitive variables (h, u, v, H, c)
    vary linearly: every limiter with phi(s, s) = s returns the slope itself,
    so depth, bed and concentration faces are the values at the faces, the
    face hc is c(face) * h(face), and the face momentum is the first-order
    Taylor expansion of h*u about the centre (product rule: h u' + u h')."""
    rng = np.random.default_rng(11)
    for _ in range(400):
        h0, a = rng.uniform(1, 5), rng.uniform(-0.4, 0.4)
        u0, b = rng.uniform(-2, 2), rng.uniform(-0.3, 0.3)
        v0, e = rng.uniform(-2, 2), rng.uniform(-0.3, 0.3)
        H0, dH = rng.uniform(-2, 8), rng.uniform(-0.5, 0.5)
        c0, dc = rng.uniform(0.01, 0.05), rng.uniform(-0.004, 0.004)
        k = np.array([-1.0, 0.0, 1.0])
        m, c, p = _triple(h0 + a * k, u0 + b * k, v0 + e * k, H0 + dH * k, c0 + dc * k)
        # the surface must stay wet across the triple for the hydrostatic
        # depths to be the depths themselves
        fm, fp, deta = oracle.cell_faces(lim, m, c, p)
        for s, f in ((-0.5, fm), (0.5, fp)):
            hf, cf = h0 + s * a, c0 + s * dc
            assert f[0] == pytest.approx(hf, rel=1e-13)
            assert f[3] == pytest.approx(H0 + s * dH, rel=1e-13, abs=1e-13)
            assert f[4] == pytest.approx(cf * hf, rel=1e-11)
            assert f[1] == pytest.approx(h0 * u0 + s * (h0 * b + u0 * a), rel=1e-12, abs=1e-12)
            assert f[2] == pytest.approx(h0 * v0 + s * (h0 * e + v0 * a), rel=1e-12, abs=1e-12)
        # the limited surface slope is the slope of eta = h - H
        assert deta == pytest.approx(a - dH, rel=1e-12, abs=1e-13)


def test_faces_depth_positivity_and_mean(oracle):
    """Whatever the neighbours, the two depth faces of a cell lie in
    [0, 2 h_i] and average to h_i (Reconstructor.cu:150-152 bounds the depth
    slope so that the reconstruction is positivity-preserving and
    conservative); concentration faces are non-negative.  Neighbour depths
    are drawn wet, dry and (as a careless caller could pass) negative,
    which is what drives the limited slope past 2 h_i."""
    rng = np.random.default_rng(12)
    clipped = 0
    for it in range(3000):
        hi = rng.choice([rng.uniform(0, 1e-3), rng.uniform(0, 3)])
        hs = [rng.choice([0.0, rng.uniform(0, 6), -rng.uniform(0, 4)]) for _ in range(2)]
        h = np.array([hs[0], hi, hs[1]])
        u = rng.normal(0, 1, 3)
        Hb = rng.uniform(-3, 5, 3)
        cc = rng.uniform(0, 0.05, 3)
        lim = S.LIM_VANLEER if it % 2 else S.LIM_MINMOD
        m, c, p = [np.array([h[k], h[k] * u[k], 0.0, Hb[k], h[k] * cc[k]]) for k in range(3)]
        fm, fp, _ = oracle.cell_faces(lim, m, c, p)
        assert 0.0 <= fm[0] <= 2 * hi and 0.0 <= fp[0] <= 2 * hi, (h, fm[0], fp[0])
        assert 0.5 * (fm[0] + fp[0]) == pytest.approx(hi, rel=1e-15, abs=1e-300), (h, fm[0], fp[0])
        assert fm[4] >= 0.0 and fp[4] >= 0.0
        clipped += fm[0] == 0.0 or fp[0] == 0.0
    assert clipped > 100          # the bound was exercised


# ----------------------------------------------- suspended-load edge row --

@pytest.mark.parametrize("case", [(2.0, 0.5, 0.1, 1.0, -0.3, 0.2), (1.0, 0.0, 0.0, 3.0, 0.0, 0.0),
                                  (0.5, 2.5, 0.0, 0.4, 2.2, 1.0), (3.0, -3.0, 0.3, 2.9, -2.5, 0.0),
                                  (1.2, 0.4, 0.0, 1.0, 0.45, 0.0)])
def test_edge_suspended_row_is_textbook_hll(oracle, sed, case):
    """On a flat bed the suspended-load row is the HLL flux of a passive
    scalar carried with the water, with the same wave speeds:
    (S_R F_L - S_L F_R + S_L S_R (hc_R - hc_L)) / (S_R - S_L), F = hc u."""
    p, s = sed
    s = dict(s, bedload=0)
    hL, uL, vL, hR, uR, vR = case
    cL, cR = 0.013, 0.041
    L = np.array([hL, hL * uL, hL * vL, 0.0, cL * hL])
    R = np.array([hR, hR * uR, hR * vR, 0.0, cR * hR])
    dm, dp = oracle.edge_flux(p, s, L, R, (1, 0), 0)
    sL, sR = math.sqrt(G * hL), math.sqrt(G * hR)
    ut = (math.sqrt(hL) * uL + math.sqrt(hR) * uR) / (math.sqrt(hL) + math.sqrt(hR))
    ct = math.sqrt(G * 0.5 * (hL + hR))
    SL, SR = min(uL - sL, ut - ct), max(uR + sR, ut + ct)
    FL, FR = cL * hL * uL, cR * hR * uR
    if SL >= 0:
        want = FL
    elif SR <= 0:
        want = FR
    else:
        want = (SR * FL - SL * FR + SL * SR * (cR * hR - cL * hL)) / (SR - SL)
    assert dm[4] == pytest.approx(want, rel=1e-12, abs=1e-15)
    assert dp[4] == pytest.approx(-dm[4], rel=1e-14, abs=1e-300)
    assert dm[3] == 0.0 and dp[3] == 0.0


# ----------------------------------------------------- non-erodible mask --

def test_nonerodible_cells_keep_their_bed(oracle):
    """fixH = 0 marks a non-erodible cell: bedload across every edge of such
    a cell is zero (SedFlux.cuh:175; interior edges take fixH[p0]*fixH[p1],
    boundary edges the cell's own flag, VolumeSolver.cu:256,297,313,347,
    386,402), so with suspension off its bed stays bit-identical while the
    bed elsewhere moves; with every cell non-erodible nothing moves."""
    sc = S.gaussian_source(n=40, dx=50.0, seed=2, suspension=0)
    V = sc.V.copy()
    V[1] = 0.8 * V[0]            # a current, so bedload runs everywhere wet
    fix = S.nonerodible_mask(40, 40, 3)
    fix[:, :3] = 0               # a masked strip along the left wall
    o = oracle.OracleSim(sc.params, sc.scheme, V, fix)
    o.step(30)
    W = o.state()
    masked = fix == 0
    assert masked.sum() > 50
    np.testing.assert_array_equal(W[3][masked], V[3][masked])
    assert np.abs(W[3] - V[3])[~masked].max() > 1e-6
    assert abs(W[3].sum() - V[3].sum()) <= 1e-13 * abs(V[3].sum())
    o2 = oracle.OracleSim(sc.params, sc.scheme, V, np.zeros_like(fix))
    o2.step(30)
    np.testing.assert_array_equal(o2.state()[3], V[3])


# ----------------------------------------- bed waves (linear Exner theory) --

def _bedload_am(u, Q, Opar, OC):
    th = Opar * u * u
    return 17.0 * Q * max(th - OC, 0.0) * max(math.sqrt(th) - math.sqrt(OC), 0.0)


def _steady_depth(q, E, H, subcritical):
    """Depth of steady frictionless flow with discharge q and head E over
    ground -H: q^2 / (2 h^2) + g (h - H) = E, on the sub- or supercritical
    branch (Newton from the matching side of the critical depth)."""
    h = (q * q / G) ** (1.0 / 3.0) * (3.0 if subcritical else 0.3)
    for _ in range(60):
        h -= (q * q / (2 * h * h) + G * (h - H) - E) / (G - q * q / h ** 3)
    return h


def _bed_hump_run(oracle, u0, h0, Q, Opar, A=0.05, sig=6.0, n=200, disp=20.0):
    p = S.base_params(n, 1, 1.0, 1.0, friction=0.0)
    p.update(Q=Q, Opar=Opar, FR=0.0)
    s = S.default_scheme(bc=[S.BC_PERIODIC, S.BC_PERIODIC, S.BC_WALL, S.BC_WALL], suspension=0)
    x = np.arange(n) + 0.5
    H0 = 10.0
    bump = A * np.exp(-(x - n / 2) ** 2 / (2 * sig * sig))     # bed raised by the hump
    H = H0 - bump
    q = u0 * h0
    E = 0.5 * u0 * u0 + G * (h0 - H0)
    fr2 = u0 * u0 / (G * h0)
    h = np.array([_steady_depth(q, E, Hi, fr2 < 1) for Hi in H])
    V = S.pack_state(h[None], np.full((1, n), q), np.zeros((1, n)), H[None])
    o = oracle.OracleSim(p, s, V, np.ones((1, n), np.int32))
    # linearised Exner celerity under quasi-steady flow:
    # c_b = q_b'(u0) u0 / (h0 (1 - Fr^2))
    du = 1e-6 * u0
    dqb = (_bedload_am(u0 + du, Q, Opar, p["OC"]) - _bedload_am(u0 - du, Q, Opar, p["OC"])) / (2 * du)
    cb = dqb * u0 / (h0 * (1.0 - fr2))
    T = abs(disp / cb)
    while o.t < T:
        o.step(1)
    m = H0 - o.state()[3, 0]
    shift = (x * m).sum() / m.sum() - (x * bump).sum() / bump.sum()
    return shift, cb * o.t, m, A


@pytest.mark.parametrize("regime", ["subcritical", "supercritical"])
def test_bed_hump_migrates_at_the_exner_celerity(oracle, regime):
    """Bedload only, steady current over a small bed hump on a periodic
    channel.  Linear Exner theory: the hump travels at
    c_b = q_b'(u) u / (h (1 - Fr^2)), downstream in subcritical flow and
    upstream (antidunes) in supercritical flow.  The bed row of the edge
    flux is upwinded by the sign of that celerity (VolumeSolver.cu:100-120),
    so the hump must move the right way at about c_b and stay a hump: no
    growth, no new trough ahead of or behind it."""
    if regime == "subcritical":
        shift, want, m, A = _bed_hump_run(oracle, 1.0, 1.0, 9e-4, 1.0)
    else:
        shift, want, m, A = _bed_hump_run(oracle, 4.0, 0.5, 3.5e-5, 1.0)
    assert (shift > 0) == (regime == "subcritical")
    assert shift == pytest.approx(want, rel=0.15), (shift, want)
    assert m.max() <= A
    assert m.min() >= -0.005 * A, m.min()


# ------------------------------------------- suspended load in a current --

def test_suspended_pulse_advects_with_the_flow(oracle):
    """Suspension and bedload off, uniform current on a flat bed, a
    concentration plateau: hc is a passive scalar, so its total is
    conserved, its centroid moves at u t (exactly at first order; the
    limited slopes' sum shifts it by far less than 0.1 % here), and a TVD
    reconstruction with an HLL flux makes no new extrema (0 <= c <= c_max).  The concentration faces
    (Reconstructor.cu:188-194) and the hc row's viscosity
    (VolumeSolver.cu:98,105) both enter."""
    n = 256
    p = S.base_params(n, 1, 2.0, 2.0, friction=0.0)
    s = S.default_scheme(bc=[S.BC_PERIODIC, S.BC_PERIODIC, S.BC_WALL, S.BC_WALL], suspension=0, bedload=0)
    x = (np.arange(n) + 0.5) * 2.0
    h, u = 2.0, 1.5
    c = np.where(np.abs(x - 150.0) < 30.0, 0.02, 0.0) + 0.01 * np.exp(-((x - 300.0) / 12.0) ** 2)
    V = S.pack_state(np.full((1, n), h), np.full((1, n), h * u), np.zeros((1, n)), np.full((1, n), 5.0),
                     (h * c)[None])
    o = oracle.OracleSim(p, s, V, np.ones((1, n), np.int32))
    while o.t < 60.0:
        o.step(1)
    W = o.state()[:, 0]
    hc0, hc1 = V[4, 0], W[4]
    assert abs(hc1.sum() - hc0.sum()) <= 1e-13 * hc0.sum()
    shift = (x * hc1).sum() / hc1.sum() - (x * hc0).sum() / hc0.sum()
    assert shift == pytest.approx(u * o.t, rel=1e-3)
    cn = hc1 / W[0]
    assert cn.min() >= 0.0 and cn.max() <= 0.02 * (1 + 1e-12), (cn.min(), cn.max())
    # and the plateau was smeared but not destroyed
    assert cn.max() > 0.018


# ------------------------------------------ exact symmetries of the grid --

def _transpose_state(V):
    return np.ascontiguousarray(np.stack([V[0].T, V[2].T, V[1].T, V[3].T, V[4].T]))


def test_transpose_symmetry_non_square_cells(oracle):
    """Transposing the grid, swapping qx with qy and dx with dy gives the
    transposed result bit for bit: the x and y sweeps, the flux divergence
    (dt/dx, dt/dy) and the cell dt (dx / lambda_x, dy / lambda_y) each use
    their own axis's spacing."""
    sc = S.gaussian_source(n=32, dx=60.0, seed=4)
    pa = dict(sc.params, dy=35.0, ymax=32 * 35.0)
    pb = dict(sc.params, dx=35.0, xmax=32 * 35.0)
    V = sc.V.copy()
    V[1] = 0.3 * V[0]
    V[2] = -0.2 * V[0]
    a = oracle.OracleSim(pa, sc.scheme, V, sc.fixH)
    b = oracle.OracleSim(pb, sc.scheme, _transpose_state(V), np.ascontiguousarray(sc.fixH.T))
    assert a.dt == b.dt
    a.step(25)
    b.step(25)
    np.testing.assert_array_equal(a.state(), _transpose_state(b.state()).reshape(a.state().shape))
    assert a.dt == b.dt and a.t == b.t


def test_cell_dt_non_square(oracle, sed):
    p, _ = sed
    p = dict(p, dx=12.0, dy=5.0)
    h, qx, qy = 3.0, 1.5, 4.5
    c = math.sqrt(G * h)
    want = p["CFL"] * 0.5 * min(12.0 / (0.5 + c), 5.0 / (1.5 + c))
    assert oracle.cell_dt(p, h, qx, qy) == pytest.approx(want, rel=1e-15)
    assert oracle.cell_dt(p, h, qy, qx) == pytest.approx(p["CFL"] * 0.5 * min(12.0 / (1.5 + c), 5.0 / (0.5 + c)),
                                                         rel=1e-15)


def test_periodic_shift_invariance(oracle):
    """On a doubly periodic grid the scheme commutes with a cyclic shift of
    the cells: neighbour wrap-around (Reconstructor.cu:130-133, 209-212) and
    the periodic edges (VolumeSolver.cu:217-408) must treat the seam like
    any interior edge."""
    n, m = 24, 20
    rng = np.random.default_rng(21)
    p = S.base_params(n, m, 15.0, 15.0)
    s = S.default_scheme(bc=[S.BC_PERIODIC] * 4)
    x, y = S.cell_centres(n, m, 15.0, 15.0)
    X, Y = np.meshgrid(x, y)
    H = 3.0 + 1.5 * np.sin(2 * np.pi * X / (n * 15.0)) * np.cos(2 * np.pi * Y / (m * 15.0)) \
        + 0.2 * rng.standard_normal((m, n))
    h = np.maximum(H + 0.6 * np.exp(-((X - 100) ** 2 + (Y - 180) ** 2) / 3000.0), 0.0)
    V = S.pack_state(h, 0.9 * h, -0.4 * h, H, 2e-3 * h)
    ones = np.ones((m, n), np.int32)
    sh = (7, 5)
    a = oracle.OracleSim(p, s, V, ones)
    b = oracle.OracleSim(p, s, np.ascontiguousarray(np.roll(V, sh, axis=(1, 2))), ones)
    a.step(30)
    b.step(30)
    np.testing.assert_array_equal(np.roll(a.state(), sh, axis=(1, 2)), b.state())
    assert np.abs(a.state() - V).max() > 1e-3


def test_running_extrema_brute_force(oracle):
    """The running extrema (Volume_kernel.cu:79-110) after every step: max h
    and max eta over the steps where h > epsilon_h, max H over all steps;
    checked against the states read back one step at a time."""
    sc = S.gaussian_source(n=24, dx=80.0, seed=6, extrema=1)
    V = sc.V.copy()
    V[1] = 1.5 * V[0]
    o = oracle.OracleSim(sc.params, sc.scheme, V, sc.fixH)
    eh = sc.params["epsilon_h"]
    hmax = np.full(V[0].shape, -1e30)
    bmax = np.full(V[0].shape, -1e30)
    emax = np.full(V[0].shape, -1e30)
    for _ in range(25):
        o.step(1)
        W = o.state()
        wet = W[0] > eh
        hmax = np.where(wet, np.maximum(hmax, W[0]), hmax)
        emax = np.where(wet, np.maximum(emax, W[0] - W[3]), emax)
        bmax = np.maximum(bmax, W[3])
    got = o.extrema()
    np.testing.assert_array_equal(got[0], hmax)
    np.testing.assert_array_equal(got[1], bmax)
    np.testing.assert_array_equal(got[2], emax)
    assert (hmax > -1e30).any() and (hmax == -1e30).any()   # wet and never-wet cells


def _inflow_case(side, n=24, m=16, dx=40.0):
    """A wave entering through one side (0 left, 1 right, 2 bottom, 3 top) of
    a basin whose depth varies along the entry direction; the CC slots of
    that side carry (eta(t), u, v, H) as VolumeSolver.cu:172-202 reads them
    (bottom i, right npx+j, top npx+npy+i, left 2npx+npy+j).  Built for the
    left side and mapped to the others by mirror and transpose."""
    x, _ = S.cell_centres(n, m, dx, dx)
    H = np.broadcast_to((12.0 - 8.0 * x / (n * dx))[None, :], (m, n)).copy()
    H[5, 9] += 1.5
    h = np.maximum(H, 0.0)
    tCC = np.arange(0.0, 400.0 + 1e-9, 20.0)
    eta = S.n_wave(tCC, 0.8, 200.0)
    ring = 2 * n + 2 * m
    CC = np.zeros((len(tCC), ring, 4))
    CC[:, 2 * n + m:, 0] = eta[:, None]
    CC[:, 2 * n + m:, 1] = eta[:, None] * np.sqrt(G / H[:, 0])[None, :]
    CC[:, 2 * n + m:, 2] = 0.05 * eta[:, None]
    CC[:, 2 * n + m:, 3] = H[:, 0][None, :]
    V = S.pack_state(h, 0 * h, 0 * h, H, 0 * h)
    fix = np.ones((m, n), np.int32)
    bc = [S.BC_INFLOW, S.BC_WALL, S.BC_WALL, S.BC_WALL]
    p = S.base_params(n, m, dx, dx)
    if side == 0:
        return p, bc, V, fix, CC, tCC
    if side == 1:       # mirror x: left slot j -> right slot npx+j, u -> -u
        V2 = V[:, :, ::-1].copy()
        V2[1] *= -1
        C2 = np.zeros_like(CC)
        C2[:, n:n + m] = CC[:, 2 * n + m:] * np.array([1, -1, 1, 1.0])
        return p, [S.BC_WALL, S.BC_INFLOW, S.BC_WALL, S.BC_WALL], np.ascontiguousarray(V2), fix, C2, tCC
    # transpose: left slot j -> bottom slot j on the n x m grid's transpose
    Vt = _transpose_state(V)
    pt = S.base_params(m, n, dx, dx)
    Ct = np.zeros((len(tCC), ring, 4))
    Ct[:, 0:m] = CC[:, 2 * n + m:][:, :, [0, 2, 1, 3]]
    fixt = np.ones((n, m), np.int32)
    if side == 2:
        return pt, [S.BC_WALL, S.BC_WALL, S.BC_INFLOW, S.BC_WALL], Vt, fixt, Ct, tCC
    # side 3, top: mirror the transposed case in y (bottom slot i -> top slot npx+npy+i)
    V3 = Vt[:, ::-1, :].copy()
    V3[2] *= -1
    C3 = np.zeros_like(Ct)
    C3[:, m + n:2 * m + n] = Ct[:, 0:m] * np.array([1, 1, -1, 1.0])
    return pt, [S.BC_WALL, S.BC_WALL, S.BC_WALL, S.BC_INFLOW], np.ascontiguousarray(V3), fixt, C3, tCC


def test_inflow_slots_of_the_four_sides(oracle):
    """The same inflow through each side of the basin, with the state
    mirrored or transposed to match, gives the mirrored or transposed
    result bit for bit; so each side reads its own slots of the CC ring in
    the reference's order, with its velocity components in the grid frame."""
    res = []
    for side in range(4):
        p, bc, V, fix, CC, tCC = _inflow_case(side)
        s = S.default_scheme(bc=bc, suspension=0, bedload=0)
        o = oracle.OracleSim(p, s, V, fix, CC, tCC)
        o.step(60)
        res.append((o.state(), o.t))
    W0 = res[0][0]
    assert np.abs(W0[0] - _inflow_case(0)[2][0]).max() > 0.1      # the wave came in
    W1 = res[1][0][:, :, ::-1].copy()
    W1[1] *= -1
    np.testing.assert_array_equal(W1, W0)
    W2 = _transpose_state(res[2][0])
    np.testing.assert_array_equal(W2, W0)
    W3 = res[3][0][:, ::-1, :].copy()
    W3[2] *= -1
    np.testing.assert_array_equal(_transpose_state(W3), W0)
    assert res[0][1] == res[1][1] == res[2][1] == res[3][1]


def test_shared_denominator_quotient_tiny_numerators(oracle):
    """Below 2^-968 the residual step of the shared-denominator quotient can
    underflow and miss x / d; the device then takes the IEEE division.  On
    numerators from the smallest subnormal to 2^-900 the unguarded form must
    miss somewhere (so the threshold matters) and the guarded form never."""
    L = oracle.lib()
    assert L.orc_shared_quotient_tiny_mismatches(1_000_000, 777, 0) > 0
    assert L.orc_shared_quotient_tiny_mismatches(1_000_000, 777, 1) == 0
