"""Seeded, deterministic input generators for the benchmark workloads.

Each generator returns a :class:`Scenario`: run parameters with the meaning
of the reference ``SimParams`` (TSUNER.h:10-30, derived fields as in
Problem_loader.cxx:337-342), the scheme switches that the reference fixes at
compile time (Sim_parameters.hxx:40-54), and the initial planar SoA state
``V[f, j, i]`` with f = h, qx, qy, H, hc (GPU/main_TSUNER.cu:104-111).

H is depth below datum (positive down), bed z = -H, and the initial depth is
h = max(0, eta + H) (Problem_loader.cxx:586-588, cini.cxx:34-42).  The
initial hc plane is zero (GPU/main_TSUNER.cu:110).

Inflow series use the reference CC layout (VolumeSolver.cu:172-202,
interpolations.cxx:246-332): per time level, (2*npx + 2*npy) boundary slots of
four values (eta, u, v, H); slot i is the bottom side, npx+j the right side,
npx+npy+i the top side and 2*npx+npy+j the left side.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

G = 9.81

BC_FREE, BC_WALL, BC_PERIODIC, BC_INFLOW = 0, 1, 2, 3
LIM_VANLEER, LIM_MINMOD = 0, 1
BEDLOAD_AM, BEDLOAD_GRASS = 0, 1
FRIC_CHEZY, FRIC_MANNING = 0, 1


def derive_params(p: dict) -> dict:
    """Derived fields exactly as Problem_loader.cxx:337-342 computes them."""
    p = dict(p)
    p["FR"] = p["FR"] / 8.0
    p["HQB"] = 5.0 * p["DS"]
    p["Q"] = p["DS"] * math.sqrt(G * (p["R"] - 1.0) * p["DS"]) / (1.0 - p["porosity"])
    p["Opar"] = p["FR"] / ((p["R"] - 1.0) * G * p["DS"])
    return p


def base_params(npx: int, npy: int, dx: float, dy: float, *, friction: float = 0.025,
                R: float = 2.65, DS: float = 2e-4, porosity: float = 0.4,
                OC: float = 0.0047, Dsg: float = 1e-4, CFL: float = 0.9,
                epsilon_h: float = 1e-4, total_time: float = 1e30,
                save_time: float = -1.0) -> dict:
    """data.dat values (data.dat:1-45) for a Cartesian grid anchored at 0."""
    p = dict(xmin=0.0, xmax=npx * dx, ymin=0.0, ymax=npy * dy, total_time=total_time,
             save_time=save_time, npx=npx, npy=npy, nvol=npx * npy, CFL=CFL, Q=0.0, R=R,
             porosity=porosity, DS=DS, FR=friction, Opar=0.0, OC=OC, Dsg=Dsg, HQB=0.0,
             vmax=10.0, epsilon_h=epsilon_h, dx=dx, dy=dy)
    return derive_params(p)


def default_scheme(**kw) -> dict:
    """The reference's compile-time configuration (Sim_parameters.hxx:45-51):
    second order, bedload and suspension on, walls on all sides, van Leer
    limiter, Ashida-Michiue bedload, quadratic friction, float-narrowed dt."""
    s = dict(order=2, bedload=1, suspension=1, bc=[BC_WALL] * 4, limiter=LIM_VANLEER,
             bedload_law=BEDLOAD_AM, friction_law=FRIC_CHEZY, dt_narrow=1, extrema=0,
             grass_ag=1e-3, manning_n=0.025)
    s.update(kw)
    return s


@dataclasses.dataclass
class Scenario:
    name: str
    params: dict
    scheme: dict
    V: np.ndarray                 # (5, npy, npx) float64
    fixH: np.ndarray              # (npy, npx) int32
    CC: Optional[np.ndarray] = None   # (T, 2*npx+2*npy, 4) float64
    tCC: Optional[np.ndarray] = None  # (T,) float64
    steps: int = 0

    @property
    def npx(self) -> int:
        return int(self.params["npx"])

    @property
    def npy(self) -> int:
        return int(self.params["npy"])

    @property
    def sediment(self) -> bool:
        return bool(self.scheme["bedload"] or self.scheme["suspension"])


def cell_centres(npx: int, npy: int, dx: float, dy: float):
    x = (np.arange(npx, dtype=np.float64) + 0.5) * dx
    y = (np.arange(npy, dtype=np.float64) + 0.5) * dy
    return x, y


def fractal_field(npx: int, npy: int, seed: int, rms: float = 2.0, exponent: float = -3.0) -> np.ndarray:
    """Isotropic Gaussian random field with power spectrum P(k) ~ k**exponent,
    scaled to the given RMS, from seeded white noise shaped in Fourier space."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((npy, npx))
    W = np.fft.rfft2(w)
    ky = np.fft.fftfreq(npy)[:, None]
    kx = np.fft.rfftfreq(npx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    amp = np.zeros_like(k)
    nz = k > 0
    amp[nz] = k[nz] ** (exponent / 2.0)
    f = np.fft.irfft2(W * amp, s=(npy, npx))
    f -= f.mean()
    f *= rms / np.sqrt(np.mean(f * f))
    return f


def nonerodible_mask(npx: int, npy: int, seed: int, nrect: int = 40, frac: float = 0.05) -> np.ndarray:
    """fixH (1 = erodible, 0 = fixed bed) with ``nrect`` seeded rectangles
    whose union covers ``frac`` of the domain (scale found by bisection)."""
    rng = np.random.default_rng(seed)
    fx = rng.uniform(0, 1, nrect)
    fy = rng.uniform(0, 1, nrect)
    rw = rng.uniform(0.5, 2.0, nrect)
    rh = rng.uniform(0.5, 2.0, nrect)

    def build(scale, nx, ny):
        m = np.ones((ny, nx), dtype=np.int32)
        for a, b, w_, h_ in zip(fx * nx, fy * ny, rw, rh):
            hw = max(1, int(round(scale * w_ * nx / 2)))
            hh = max(1, int(round(scale * h_ * ny / 2)))
            i0, i1 = max(0, int(a) - hw), min(nx, int(a) + hw)
            j0, j1 = max(0, int(b) - hh), min(ny, int(b) + hh)
            m[j0:j1, i0:i1] = 0
        return m

    # the covered fraction depends on the scale only, so bisect on a small
    # proxy grid and build the full-size mask once
    px, py = min(npx, 512), min(npy, 512)
    lo, hi = 0.0, 0.5
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        if 1.0 - build(mid, px, py).mean() < frac:
            lo = mid
        else:
            hi = mid
    return build(0.5 * (lo + hi), npx, npy)


def n_wave(t: np.ndarray, amplitude: float, period: float) -> np.ndarray:
    """Boundary N-wave: one full sine period (crest then trough) of the given
    amplitude, zero afterwards."""
    t = np.asarray(t, dtype=np.float64)
    eta = amplitude * np.sin(2.0 * np.pi * t / period)
    return np.where((t >= 0) & (t <= period), eta, 0.0)


def inflow_series(npx: int, npy: int, H: np.ndarray, eta_left: np.ndarray, tCC: np.ndarray) -> np.ndarray:
    """CC array (T, 2npx+2npy, 4) with the left side driven by eta_left(t) and
    u = eta * sqrt(g / H_b); other slots hold still water (eta 0, u v 0) over the
    adjacent boundary cell's depth, as interpolate_CC fills them."""
    T = len(tCC)
    ring = 2 * npx + 2 * npy
    CC = np.zeros((T, ring, 4), dtype=np.float64)
    CC[:, 0:npx, 3] = H[0, :]
    CC[:, npx:npx + npy, 3] = H[:, npx - 1]
    CC[:, npx + npy:2 * npx + npy, 3] = H[npy - 1, :]
    Hb = H[:, 0]
    CC[:, 2 * npx + npy:, 3] = Hb
    c = np.sqrt(G / np.maximum(Hb, 1e-3))
    CC[:, 2 * npx + npy:, 0] = eta_left[:, None]
    CC[:, 2 * npx + npy:, 1] = eta_left[:, None] * c[None, :]
    return CC


def pack_state(h, qx, qy, H, hc=None) -> np.ndarray:
    npy, npx = h.shape
    V = np.zeros((5, npy, npx), dtype=np.float64)
    V[0], V[1], V[2], V[3] = h, qx, qy, H
    if hc is not None:
        V[4] = hc
    return V


def plain_swe(n: int = 512, dx: float = 20.0, steps: int = 2000, ny: Optional[int] = None,
              **scheme_kw) -> Scenario:
    """Plain shallow water, no sediment: a runup beach (bed z linear from
    -50 m at x=0 to +10 m at x=max, uniform in y), still water plus a
    Gaussian hump (2 m, sigma 500 m, centred at (3 km, mid-line)), walls,
    friction coefficient 0.025, CFL 0.9.  The default size is 512^2 at 20 m."""
    ny = n if ny is None else ny
    p = base_params(n, ny, dx, dx, friction=0.025)
    x, y = cell_centres(n, ny, dx, dx)
    Lx, Ly = n * dx, ny * dx
    z = -50.0 + 60.0 * x / Lx
    H = np.broadcast_to(-z[None, :], (ny, n)).copy()
    X, Y = np.meshgrid(x, y)
    eta0 = 2.0 * np.exp(-((X - 3000.0 * Lx / 10240.0) ** 2 + (Y - 0.5 * Ly) ** 2) / (2.0 * (500.0 * Lx / 10240.0) ** 2))
    h = np.maximum(0.0, eta0 + H)
    V = pack_state(h, np.zeros_like(h), np.zeros_like(h), H)
    s = default_scheme(bedload=0, suspension=0)
    s.update(scheme_kw)
    return Scenario("plain_swe_%dx%d" % (n, ny), p, s, V, np.ones((ny, n), np.int32), steps=steps)


def shelf_bed(npx: int, npy: int, dx: float, seed: int) -> np.ndarray:
    """H (depth, positive down) of the sediment workloads: linear shelf
    z = -40 + 48 x/Lx plus a k^-3 fractal perturbation of RMS 2 m."""
    x, _ = cell_centres(npx, npy, dx, dx)
    Lx = npx * dx
    z = -40.0 + 48.0 * x[None, :] / Lx + fractal_field(npx, npy, seed)
    return -z


def sediment_shelf(n: int = 2048, dx: float = 10.0, steps: int = 5000, seed: int = 2,
                   ny: Optional[int] = None, inflow: bool = True, t_series: float = 4000.0,
                   **scheme_kw) -> Scenario:
    """Sediment morphodynamics over a fractal shelf: erodible sand
    (d50 0.2 mm, rho_s 2650 kg/m3 so R = 2.65, porosity 0.4), non-erodible
    mask of 40 seeded rectangles covering 5 %, still water, and an N-wave of
    3 m amplitude and 600 s period entering at x = 0 (other sides walls)."""
    ny = n if ny is None else ny
    p = base_params(n, ny, dx, dx, friction=0.025, R=2.65, DS=2e-4, porosity=0.4)
    H = shelf_bed(n, ny, dx, seed)
    h = np.maximum(0.0, H)
    V = pack_state(h, np.zeros_like(h), np.zeros_like(h), H)
    fix = nonerodible_mask(n, ny, seed + 1000)
    s = default_scheme()
    CC = tCC = None
    if inflow:
        s["bc"] = [BC_INFLOW, BC_WALL, BC_WALL, BC_WALL]
        tCC = np.arange(0.0, t_series + 1e-9, 10.0)
        CC = inflow_series(n, ny, H, n_wave(tCC, 3.0, 600.0), tCC)
    s.update(scheme_kw)
    return Scenario("sediment_shelf_%dx%d" % (n, ny), p, s, V, fix, CC, tCC, steps=steps)


def roofline_domain(n: int = 8192, dx: float = 5.0, steps: int = 1000, seed: int = 3,
                    t_series: float = 400.0, **kw) -> Scenario:
    """The single large domain: 8192^2 cells at 5 m with the sediment-shelf
    bed, sediment parameters and N-wave inflow (the series covers 400 s,
    longer than 1000 steps at dt ~ 0.1 s)."""
    sc = sediment_shelf(n=n, dx=dx, steps=steps, seed=seed, t_series=t_series, **kw)
    sc.name = "roofline_%dx%d" % (n, n)
    return sc


def gaussian_source(n: int, dx: float, seed: int, steps: int = 3000, **scheme_kw) -> Scenario:
    """One ensemble member: fractal shelf bed (sediment-shelf generator),
    Gaussian eta source with amplitude U(0.5, 4) m, sigma U(200, 1500) m,
    centre uniform in the offshore (deep, small-x) third; walls."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.5, 4.0)
    sig = rng.uniform(200.0, 1500.0)
    L = n * dx
    cx = rng.uniform(0.0, L / 3.0)
    cy = rng.uniform(0.0, L)
    p = base_params(n, n, dx, dx, friction=0.025)
    H = shelf_bed(n, n, dx, seed)
    x, y = cell_centres(n, n, dx, dx)
    X, Y = np.meshgrid(x, y)
    eta0 = A * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * sig * sig))
    h = np.maximum(0.0, eta0 + H)
    V = pack_state(h, np.zeros_like(h), np.zeros_like(h), H)
    s = default_scheme()
    s.update(scheme_kw)
    fix = nonerodible_mask(n, n, seed + 1000)
    return Scenario("gauss_%d_s%d" % (n, seed), p, s, V, fix, steps=steps)


def lake_at_rest(n: int = 64, dx: float = 10.0, seed: int = 5, **scheme_kw) -> Scenario:
    """Still water (eta = 0, u = 0) over the fractal shelf with emerged land:
    a well-balanced scheme keeps it at rest."""
    sc = sediment_shelf(n=n, dx=dx, steps=0, seed=seed, inflow=False, **scheme_kw)
    sc.name = "lake_at_rest_%d" % n
    return sc
