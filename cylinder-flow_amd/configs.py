"""The five benchmark/validation workloads as solver configurations.

Lattice units throughout.  nu = u*D/Re and tau0 = 3*nu (the solver relaxes with
tau = tau0 + 0.5, as the reference does: 3D MPI+KOKKOS+MPIIO/System.cpp:18-19,
lbm.cpp:269).  Every workload uses the plain boundary mode: half-way bounce-back on the
body and the side walls, Zou-He velocity inlet (regularised at the inlet node),
zero-gradient outlet, and an initial state
p = 0, u = (u_in + 1e-3 xi, 0, 0) with xi uniform in [-1, 1) from the seeded counter RNG.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import lbm as L

SEED = 5374857  # the reference's Random_XorShift64_Pool seed (3D MPI+KOKKOS+MPIIO/lbm.cpp:225)


@dataclass
class Workload:
    name: str
    cfg: L.LbmConfig
    steps: int
    D: float                      # body diameter
    area: float                   # reference area for Cd (D in 2D, pi D^2/4 sphere, D*nz cylinder)
    u0: float
    window: tuple = (0, 0)        # adjoint objective window (1-based steps)
    checkpoint: int = 0
    extra: dict = field(default_factory=dict)

    def cd_scale(self) -> float:
        return 2.0 / (self.u0 * self.u0 * self.area)


def cylinder2d_400x100(precision=L.FP64, **kw) -> Workload:
    """D2Q9 400x100, cylinder D=20 at (80,50), Re=100, u=0.05, 20 000 steps, forces every 10."""
    u, D, Re = 0.05, 20.0, 100.0
    c = L.make_config(9, 400, 100, tau0=3 * u * D / Re, u0=u, precision=precision, body=L.BODY_SPHERE,
                      center=(80.0, 50.0, 0.0), r2=(D / 2) ** 2, seed=SEED, perturb=1e-3,
                      **{"force_interval": 10, **kw})
    return Workload("d2q9_cylinder_400x100_re100", c, 20000, D, D, u)


def sphere3d_256x128x128(precision=L.FP32, **kw) -> Workload:
    """D3Q27 256x128x128, sphere D=24 at (64,64,64), Re=200, u=0.05, 10 000 steps."""
    u, D, Re = 0.05, 24.0, 200.0
    c = L.make_config(27, 256, 128, 128, tau0=3 * u * D / Re, u0=u, precision=precision, body=L.BODY_SPHERE,
                      center=(64.0, 64.0, 64.0), r2=(D / 2) ** 2, seed=SEED, perturb=1e-3,
                      **{"force_interval": 10, **kw})
    return Workload("d3q27_sphere_256x128x128_re200", c, 10000, D, math.pi * D * D / 4, u)


def zcylinder3d_512x256x256(precision=L.FP16S, **kw) -> Workload:
    """D3Q27 512x256x256, z-spanning cylinder D=40 at (128,128), Re=300, u=0.05, 20 000 steps."""
    u, D, Re = 0.05, 40.0, 300.0
    c = L.make_config(27, 512, 256, 256, tau0=3 * u * D / Re, u0=u, precision=precision, body=L.BODY_ZCYL,
                      center=(128.0, 128.0, 0.0), r2=(D / 2) ** 2, seed=SEED, perturb=1e-3,
                      **{"force_interval": 10, **kw})
    return Workload("d3q27_zcylinder_512x256x256_re300", c, 20000, D, D * 256, u)


def smoothed_disk(nx, ny, cx, cy, R, width=1.0, noise=0.01, seed=SEED):
    """Solid fraction: 0.5*(1 - tanh((r - R)/width)) plus seeded N(0, noise^2), clamped to [0,1]."""
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    r = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2)
    s = 0.5 * (1.0 - np.tanh((r - R) / width))
    s = s + noise * np.random.default_rng(seed).standard_normal(s.shape)
    return np.clip(s, 0.0, 1.0)[None]


def smoothed_ball(nx, ny, nz, c, R, width=1.0, noise=0.01, seed=SEED):
    zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx].astype(np.float64)
    r = np.sqrt((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2)
    s = 0.5 * (1.0 - np.tanh((r - R) / width))
    s = s + noise * np.random.default_rng(seed).standard_normal(s.shape)
    return np.clip(s, 0.0, 1.0)


def adjoint2d_400x100(**kw) -> Workload:
    """D2Q9 adjoint on the 400x100 cylinder geometry: solid fraction = smoothed disk + 0.01 N(0,1); J = mean Cd
    over steps 4000-8000; checkpoint every 500; FP64."""
    u, D, Re = 0.05, 20.0, 100.0
    c = L.make_config(9, 400, 100, tau0=3 * u * D / Re, u0=u, precision=L.FP64, body=L.BODY_NONE,
                      seed=SEED, perturb=1e-3, **kw)
    w = Workload("d2q9_adjoint_400x100_re100", c, 8000, D, D, u, window=(4000, 8000), checkpoint=500)
    w.extra["sfrac"] = smoothed_disk(400, 100, 80.0, 50.0, D / 2)
    return w


def legendre_coefficients(seed=SEED, n=8, amplitude=0.05):
    """8 seeded coefficients with sum |a_k| <= amplitude and a_0 > 0 (mean flow forward)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, n)
    a[0] = abs(a[0]) + 1.0
    return amplitude * a / np.abs(a).sum()


def adjoint3d_192x96x96(precision=L.FP32, Re=100.0, **kw) -> Workload:
    """D3Q27 adjoint 192x96x96, sphere D=20 as a solid-fraction field, parabola x Legendre inlet
    (8 seeded coefficients, sum |a_k| <= 0.05); J = mean Cd over steps 2000-6000; checkpoint every
    250; FP32 with FP64 reductions.  Re is not fixed by the workload description: Re = 100 on
    D = 20 with u = 0.05 gives tau = 0.53."""
    u, D = 0.05, 20.0
    coef = tuple(legendre_coefficients())
    c = L.make_config(27, 192, 96, 96, tau0=3 * u * D / Re, u0=u, precision=precision, body=L.BODY_NONE,
                      inlet=L.INLET_LEGENDRE, coef=coef, seed=SEED, perturb=1e-3, **kw)
    w = Workload("d3q27_adjoint_192x96x96", c, 6000, D, math.pi * D * D / 4, u, window=(2000, 6000),
                 checkpoint=250)
    w.extra["sfrac"] = smoothed_ball(192, 96, 96, (48.0, 48.0, 48.0), D / 2)
    w.extra["coef"] = np.array(coef)
    return w


ALL = {f.__name__: f for f in (cylinder2d_400x100, sphere3d_256x128x128, zcylinder3d_512x256x256, adjoint2d_400x100,
                                 adjoint3d_192x96x96)}


def strouhal(cl: np.ndarray, dt_steps: float, D: float, u: float, min_periods: int = 3,
             max_cv: float = 0.2, band: tuple | None = None) -> float:
    """Shedding frequency St = D / (T u) from the upward crossings of the lift signal (mean
    removed) through +h after it was below -h, h = std/4 (hysteresis, so noise near zero does
    not add crossings); crossings are interpolated linearly.  NaN unless the signal is
    periodic: at least `min_periods` full periods whose lengths vary by less than `max_cv`
    (coefficient of variation), so a wake that has not started to shed gives NaN.
    band = (St_lo, St_hi) first keeps only those frequencies (FFT band-pass), e.g. to remove a
    channel's transverse acoustic mode (period 2 ny / c_s steps) from the lift."""
    x = np.asarray(cl, dtype=np.float64) - np.mean(cl)
    if band is not None:
        X = np.fft.rfft(x)
        st = np.fft.rfftfreq(len(x), d=dt_steps) * D / u
        X[(st < band[0]) | (st > band[1])] = 0.0
        x = np.fft.irfft(X, n=len(x))
    h = 0.25 * x.std()
    if not h > 0:
        return float("nan")
    t, low = [], False
    for i in range(1, len(x)):
        if x[i] < -h:
            low = True
        elif low and x[i] >= h and x[i - 1] < h:
            t.append(i - 1 + (h - x[i - 1]) / (x[i] - x[i - 1]))
            low = False
    if len(t) < min_periods + 1:
        return float("nan")
    periods = np.diff(t) * dt_steps
    if periods.std() > max_cv * periods.mean():
        return float("nan")
    return D / (periods.mean() * u)
