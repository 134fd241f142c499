"""Synthetic SWD workloads with the shapes of the benchmark configurations.

Each builder returns (problem dict for ``engine.Engine``, models dict).  The
models are synthetic (no field data is read): linear-gradient, checkerboard
and faulted layered Vs, Vp = 1.73 Vs.  Density is fixed to 1 by the
reference wrappers (core/staggerfd_eigen.m:38), so rho does not enter.
"""
from __future__ import annotations

import numpy as np

from .fd import ricker


def checkerboard(nz, nx, cell, base, frac):
    """+-frac checkerboard of cell x cell squares (core/generateCheckerboard.m:2-12 idea)."""
    iz = np.arange(nz)[:, None] // cell
    ix = np.arange(nx)[None, :] // cell
    even = ((iz + ix) % 2) == 0
    return np.where(even, base * (1 + frac), base * (1 - frac)).astype(np.float64)


def production_shots(nshots=60, nt=6000, beta=10.0):
    """Single-model production shots: 100 x 600 cells at 0.5 m, 30-cell borders,
    60 Ricker 25 Hz shots 5 m apart, 600 receivers, dt = 8e-5 s, nt = 6000.
    True model: 300 m/s with a +-20% checkerboard of 10 m cells; current: 300 m/s."""
    nz, nx, nbc, dx, dt = 100, 600, 30, 0.5, 8e-5
    sx = 1 + 10 * np.arange(nshots)
    prob = dict(nz=nz, nx=nx, nbc=nbc, fd_order=22, nt=nt, in_wf=5, dx=dx, dt=dt, sx=sx,
                sz=np.ones(nshots, dtype=int), ng=600, gx0=1, dg=1, gz=1, df=0.5, fmin=5.0, fmax=80.0,
                vmin=100.0, vmax=800.0, ds=0.5, offset=600, m=0, M=10, ini=10, w=3.0, beta=beta, ns=0)
    vs_true = checkerboard(nz, nx, 20, 300.0, 0.2)
    vs_cur = np.full((nz, nx), 300.0)
    return prob, {"vs_true": vs_true, "vp_true": 1.73 * vs_true, "vs": vs_cur, "vp": 1.73 * vs_cur,
                  "s": ricker(25.0, dt, nt)}


def plain_single_shot(nt=3000, beta=10.0):
    """60 x 200 cells at 0.5 m, 20-cell borders, Vs 200 -> 600 m/s with depth,
    one 30 Hz shot at x = 2 m (trace 5), 199 receivers; observed model with a
    +10% Gaussian Vs anomaly (sigma 5 m).  Windows use is = 5, M = 1, w = 0 so
    that the source trace is 5 and both sides pass the branch guards."""
    nz, nx, nbc, dx, dt = 60, 200, 20, 0.5, 1e-4
    vs = np.tile(np.linspace(200, 600, nz)[:, None], (1, nx))
    z, x = np.meshgrid(np.arange(nz) * dx, np.arange(nx) * dx, indexing="ij")
    vs_true = vs * (1 + 0.1 * np.exp(-((z - 15.0) ** 2 + (x - 50.0) ** 2) / (2 * 5.0 ** 2)))
    prob = dict(nz=nz, nx=nx, nbc=nbc, fd_order=22, nt=nt, in_wf=5, dx=dx, dt=dt, sx=np.array([5]),
                sz=np.array([1]), ng=199, gx0=1, dg=1, gz=1, df=0.5, fmin=5.0, fmax=80.0, vmin=100.0,
                vmax=800.0, ds=0.5, offset=199, m=0, M=1, ini=10, w=0.0, beta=beta, ns=5,
                is_index=np.array([5]))
    return prob, {"vs_true": vs_true, "vp_true": 1.73 * vs_true, "vs": vs, "vp": 1.73 * vs,
                  "s": ricker(30.0, dt, nt)}


def faulted_layers(nz, nx, rng, vmin=150.0, vmax=700.0, throw_m=10.0, dx=0.5):
    """Layered Vs increasing with depth, offset across a vertical fault at nx/2."""
    nl = int(rng.integers(4, 8))
    tops = np.sort(rng.uniform(0, nz * dx, nl - 1))
    vel = np.sort(rng.uniform(vmin, vmax, nl))
    z = np.arange(nz)[:, None] * dx
    vs = np.empty((nz, nx))
    for half, shift in ((slice(0, nx // 2), 0.0), (slice(nx // 2, nx), throw_m)):
        zz = np.broadcast_to(z - shift, (nz, nx))[:, half]
        vs[:, half] = vel[np.searchsorted(tops, zz, side="right")]
    return vs


def uneven_models(seed=20261203, shots=16, sizes=(200, 300, 450, 600, 800, 900, 1000, 1200), nt_range=(2000, 8000),
                  f0=20.0):
    """Eight models of different size and nt, each with `shots` shots (dx 0.5 m,
    30-cell borders, dt 1e-4 s, f0 Hz Ricker), faulted layered Vs 150-700 m/s."""
    rng = np.random.default_rng(seed)
    out = []
    for nx in sizes:
        nz = int(rng.integers(40, 151))
        nt = int(rng.integers(nt_range[0] // 5, nt_range[1] // 5 + 1)) * 5
        dx, dt = 0.5, 1e-4
        sx = np.linspace(1, nx, shots).round().astype(int)
        vs = faulted_layers(nz, nx, rng, throw_m=float(rng.uniform(5, 15)))
        vs_cur = np.tile(np.linspace(vs.min(), vs.max(), nz)[:, None], (1, nx))
        ds_tr = 1
        M = max(1, int(round((sx[1] - sx[0]) / ds_tr))) if shots > 1 else 1
        prob = dict(nz=nz, nx=nx, nbc=30, fd_order=22, nt=nt, in_wf=5, dx=dx, dt=dt, sx=sx,
                    sz=np.ones(shots, dtype=int), ng=nx, gx0=1, dg=1, gz=1, df=0.5, fmin=5.0, fmax=80.0,
                    vmin=100.0, vmax=800.0, ds=dx, offset=nx, m=0, M=1, ini=10, w=0.0, beta=10.0, ns=nx,
                    is_index=sx)
        out.append((prob, {"vs_true": vs, "vp_true": 1.73 * vs, "vs": vs_cur, "vp": 1.73 * vs_cur,
                           "s": ricker(f0, dt, nt)}))
    return out
