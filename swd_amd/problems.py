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


def chirp_phase_velocity(f):
    """Phase velocity of the synthetic dispersive wave, c(f) = 250 + 300 exp(-f/20) m/s."""
    return 250.0 + 300.0 * np.exp(-np.asarray(f, dtype=np.float64) / 20.0)


def chirp_gathers(nshots=60, nt=6000, ng=600, dt=8e-5, dx=0.5, sx=None, noise=0.1, seed=20261203,
                  fmin=5.0, fmax=80.0, df=0.5):
    """Dispersion-path gathers [shot][t][g] (float32): Gaussian noise band-
    limited to fmin..fmax plus a dispersive surface wave with phase velocity
    chirp_phase_velocity(f), radiating from each shot's source trace x_s
    (1-based trace sx[s]) with spectrum exp(-i 2 pi f (t0 + |x - x_s| / c(f))).
    The wave is synthesised on a 1/df-second grid (the Radon's own frequency
    spacing) and cut to the nt-sample record, so arrivals beyond the record
    are absent as in a real gather.  noise = noise RMS / signal RMS."""
    rng = np.random.default_rng(seed)
    if sx is None:
        sx = 1 + 10 * np.arange(nshots)
    N = max(nt, int(round(1.0 / df / dt)))
    f = np.fft.rfftfreq(N, dt)
    band = (f >= fmin) & (f <= fmax)
    taper = np.zeros_like(f)
    w = 2.0  # cosine tapers over 2 Hz at both band edges
    taper[band] = 1.0
    lo = (f >= fmin) & (f < fmin + w)
    hi = (f > fmax - w) & (f <= fmax)
    taper[lo] = 0.5 - 0.5 * np.cos(np.pi * (f[lo] - fmin) / w)
    taper[hi] = 0.5 - 0.5 * np.cos(np.pi * (fmax - f[hi]) / w)
    c = np.where(band, chirp_phase_velocity(np.maximum(f, 1e-6)), 1.0)
    # the wave depends on the offset |x - x_s| only: synthesise every offset once
    t0 = 0.02  # onset delay (s)
    offs = np.arange(ng) * dx
    sig_by_off = np.empty((nt, ng))
    for j0 in range(0, ng, 100):  # offset blocks keep the complex spectra small
        o = offs[j0:j0 + 100]
        ph = np.exp(-2j * np.pi * f[:, None] * (o[None, :] / c[:, None] + t0))
        sig_by_off[:, j0:j0 + 100] = np.fft.irfft(taper[:, None] * ph, n=N, axis=0)[:nt]
    # noise: band-limited on the record's own grid
    fr = np.fft.rfftfreq(nt, dt)
    nband = ((fr >= fmin) & (fr <= fmax)).astype(np.float64)
    out = np.empty((nshots, nt, ng), dtype=np.float32)
    g_idx = np.arange(ng)
    for s in range(nshots):
        sig = sig_by_off[:, np.abs(g_idx - (int(sx[s]) - 1))]
        nz = np.fft.irfft(nband[:, None] * np.fft.rfft(rng.standard_normal((nt, ng)), axis=0), n=nt, axis=0)
        sr = np.sqrt(np.mean(sig ** 2))
        nr = np.sqrt(np.mean(nz ** 2))
        out[s] = (sig + noise * sr / max(nr, 1e-30) * nz).astype(np.float32)
    return out
