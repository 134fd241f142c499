"""Field-data inputs of SWD: the variants the field scripts run, their
manual-pick cubes, and the outcrop survey geometry.

The two field examples (fieldexamples/outcrop/outcrop.m and
fieldexamples/QadimahFault/QademahSWD3.m) run the SWD.m loop with their own
copies of core/: predicted gathers normalised trace by trace (norm_trace),
picks started at the spectrum's strongest column, a misfit that leaves out
the band edges, adjoint sources scaled to unit peak, and vp tied to vs at
every model change.  :data:`VARIANTS` names these settings for
``Engine.set_variant`` / ``DispersionBatch.set_variant``
(swd_engine_set_variant in include/swd_abi.h).  The observed picks are
manual, read from MATLAB cubes (:func:`load_manual_cubes`,
:func:`manual_picks`).
"""
from __future__ import annotations

import ctypes

import numpy as np

from .fd import ricker


class Variant(ctypes.Structure):
    _fields_ = [("norm_trace", ctypes.c_int), ("auto_start", ctypes.c_int), ("res_first", ctypes.c_int),
                ("res_last", ctypes.c_int), ("adj_peak_norm", ctypes.c_int), ("vp_ratio", ctypes.c_double),
                ("fk_filter", ctypes.c_int)]


VARIANTS = {
    # SWD.m with core/: nothing changed
    "swd": {},
    # fieldexamples/outcrop: norm_trace (outcrop.m:164), LHDispPick start at max(max(mlr)) and
    # misfit over cr_pre(15:end-15) (outcrop/core/LHDispPick.m:3-21), ADWDgrad_1.m:79 peak
    # normalisation, vp = vs1*1.732 (outcrop.m:217)
    "outcrop": dict(norm_trace=1, auto_start=1, res_first=15, res_last=15, adj_peak_norm=1, vp_ratio=1.732),
    # fieldexamples/QadimahFault: as the outcrop with the misfit over cr_pre(50:end-30)
    # (QadimahFault/core/LHDispPick.m:25); FK = 0 there (QademahSWD3.m:119), fk_filter=1 is its FK = 1
    "qadimah": dict(norm_trace=1, auto_start=1, res_first=50, res_last=30, adj_peak_norm=1, vp_ratio=1.732),
}


def variant_struct(variant="swd") -> Variant:
    """The swd_variant struct of a variant name or dict."""
    v = VARIANTS[variant] if isinstance(variant, str) else dict(variant)
    unknown = set(v) - {f for f, _ in Variant._fields_}
    if unknown:
        raise ValueError(f"unknown variant fields {sorted(unknown)}")
    return Variant(**{k: (float(x) if k == "vp_ratio" else int(x)) for k, x in v.items()})


# --------------------------------------------------------------------------
# Manual picks
# --------------------------------------------------------------------------

def load_manual_cubes(path: str, names=("crr_all_manual", "crl_all_manual")) -> tuple:
    """Read pick cubes (nfreq x nshots x noffsets, MATLAB double) from a
    MATLAB v5 .mat file, e.g. the outcrop's outcropDisp.mat
    (crr_all_manual, crl_all_manual; outcrop.m:139-140) or the Qadimah
    survey's manualDispfile3of3.mat (crObsManualSR, crObsManualSL;
    QademahSWD3.m:73).  scipy's reader parses arrays only."""
    import scipy.io
    d = scipy.io.loadmat(path, variable_names=list(names))
    missing = [n for n in names if n not in d]
    if missing:
        raise KeyError(f"{path} has no variables {missing}")
    return tuple(np.asarray(d[n], dtype=np.float64) for n in names)


def manual_picks(cube, fmin: float, fmax: float, df: float, offset: int, offset_step: int, smooth=None):
    """Observed picks of one offset from a manual cube whose rows start at
    1 Hz: cube(fstart:fband, :, offset/offset_step) with
    fband = round((fmax-1)/df+1), fstart = round((fmin-1)/df+1)
    (outcrop.m:148-157 with offset_step 5; QademahSWD3.m:149-153 with 3,
    which then smooths by smooth2a(., 2, 1): pass smooth=(2, 1)).
    Returns an npair x nshots array."""
    def mround(x):
        return int(np.floor(x + 0.5)) if x >= 0 else -int(np.floor(-x + 0.5))
    fband = mround((fmax - 1) / df + 1)
    fstart = mround((fmin - 1) / df + 1)
    k = offset / offset_step
    if k != int(k) or not 1 <= k <= cube.shape[2]:
        raise ValueError(f"offset {offset} does not index the cube (step {offset_step}, {cube.shape[2]} offsets)")
    if fstart < 1 or fband > cube.shape[0]:
        raise ValueError(f"rows {fstart}:{fband} outside the cube's {cube.shape[0]} frequencies")
    cr = np.array(cube[fstart - 1:fband, :, int(k) - 1], dtype=np.float64, order="F")
    if smooth is not None:
        from .inversion import smooth2a
        cr = smooth2a(cr, *smooth)
    return cr


# --------------------------------------------------------------------------
# The outcrop survey (fieldexamples/outcrop/outcrop.m:78-120)
# --------------------------------------------------------------------------

OUTCROP_VS_BOUNDS = (90.0, 671 * 1.6)      # vsmin, vsmax (outcrop.m:80)


def outcrop_fmin(offset: int) -> float:
    """fmin_all = [25, 20 x 9]; fmin = round(fmin_all(offset/5), 1) (outcrop.m:146-147)."""
    fmin_all = [25.0] + [20.0] * 9
    return round(fmin_all[offset // 5 - 1], 1)


def outcrop(offset: int = 50, nshots: int | None = None, nt: int | None = None):
    """The outcrop survey: 40 x 240 cells at 0.25 m, 40-cell borders, 60
    shots and 60 receivers every 4 cells (1 m traces, M = 1), 30 Hz Ricker,
    dt = dx/vpmax/2 with vpmax = 1.732 * 671 * 1.6, nt from outcrop.m:92,
    snapshots every 5 steps; Radon 100-800 m/s at df = 0.1 Hz up to 85 Hz,
    windows of ``offset`` traces, w = 4, SoftArgNorm 1e4.  Initial model
    vs = linspace(150, 700) with depth, vp = 1.732 vs (outcrop.m:84-89).
    ``nshots`` keeps the first shots and ``nt`` overrides the trace length
    (shorter test runs); ns of the branch guards stays 60."""
    vsmax = 671 * 1.6
    vpmax = vsmax * 1.732
    dx = 0.25
    dt = dx / vpmax * 0.5
    nz, nx, nbc = 40, 240, 40
    vs = np.tile(np.linspace(150, 700, nz)[:, None], (1, nx))
    vp = vs * 1.732
    nt_full = int(np.floor((nx * dx / vs.min() / dt + 1000) / 10) * 10 + 500)
    nt = nt_full if nt is None else int(nt)
    sx = np.arange(1, nx + 1, 4)
    ns = sx.size
    nshots = ns if nshots is None else int(nshots)
    prob = dict(nz=nz, nx=nx, nbc=nbc, fd_order=22, nt=nt, in_wf=5, dx=dx, dt=dt, sx=sx[:nshots],
                sz=np.ones(nshots, dtype=int), ng=60, gx0=1, dg=4, gz=1, df=0.1, fmin=outcrop_fmin(offset),
                fmax=85.0, vmin=100.0, vmax=800.0, ds=4 * dx, offset=offset, m=0, M=1, ini=200, w=4.0, beta=1e4,
                ns=ns, is_index=np.arange(1, nshots + 1))
    return prob, {"vs": vs, "vp": vp, "s": ricker(30.0, dt, nt)}


# --------------------------------------------------------------------------
# The Qadimah-fault survey (fieldexamples/QadimahFault/QademahSWD3.m:75-136)
# --------------------------------------------------------------------------

QADIMAH_VS_BOUNDS = (100.0, 1200.0)        # vsmin, vsmax (QademahSWD3.m:76)


def qadimah(offset: int = 39, nshots: int | None = None, nt: int | None = None):
    """The Qadimah-fault survey: 120 x 2400 cells at 0.25 m, 40-cell borders,
    120 shots every 20 cells and 240 receivers every 10 cells (2.5 m traces,
    M = 2), 30 Hz Ricker, dt = dx/vpmax/2 with vpmax = 1.732 * 1200, nt 8500,
    snapshots every 10 steps (dt_wf = 10 dt); Radon 100-1000 m/s at df = 0.1 Hz
    over 10-60 Hz, windows of ``offset`` traces, w = 3, SoftArgNorm 1e4.
    Initial model vs = linspace(200, 1100) with depth, vp = 1.732 vs
    (QademahSWD3.m:75-123).  The pick starts at the strongest column (the
    "qadimah" variant), so ``ini`` is not used.  ``nshots`` keeps the first
    shots and ``nt`` overrides the trace length (shorter test runs); ns of the
    branch guards stays 120."""
    vsmax = 1200.0
    vpmax = vsmax * 1.732
    dx = 0.25
    dt = dx / vpmax * 0.5
    nz, nx, nbc = 120, 2400, 40
    vs = np.tile(np.linspace(200, 1100, nz)[:, None], (1, nx))
    vp = vs * 1.732
    nt = 8500 if nt is None else int(nt)
    sx = np.arange(1, nx + 1, 20)
    ns = sx.size
    nshots = ns if nshots is None else int(nshots)
    prob = dict(nz=nz, nx=nx, nbc=nbc, fd_order=22, nt=nt, in_wf=10, dx=dx, dt=dt, sx=sx[:nshots],
                sz=np.ones(nshots, dtype=int), ng=240, gx0=1, dg=10, gz=1, df=0.1, fmin=10.0,
                fmax=60.0, vmin=100.0, vmax=1000.0, ds=10 * dx, offset=offset, m=0, M=2, ini=1, w=3.0, beta=1e4,
                ns=ns, is_index=np.arange(1, nshots + 1))
    return prob, {"vs": vs, "vp": vp, "s": ricker(30.0, dt, nt)}
