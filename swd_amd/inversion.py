"""The SWD inversion loop (SWD.m:146-335) on the device-resident engine.

Each iteration is one ``swd_engine_iterate`` call: gradient, smooth2a, line
search over forward-only batches and the model update all run on the GPU; the
host only carries the script's scalars (f1, f2, res2, the residual history)
and applies the offset schedule of SWD.m:326-333, which changes the Radon
windows and so needs a new engine and new observed picks.
"""
from __future__ import annotations

import numpy as np

from ._lib import check, load, ptr
from .engine import Engine


def smooth2a(A, nr: int, nc: int) -> np.ndarray:
    """core/smooth2a.m on the GPU: (2nr+1) x (2nc+1) box mean, truncated edges, NaNs ignored."""
    a = np.asfortranarray(A, dtype=np.float64)
    out = np.zeros_like(a, order="F")
    check(load().swd_smooth2a(ptr(a), a.shape[0], a.shape[1], int(nr), int(nc), ptr(out)))
    return out


def invert(problem: dict, models: dict, iterations: int, pur: float = -0.05, err: float = 0.01,
           smooth_z: int = 1, smooth_x: int = 2, min_offset: int = 3, log=None, *, offset_step: int = 3,
           variant="swd", observed=None, vs_bounds=None, autotune: bool = False, graph: bool = True) -> dict:
    """Run SWD iterations from models["vs"] (with models["vp"]).

    Observed picks: by default from a forward of models["vs_true"] (synthetic
    data, SWD.m:138-162); ``observed(offset, npair)`` -> (cr0_r, cr0_l) gives
    them instead, e.g. :func:`swd_amd.field.manual_picks` of a manual cube
    (outcrop.m:148-157).  ``vs_bounds`` = (vsmin, vsmax) of the clipping
    (default: the true model's extremes, SWD.m:76).  After an iteration whose
    relative misfit drop is below ``err`` the offset shrinks by
    ``offset_step`` (SWD.m:326-329: 3; outcrop.m:342-345: 5) and the loop
    stops once it reaches ``min_offset`` or less (SWD.m:331-333; the outcrop
    script's WD gradient below offset 11 is not part of this engine).
    ``variant`` is a field-script variant of :mod:`swd_amd.field`.
    ``autotune`` times the FD launch shapes for each engine (offset window)
    before its first iteration; ``graph`` replays the FD phases as captured
    HIP graphs.  Neither changes the results.  Returns the final model, the
    residual history and the per-iteration records."""
    if vs_bounds is None:
        vs_bounds = (float(np.min(models["vs_true"])), float(np.max(models["vs_true"])))
    vsmin, vsmax = (float(v) for v in vs_bounds)
    prob = dict(problem)
    vs = np.asfortranarray(models["vs"], dtype=np.float64)
    vp = np.asfortranarray(models["vp"], dtype=np.float64)
    state = {"f1": 0.5}
    residual, records = [], []

    def engine_for(q, vp_now, vs_now):
        e = Engine(q)
        e.set_variant(variant)
        e.set_wavelet(models["s"])
        if observed is None:
            e.set_model(models["vp_true"], models["vs_true"])
            e.observe()                       # observed picks for these windows
        else:
            e.set_observed_picks(*observed(q["offset"], e.npair))
        e.set_model(vp_now, vs_now)
        if autotune:
            e.autotune()
        e.set_graph(graph)
        return e

    e = engine_for(prob, vp, vs)
    ratio = e.variant.vp_ratio
    try:
        for k in range(1, iterations + 1):
            r = e.iterate(vsmin, vsmax, pur, smooth_z, smooth_x, state)
            state = {"f1": r["f1"], "f2": r["f2"], "res2": r["res2"]}
            residual.append(r["residual"])
            records.append(r)
            if log:
                log(f"k={k} offset={prob['offset']} residual={r['residual']:.6g} res0={r['res0']:.6g} "
                    f"gama={r['gama']:.4g} forwards={r['forwards']}")
            if k > 1 and (residual[-1] - r["res0"]) / residual[-1] < err:  # SWD.m:326-329
                prob["offset"] = prob["offset"] - offset_step
                state["f1"] = 0.5
                if prob["offset"] <= min_offset:                          # SWD.m:331-333
                    break
                vs = e.model()
                if ratio > 0:
                    vp = vs * ratio
                e.close()
                e = engine_for(prob, vp, vs)
        vs = e.model()
    finally:
        e.close()
    return {"vs": vs, "residual": np.array(residual), "records": records, "offset": prob["offset"]}
