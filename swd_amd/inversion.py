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
           smooth_z: int = 1, smooth_x: int = 2, min_offset: int = 3, log=None) -> dict:
    """Run SWD iterations from models["vs"] towards picks observed on
    models["vs_true"] (synthetic data: the observed picks come from a forward
    of the true model, SWD.m:138-162).  vsmin / vsmax are the true model's
    extremes (SWD.m:76).  Returns the final model, the residual history and
    the per-iteration records."""
    vsmin, vsmax = float(np.min(models["vs_true"])), float(np.max(models["vs_true"]))
    prob = dict(problem)
    vs = np.asfortranarray(models["vs"], dtype=np.float64)
    state = {"f1": 0.5}
    residual, records = [], []

    def engine_for(q, vs_now):
        e = Engine(q)
        e.set_wavelet(models["s"])
        e.set_model(models["vp_true"], models["vs_true"])
        e.observe()                       # observed picks for these windows
        e.set_model(models["vp"], vs_now)
        return e

    e = engine_for(prob, vs)
    try:
        for k in range(1, iterations + 1):
            r = e.iterate(vsmin, vsmax, pur, smooth_z, smooth_x, state)
            state = {"f1": r["f1"], "f2": r["f2"], "res2": r["res2"]}
            residual.append(r["residual"])
            records.append(r)
            if log:
                log(f"k={k} residual={r['residual']:.6g} res0={r['res0']:.6g} gama={r['gama']:.4g} "
                    f"forwards={r['forwards']}")
            if k > 1 and (residual[-1] - r["res0"]) / residual[-1] < err:  # SWD.m:326-329
                prob["offset"] = prob["offset"] - 3
                state["f1"] = 0.5
                if prob["offset"] == min_offset:                          # SWD.m:331-333
                    break
                vs = e.model()
                e.close()
                e = engine_for(prob, vs)
        vs = e.model()
    finally:
        e.close()
    return {"vs": vs, "residual": np.array(residual), "records": records, "offset": prob["offset"]}
