"""Host mirror of SWD's dispersion functions over libswd (FP32 on the GPU).

``RTr``, ``RTl``, ``RTrAD``, ``RTlAD`` (core/RTr.m:2, core/RTl.m:2,
core/RTrAD.m:2, core/RTlAD.m:2), ``LHDispPick`` (core/LHDispPick.m:2) and
``ADWDgrad_1`` (core/ADWDgrad_1.m:1) keep the reference argument lists and
return values.  The AD variants return the ``saveForBackward`` struct as a
dict with the reference fields (mm, ll0, lf, nf, ccn, mlr/mll,
uxtposr/uxtposl) plus the scalars the device path rebuilds ll0 from
(np, vmin, df, ds, m).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from ._lib import check, fortran, load, ptr


def _radon(side, seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M, want_ad):
    seismo = np.asarray(seismo)
    single = seismo.dtype == np.float32
    seismo = fortran(seismo, np.float32 if single else np.float64)
    nt, ng = seismo.shape
    lf = int(math.floor(fmin / df + 0.5)) + 1
    nf = int(math.floor(fmax / df + 0.5)) + 1
    npair = nf - lf + 1
    ml = np.zeros((np_, npair), dtype=np.float32, order="F")
    mlr = np.zeros((np_, npair), dtype=np.float32, order="F") if want_ad else None
    mm = np.zeros((2 * np_, npair), dtype=np.float32, order="F") if want_ad else None
    L = ctypes.c_int(0)
    info = np.zeros(4, dtype=np.int32)
    fn = load().swd_radon_f32 if single else load().swd_radon
    check(fn(side, ptr(seismo), nt, ng, int(is_), df, dt, int(np_), float(vmin), float(vmax), float(fmin),
             float(fmax), float(ds), int(offset), int(m), int(M), ptr(ml), ctypes.byref(L), ptr(mlr), ptr(mm),
             ptr(info)))
    if not want_ad:
        return ml, L.value
    lf, nf, ccn, pos0 = (int(v) for v in info)
    mmc = np.zeros((np_, nf), dtype=np.complex64, order="F")
    mmc[:, lf - 1:nf] = mm[0::2] + 1j * mm[1::2]
    Lv = L.value
    x = (m + np.arange(Lv)) * ds
    pp = 1.0 / np.arange(vmin, vmax + 1, dtype=np.float64)
    sfb = {"mm": mmc, "ll0": 1j * 2 * np.pi * df * np.outer(pp, x), "lf": lf, "nf": nf, "ccn": ccn,
           "np": int(np_), "vmin": float(vmin), "df": float(df), "ds": float(ds), "m": int(m), "pos0": pos0}
    if side == 0:
        sfb["mlr"] = mlr
        sfb["uxtposr"] = np.arange(pos0, pos0 + Lv)
    else:
        sfb["mll"] = mlr
        sfb["uxtposl"] = np.arange(pos0 - Lv + 1, pos0 + 1)
    return ml, Lv, sfb


def RTr(seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M):
    """[ml, dataLen] = RTr(...) (right window)."""
    return _radon(0, seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M, False)


def RTl(seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M):
    """[ml, dataLen] = RTl(...) (left window, traces flipped)."""
    return _radon(1, seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M, False)


def RTrAD(seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M):
    """[ml, dataLen, saveForBackward] = RTrAD(...)."""
    return _radon(0, seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M, True)


def RTlAD(seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M):
    """[ml, dataLen, saveForBackward] = RTlAD(...)."""
    return _radon(1, seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M, True)


def LHDispPick(ml, vmin, cr_0, method, ini):
    """[res, cr_pre] = LHDispPick(ml, vmin, cr_0, method, ini); FP64 spectra
    are picked in FP64, others in FP32."""
    f64 = np.asarray(ml).dtype == np.float64
    ml = fortran(ml, np.float64 if f64 else np.float32)
    np_, npair = ml.shape
    cr_0 = np.ascontiguousarray(np.ravel(cr_0), dtype=np.float64)
    if cr_0.size != npair:
        raise ValueError("cr_0 must have one entry per frequency column")
    cr = np.zeros(npair)
    res = ctypes.c_double(0)
    fn = load().swd_lhdisppick_f64 if f64 else load().swd_lhdisppick
    check(fn(ptr(ml), np_, npair, float(vmin), ptr(cr_0), int(method), int(ini), ctypes.byref(res), ptr(cr)))
    return res.value, cr


def _side(sfb, key):
    if not isinstance(sfb, dict):
        return None, None, 0, 0
    lf, nf = sfb["lf"], sfb["nf"]
    mm = np.asarray(sfb["mm"])[:, lf - 1:nf]
    mmi = np.zeros((2 * mm.shape[0], mm.shape[1]), dtype=np.float32, order="F")
    mmi[0::2] = mm.real
    mmi[1::2] = mm.imag
    pos = sfb["uxtposr"] if key == "mlr" else sfb["uxtposl"]
    pos0 = int(pos[0]) if key == "mlr" else int(pos[-1])
    return mmi, fortran(sfb[key], np.float32), pos0, len(pos)


def ADWDgrad_1(nt, ng, ns, npair, is_, w, m, M, SoftArgNorm, grad_outputr, grad_outputl, space_M,
               saveForBackwardr, saveForBackwardl):
    """a_data_res = ADWDgrad_1(...) (nt x ng, float32)."""
    mmr, mlr, posr, Lr = _side(saveForBackwardr, "mlr")
    mml, mll, posl, Ll = _side(saveForBackwardl, "mll")
    ref = saveForBackwardr if isinstance(saveForBackwardr, dict) else saveForBackwardl
    if not isinstance(ref, dict):
        return np.zeros((nt, ng), dtype=np.float32, order="F")
    gor = np.ascontiguousarray(np.ravel(grad_outputr), dtype=np.float64)
    gol = np.ascontiguousarray(np.ravel(grad_outputl), dtype=np.float64)
    out = np.zeros((nt, ng), dtype=np.float32, order="F")
    check(load().swd_adwdgrad1(int(nt), int(ng), int(ns), int(npair), int(is_), float(w), int(m), int(M),
                               float(SoftArgNorm), ptr(gor), ptr(gol), ptr(mmr), ptr(mlr), posr, Lr, ptr(mml),
                               ptr(mll), posl, Ll, int(ref["np"]), float(ref["vmin"]), float(ref["df"]),
                               float(ref["ds"]), int(ref["lf"]), int(ref["ccn"]), ptr(out)))
    return out


def set_radon_precision(mode: int) -> int:
    """Radon DFT / inverse-DFT GEMM arithmetic: 0 = FP32 MFMA, 1 = BF16 split
    precision (bf16 hi + lo, three bf16 MFMA products).  Returns the old mode."""
    rc = load().swd_set_radon_precision(int(mode))
    if rc < 0:
        check(rc)
    return rc


class DispersionBatch:
    """The dispersion path alone on caller-provided gathers, batched over
    every (shot, side) column (swd_dispersion_* in include/swd_abi.h): RTr /
    RTl + LHDispPick, then RTrAD / RTlAD + ADWDgrad_1, as SWD.m:173-190 runs
    them per shot."""

    def __init__(self, problem: dict):
        import ctypes
        from .engine import make_problem
        self.problem = dict(problem)
        p, self._keep = make_problem(self.problem)
        h = ctypes.c_void_p()
        check(load().swd_dispersion_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        v = np.zeros(8, dtype=np.int64)
        check(load().swd_dispersion_info(self.h, ptr(v), 8))
        self.npair, self.np, self.lf, self.nf, self.ccn, self.lmax, self.sum_l, self.ncol = (int(x) for x in v)
        self.nshots = int(np.size(self.problem["sx"]))

    def set_gathers(self, gathers):
        """gathers: (nshots, nt, ng) float32, g fastest."""
        g = np.ascontiguousarray(gathers, dtype=np.float32)
        q = self.problem
        if g.shape != (self.nshots, q["nt"], q["ng"]):
            raise ValueError(f"gathers must be {(self.nshots, q['nt'], q['ng'])}, got {g.shape}")
        check(load().swd_dispersion_set_gathers(self.h, ptr(g)))

    def set_variant(self, variant="swd"):
        """Field-script variant (swd_dispersion_set_variant; vp_ratio ignored)."""
        import ctypes
        from .field import variant_struct
        v = variant_struct(variant)
        check(load().swd_dispersion_set_variant(self.h, ctypes.byref(v)))

    def set_observed_picks(self, cr0_r, cr0_l):
        check(load().swd_dispersion_set_observed_picks(self.h, ptr(fortran(cr0_r, np.float64)),
                                                        ptr(fortran(cr0_l, np.float64))))

    def pick(self):
        """Picks and misfits of the resident gathers: dict cr_r, cr_l
        (npair x nshots), res_r, res_l (nshots), ms (device time)."""
        import ctypes
        cr = np.zeros((self.npair, self.nshots), order="F")
        cl = np.zeros((self.npair, self.nshots), order="F")
        rr, rl = np.zeros(self.nshots), np.zeros(self.nshots)
        ms = ctypes.c_double(0)
        check(load().swd_dispersion_pick(self.h, ptr(cr), ptr(cl), ptr(rr), ptr(rl), ctypes.byref(ms)))
        return {"cr_r": cr, "cr_l": cl, "res_r": rr, "res_l": rl, "ms": ms.value}

    def adjoint(self, beta: float = 0.0, want=True):
        """Adjoint sources (nshots, nt, ng) of the last picks; ms = device time."""
        import ctypes
        q = self.problem
        out = np.zeros((self.nshots, q["nt"], q["ng"]), dtype=np.float32) if want else None
        ms = ctypes.c_double(0)
        check(load().swd_dispersion_adjoint(self.h, float(beta), ptr(out) if want else None, ctypes.byref(ms)))
        return out, ms.value

    def close(self):
        if getattr(self, "h", None):
            load().swd_dispersion_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
