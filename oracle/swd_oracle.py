"""TEST INFRASTRUCTURE ONLY - the CPU oracle for the SWD hot path.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this module, and only as the checker.  The product (``libswd.so``
and ``swd_amd``) never imports it.

What it restates (file:line in the reference, zhangchang271/SWD):

* FD forward / adjoint time loops and the wrapper preparation: the C library
  ``oracle/_build/libswd_oracle.so`` built from ``swd_oracle.c`` (see the
  citations there); this module wraps it with the MATLAB wrapper signatures of
  ``core/staggerfd_eigen(d).m:2-141`` and ``core/e2drtm_eigen(d).m:8-123``.
* Linear Radon transform ``core/RTr.m:2-33``, ``core/RTl.m:2-33``,
  ``core/RTrAD.m:2-35``, ``core/RTlAD.m:2-36``  -> :func:`rt_side`.
* Dispersion picking ``core/LHDispPick.m:2-13`` and ``core/pickFDC1.m:1-131``
  (with MATLAB ``findpeaks`` / ``interp1`` semantics) -> :func:`pickFDC1`.
* SoftMax adjoint source ``core/ADWDgrad_1.m:1-82`` -> :func:`ADWDgrad_1`.
* Gradient assembly ``SWD.m:193,199-201`` and ``core/smooth2a.m:52-76``.

Everything runs in float64 (MATLAB ``double``) with numpy's FFT standing in
for MATLAB's ``fft``/``ifft`` (same definition, both unnormalised forward and
1/N inverse).  Parity status of each function is listed in the module
``PARITY`` table below; functions pinned only by this restatement say so.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from typing import Dict, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libswd_oracle.so")

# How each oracle function is pinned (checks live in tests/test_oracle_*.py).
PARITY = {
    "oracle_prep": "hand-computed tiny cases (pad/calparam/damp_circle values, free-surface row edits)",
    "oracle_ricker": "closed form at the centre sample, symmetry, length rule",
    "oracle_staggerfd": "two-step closed form after an impulse (order 22) and one-step closed form with "
                        "S41/S42 (order 24: strains and the next velocities), linearity in s, zero source, "
                        "density invariance, receiver-on-source, zz free-surface row, FP32 twin",
    "oracle_e2drtm": "closed forms read back through the imaging condition: the first stresses after one "
                     "adjoint impulse (orders 22 and 24), the second step's coefficient-weighted velocity "
                     "half-step derived cell by cell (order 22, also on the production receiver row), and "
                     "three steps against a scatter form of the velocity half-step (orders 22 and 24, "
                     "xx_b included); illum closed form, linearity in the adjoint source, bilinearity in "
                     "the snapshots",
    "rt_side": "plane-wave peak at c0, single-trace constant |mm|, linearity, window/flip rules",
    "pickFDC1": "hand-built columns (single peak, width-jump carry-over + interpolation, ties)",
    "findpeaks": "hand-built vectors incl. plateaus and endpoints (published MATLAB rules)",
    "ADWDgrad_1": "finite differences of the soft-argmax misfit, slant-stack dot-product test, "
                  "one-sided real(ifft) closed form",
    "smooth2a": "box mean with truncated edges, brute force",
    "norm_trace": "closed form: each column's peak |x| becomes 1 with signs kept, invariance under positive "
                  "column gains, NaN samples skipped by the max, a zero column gives NaN",
    "LHDispPick_field": "hand-built spectra whose largest |mm| sits in a known column (start column and row), "
                        "closed-form windowed misfit, pickFDCValue == pickFDC1 when the given row is the argmax",
    "fk_filter": "brute-force 2-D DFT restatement of the MATLAB loops on tiny inputs (explicit sums, explicit "
                 "shift indices), real output, linearity",
    "ADWDgrad_1 peak_norm": "closed form: peak |a| becomes 1, the result is the unnormalised source over its peak",
    "swd_line_search": "quadratic misfit in the step factor: the parabola step lands on the exact minimum; "
                       "the halving branch, the doubling branch, the fallback to the better trial and the "
                       "clipping each checked on closed-form misfits",
}


def _build_if_needed() -> None:
    if os.path.exists(LIB_PATH):
        return
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_LIB = None


def lib() -> ctypes.CDLL:
    """Load the C oracle (building it with ``make`` when absent)."""
    global _LIB
    if _LIB is None:
        _build_if_needed()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        for suf in ("d", "f"):
            fn = getattr(L, f"oracle_staggerfd_{suf}")
            fn.restype = ctypes.c_int
            fn.argtypes = [P] * 15
            fn = getattr(L, f"oracle_e2drtm_{suf}")
            fn.restype = ctypes.c_int
            fn.argtypes = [P] * 16
            fn = getattr(L, f"oracle_batch_{suf}")
            fn.restype = ctypes.c_int
            fn.argtypes = [P] * 11 + [ctypes.c_int, P, P, P, P, P, ctypes.c_int]
        L.oracle_prep.restype = ctypes.c_int
        L.oracle_prep.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_double] + [P] * 7
        L.oracle_damp_circle.restype = None
        L.oracle_damp_circle.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int, P]
        L.oracle_ricker.restype = ctypes.c_int
        L.oracle_ricker.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, P]
        L.oracle_max_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _f(a, dtype=np.float64) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=dtype))


# --------------------------------------------------------------------------
# model preparation (staggerfd_eigen.m:37-107)
# --------------------------------------------------------------------------

def pad_top_of(fd_order: int) -> int:
    """MATLAB pad_top with a free surface, staggerfd_eigen.m:50-51."""
    return (fd_order - 20) // 2 + 1


def prep(vp, vs, nbc: int, dx: float, dt: float, fd_order: int = 22, den: float = 1.0) -> Dict[str, np.ndarray]:
    vp = _f(vp)
    vs = _f(vs)
    nz, nx = vp.shape
    pt = pad_top_of(fd_order)
    nzbc, nxbc = nz + pt + nbc, nx + 2 * nbc
    out = {k: np.zeros((nzbc, nxbc), order="F") for k in ("temp", "ca", "cl", "cm", "cm1", "b", "b1")}
    rc = lib().oracle_prep(_p(vp), _p(vs), nz, nx, nbc, dx, dt, fd_order, den,
                           *[_p(out[k]) for k in ("temp", "ca", "cl", "cm", "cm1", "b", "b1")])
    if rc != 0:
        raise ValueError(f"oracle_prep rejected the model (rc={rc})")
    return out


def damp_circle(vmin, nzbc, nxbc, nbc, dx, isfs, pad_top) -> np.ndarray:
    d = np.zeros((nzbc, nxbc), order="F")
    lib().oracle_damp_circle(vmin, nzbc, nxbc, nbc, dx, int(isfs), pad_top, _p(d))
    return d


def ricker(f: float, dt: float, nt: int) -> np.ndarray:
    w = np.zeros(nt)
    if lib().oracle_ricker(f, dt, nt, _p(w)) != 0:
        raise ValueError("nt is smaller than condition!")
    return w


# --------------------------------------------------------------------------
# FD time loops (C oracle) with the MEX argument lists
# --------------------------------------------------------------------------

COEF_ORDER = ("temp", "ca", "cl", "cm", "cm1", "b", "b1")


def staggerfd(iv, coefs: Dict[str, np.ndarray], s, precision: str = "d", snapshots: bool = True):
    """eigen_staggerfd_{double,single}(iv, temp, ca, cl, cm, cm1, b, b1, s)."""
    dt_ = np.float64 if precision == "d" else np.float32
    iv = _f(np.ravel(iv), dt_)
    nt, ng, in_wf, nz, nx = int(iv[0]), int(iv[4]), int(iv[12]), int(iv[13]), int(iv[14])
    cs = [_f(coefs[k], dt_) for k in COEF_ORDER]
    s = _f(np.ravel(s), dt_)
    su = np.zeros((nt, ng), dtype=dt_, order="F")
    sw = np.zeros((nt, ng), dtype=dt_, order="F")
    nrec = nt // in_wf if in_wf > 0 else 0
    if snapshots:
        sn = [np.zeros((nz, nx, nrec), dtype=dt_, order="F") for _ in range(4)]
        sp = [_p(a) for a in sn]
    else:
        sn = None
        sp = [None] * 4
    fn = getattr(lib(), f"oracle_staggerfd_{precision}")
    rc = fn(_p(iv), *[_p(c) for c in cs], _p(s), _p(su), _p(sw), *sp)
    if rc != 0:
        raise ValueError(f"oracle_staggerfd rejected the inputs (rc={rc})")
    snaps = None if sn is None else dict(zip(("fux", "fuz", "bwx", "bwz"), sn))
    return su, sw, snaps


def e2drtm(iv, coefs: Dict[str, np.ndarray], adj, snaps, precision: str = "d"):
    """eigen_e2drtm_{double,single}(iv, temp, ca, cl, cm, cm1, b, b1, adj, fux, fuz, bwx, bwz)."""
    dt_ = np.float64 if precision == "d" else np.float32
    iv = _f(np.ravel(iv), dt_)
    nz, nx = int(iv[11]), int(iv[12])
    cs = [_f(coefs[k], dt_) for k in COEF_ORDER]
    adj = _f(adj, dt_)
    sn = [_f(snaps[k], dt_) for k in ("fux", "fuz", "bwx", "bwz")]
    outs = [np.zeros((nz, nx), dtype=dt_, order="F") for _ in range(3)]
    fn = getattr(lib(), f"oracle_e2drtm_{precision}")
    rc = fn(_p(iv), *[_p(c) for c in cs], _p(adj), *[_p(a) for a in sn], *[_p(o) for o in outs])
    if rc != 0:
        raise ValueError(f"oracle_e2drtm rejected the inputs (rc={rc})")
    return tuple(outs)


SOURCE_TYPES = {"p": 1, "s": 2, "s1": 3, "z": 4, "w": 5}


def fwd_iv(nt, nzbc, nxbc, dtx, ng, sz, sx, gz1, gx1, dg, source_type, fd_order, in_wf, nz, nx, fmt=3):
    """input_vector of staggerfd_eigen.m:138 (sz/sx/gz/gx already padded, 1-based)."""
    return np.array([nt, nzbc, nxbc, dtx, ng, sz, sx, gz1, gx1, dg, source_type, fd_order, in_wf, nz, nx, fmt],
                    dtype=np.float64)


def adj_iv(nt, nzbc, nxbc, dtx, ng, nbc, gz1, gx1, dg, fd_order, in_wf, nz, nx, dt):
    """input_vector of e2drtm_eigen.m:121."""
    return np.array([nt, nzbc, nxbc, dtx, ng, nbc, gz1, gx1, dg, fd_order, in_wf, nz, nx, dt], dtype=np.float64)


def staggerfd_eigend(is_, nbc, nt, dtx, dx, dt, sx, sz, gx, gz, s, vp, vs, isfs, fd_order,
                     source_type, parameter_type, in_wf, precision="d", snapshots=True):
    """core/staggerfd_eigend.m:2-140 (precision 'f' gives staggerfd_eigen.m)."""
    if not isfs or parameter_type != 0:
        raise ValueError("only isfs=1, parameter_type=0 are defined by the wrapper")
    vp = _f(vp)
    nz, nx = vp.shape
    pt = pad_top_of(fd_order)
    c = prep(vp, vs, nbc, dx, dt, fd_order)
    gx = np.ravel(np.asarray(gx, dtype=np.float64))
    gz = np.ravel(np.asarray(gz, dtype=np.float64))
    iv = fwd_iv(nt, nz + pt + nbc, nx + 2 * nbc, dtx, gx.size, sz + pt, sx + nbc, gz[0] + pt, gx[0] + nbc,
                gx[1] - gx[0], SOURCE_TYPES[source_type], fd_order, in_wf, nz, nx)
    return staggerfd(iv, c, s, precision, snapshots)


def e2drtm_eigend(wavefield_gradient, seismo_w, is_, nbc, nt, dtx, dx, dt, gx, gz, s, vp, vs, isfs,
                  fd_order, parameter_type, in_wf, precision="d"):
    """core/e2drtm_eigend.m:8-122."""
    if not isfs or parameter_type != 0:
        raise ValueError("only isfs=1, parameter_type=0 are defined by the wrapper")
    vp = _f(vp)
    nz, nx = vp.shape
    pt = pad_top_of(fd_order)
    c = prep(vp, vs, nbc, dx, dt, fd_order)
    gx = np.ravel(np.asarray(gx, dtype=np.float64))
    gz = np.ravel(np.asarray(gz, dtype=np.float64))
    iv = adj_iv(nt, nz + pt + nbc, nx + 2 * nbc, dtx, gx.size, nbc, gz[0] + pt, gx[0] + nbc, gx[1] - gx[0],
                fd_order, in_wf, nz, nx, dt)
    return e2drtm(iv, c, seismo_w, wavefield_gradient, precision)


# --------------------------------------------------------------------------
# Linear Radon transform (RTr.m / RTl.m / RTrAD.m / RTlAD.m)
# --------------------------------------------------------------------------

def mround(x: float) -> int:
    """MATLAB round (half away from zero)."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def ccn_of(df: float, dt: float) -> int:
    """ccn = fix(1./df./dt), evaluated left to right in double (RTr.m:16)."""
    return int(math.trunc((1.0 / df) / dt))


def rt_window(side: str, is_: int, b: int, offset: int, m: int, M: int) -> Tuple[int, int]:
    """1-based first/last trace of the Radon window and L.

    right: RTr.m:5-11  (traces src+m .. nend, nend = min(src+m+offset-1, b))
    left:  RTl.m:5-11  (traces nend+1 .. src-m, nend = max(src-offset-m, 0)), read flipped.
    """
    src = M * is_ - M + 1
    if side == "r":
        nend = m + offset - 1 + src
        if nend > b:
            nend = b
        first = src + m
        return first, nend
    nend = src - offset - m
    if nend < 1:
        nend = 0
    return nend + 1, mround(src - m)


def fk_filter(d, dt=None, dx=None) -> np.ndarray:
    """fieldexamples/QadimahFault/core/fk_filter.m:48-53,283-296 (the branch
    that runs: the down-dip case): zero-pad d (ns x ntr) to 2ns x 2ntr,
    fft2, fftshift, zero the shifted quadrants (1:ns, 1:ntr) and
    (ns+1:2ns, ntr+1:2ntr), shift back, real(ifft2), crop to ns x ntr.
    dt and dx only label the axes there and do not enter."""
    d = np.asarray(d, dtype=np.float64)
    ns, ntr = d.shape
    D = np.zeros((2 * ns, 2 * ntr))
    D[:ns, :ntr] = d
    D = np.fft.fftshift(np.fft.fft2(D))
    D[:ns, :ntr] = 0
    D[ns:, ntr:] = 0
    return np.real(np.fft.ifft2(np.fft.fftshift(D)))[:ns, :ntr]


def rt_side(seismo, side: str, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M, fk: bool = False):
    """RTr/RTl (+AD) in float64.  Returns (ml, dataLen, saveForBackward dict).
    ``fk`` is the QadimahFault copies' FK == 1 branch: fk_filter on the
    (flipped) window before the FFT (fieldexamples/QadimahFault/core/RTr.m:14-16,
    RTl.m:15-17)."""
    seismo = np.asarray(seismo, dtype=np.float64)
    first, last = rt_window(side, is_, b, offset, m, M)
    uxt = seismo[:, first - 1:last]
    if side == "l":
        uxt = uxt[:, ::-1]
    if fk:
        uxt = fk_filter(uxt, dt, ds)
    L = uxt.shape[1]
    x = (m + np.linspace(0, L - 1, L)) * ds
    ccn = ccn_of(df, dt)
    nt = uxt.shape[0]
    if ccn >= nt:
        d = np.fft.fft(uxt, n=ccn, axis=0)
    else:
        d = np.fft.fft(uxt[:ccn], n=ccn, axis=0)
    d = d.T                                        # L x ccn (non-conjugate transpose)
    lf = mround(fmin / df) + 1
    nf = mround(fmax / df) + 1
    pp = 1.0 / np.arange(vmin, vmax + 1, dtype=np.float64)
    if pp.size != np_:
        raise ValueError("np must equal vmax-vmin+1")
    theta = (2 * np.pi * df) * np.outer(pp, x)     # ll0 = 1i*theta (RTr.m:22)
    mm = slant_stack(theta, d, lf, nf)
    ml = np.abs(mm)
    ml1 = ml[:, lf - 1:nf].copy()
    for luoi in range(lf, nf + 1):
        ml[:, luoi - 1] = ml[:, luoi - 1] / np.max(ml[:, luoi - 1])
    ml = ml[:, lf - 1:nf]
    sfb = {"mm": mm, "theta": theta, "lf": lf, "nf": nf, "ccn": ccn, "ml1": ml1,
           "uxtpos": np.arange(first, last + 1), "x": x}
    return ml, L, sfb


def slant_stack(theta, d, lf: int, nf: int) -> np.ndarray:
    """RTr.m:23-27: mm(:, c) = exp(1i*theta*(c-1)) * d(:, c) for c = lf..nf (1-based); other columns 0."""
    mm = np.zeros((theta.shape[0], nf), dtype=np.complex128)
    for luoj in range(lf, nf + 1):
        l = np.exp(1j * (theta * (luoj - 1)))
        mm[:, luoj - 1] = l @ d[:, luoj - 1]
    return mm


def slant_stack_adj(theta, gmm, lf: int, nf: int, ccn: int) -> np.ndarray:
    """ADWDgrad_1.m:35-40: graddr(:, c) = exp(1i*theta*(c-1))' * gmm(:, c-lf+1); L x ccn."""
    graddr = np.zeros((theta.shape[1], ccn), dtype=np.complex128)
    for luoj in range(lf, nf + 1):
        K = np.exp(1j * (theta * (luoj - 1)))
        graddr[:, luoj - 1] = K.conj().T @ gmm[:, luoj - lf]
    return graddr


def inv_dft_real(graddr, ccn: int, nt: int) -> np.ndarray:
    """ADWDgrad_1.m:42-43: real(ifft(graddr, ccn, 2))(:, 1:nt)."""
    if ccn < nt:
        raise ValueError("Index exceeds the number of array elements (ccn < nt)")
    return np.real(np.fft.ifft(graddr, n=ccn, axis=1))[:, :nt]


def RTr(seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M):
    ml, L, _ = rt_side(seismo, "r", is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M)
    return ml, L


def RTl(seismo, is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M):
    ml, L, _ = rt_side(seismo, "l", is_, df, dt, np_, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M)
    return ml, L


# --------------------------------------------------------------------------
# findpeaks / pickFDC1 / LHDispPick
# --------------------------------------------------------------------------

def findpeaks_locs(y) -> np.ndarray:
    """1-based locations of local maxima with MATLAB findpeaks defaults.

    The published rules: bookend the signal with NaN, keep the first sample of
    each run of equal values, take sign(diff); a peak is a +1 -> -1 transition,
    so endpoints are never peaks and a flat peak reports its first index.
    Samples at +Inf count as peaks (none occur here).
    """
    y = np.asarray(y, dtype=np.float64).ravel()
    n = y.size
    yt = np.concatenate(([np.nan], y, [np.nan]))
    it = np.arange(n + 2)
    fin = ~np.isnan(yt)
    keep = np.concatenate(([True], (yt[:-1] != yt[1:]) & (fin[:-1] | fin[1:])))
    # (NaN != NaN is True; the finite mask removes runs of NaN)
    idx = it[keep]
    with np.errstate(invalid="ignore"):
        s = np.sign(np.diff(yt[idx]))
        dsg = np.diff(s)
    imax = 1 + np.nonzero(dsg < 0)[0]
    locs = idx[imax]          # positions in the bookended vector == 1-based positions in y
    return locs.astype(np.int64)


def _nanmax_first(col) -> int:
    """MATLAB [~, i] = max(col): NaNs ignored, first index of the max, 1 if all NaN."""
    col = np.asarray(col, dtype=np.float64)
    if np.all(np.isnan(col)):
        return 1
    return int(np.nanargmax(col)) + 1


def pickFDC1(ml, lp: int, fg: float = 1.5, lfg: Optional[float] = None, lpvalue: Optional[int] = None) -> np.ndarray:
    """core/pickFDC1.m:1-131; returns cr_0 (1 x b, float, 1-based velocity index).
    With ``lpvalue`` it is pickFDCValue (fieldexamples/outcrop/core/pickFDCValue.m:1-134),
    which differs only in starting at the given row: idx(lp) = lpvalue (:20)."""
    ml = np.asarray(ml, dtype=np.float64)
    a, b = ml.shape
    if lfg is None:
        lfg = fg
    idx = np.zeros(b + 1, dtype=np.int64)       # 1-based arrays (slot 0 unused)
    dvmax = np.zeros(b + 1, dtype=np.int64)
    dvmin = np.zeros(b + 1, dtype=np.int64)
    qujian = np.full(b + 1, np.nan)
    chazhi = np.ones(b + 1, dtype=np.int64)

    def col(i):
        return ml[:, i - 1]

    idx[lp] = _nanmax_first(col(lp)) if lpvalue is None else int(lpvalue)
    for i in range(lp - 1, 0, -1):
        if idx[i + 1] == a:
            idx[i] = a
            dvmax[i] = dvmax[i + 1]
            dvmin[i] = dvmin[i + 1]
            continue
        if idx[i + 1] > a - 3:
            dvmax[i] = a
        else:
            locs = findpeaks_locs(-col(i)[idx[i + 1] - 1:a])
            dvmax[i] = a if locs.size == 0 else idx[i + 1] + int(locs.min())
        if idx[i + 1] < 3:
            dvmin[i] = 1
        else:
            locs = findpeaks_locs(-col(i)[0:idx[i + 1]])
            dvmin[i] = 1 if locs.size == 0 else int(locs.max())
        qujian[i] = dvmax[i] - dvmin[i]
        if qujian[i] > qujian[i + 1] * lfg:
            idx[i] = idx[i + 1]
            chazhi[i] = 0
            dvmax[i] = dvmax[i + 1]
            dvmin[i] = dvmin[i + 1]
            qujian[i] = dvmax[i] - dvmin[i]
        else:
            idx[i] = dvmin[i] + _nanmax_first(col(i)[dvmin[i] - 1:dvmax[i]]) - 1
    for i in range(lp + 1, b + 1):
        if idx[i - 1] > a - 3:
            dvmax[i] = a
        else:
            locs = findpeaks_locs(-col(i)[idx[i - 1] - 1:a])
            dvmax[i] = a if locs.size == 0 else idx[i - 1] + int(locs.min())
        if idx[i - 1] < 3:
            dvmin[i] = 1
        else:
            locs = findpeaks_locs(-col(i)[0:idx[i - 1]])
            dvmin[i] = 1 if locs.size == 0 else int(locs.max())
        qujian[i] = dvmax[i] - dvmin[i]
        if qujian[i] > fg * qujian[i - 1]:
            idx[i] = idx[i - 1]
            chazhi[i] = 0
            dvmax[i] = dvmax[i - 1]
            dvmin[i] = dvmin[i - 1]
            qujian[i] = dvmax[i] - dvmin[i]
        else:
            idx[i] = dvmin[i] + _nanmax_first(col(i)[dvmin[i] - 1:dvmax[i]]) - 1
    idx = idx[1:].astype(np.float64)
    chazhi = chazhi[1:]
    que = np.nonzero(chazhi == 1)[0] + 1                 # 1-based, ascending
    if que.size < 2:
        raise ValueError("interp1 needs at least two sample points")
    xq = np.arange(1, b + 1, dtype=np.float64)
    idxc = _interp1_linear(que.astype(np.float64), idx[que - 1], xq)
    good = np.nonzero(~np.isnan(idxc))[0] + 1
    return _interp1_nearest_extrap(good.astype(np.float64), idxc[good - 1], xq)


def _interp1_linear(x, v, xq):
    """interp1(x, v, xq, 'linear'): NaN outside [x(1), x(end)]."""
    out = np.full(xq.shape, np.nan)
    for k, q in enumerate(xq):
        if q < x[0] or q > x[-1]:
            continue
        j = int(np.searchsorted(x, q, side="right")) - 1
        if j >= x.size - 1:
            j = x.size - 2
        t = (q - x[j]) / (x[j + 1] - x[j])
        out[k] = v[j] + t * (v[j + 1] - v[j])
    return out


def _interp1_nearest_extrap(x, v, xq):
    """interp1(x, v, xq, 'nearest', 'extrap'); x integer and contiguous here."""
    out = np.empty(xq.shape)
    for k, q in enumerate(xq):
        j = int(np.argmin(np.abs(x - q)))
        # ties go to the larger sample point
        if j + 1 < x.size and abs(x[j + 1] - q) == abs(x[j] - q):
            j = j + 1
        out[k] = v[j]
    return out


def LHDispPick(ml, vmin, cr_0, method: int, ini: int):
    """core/LHDispPick.m:2-13 -> (res, cr_pre) with cr_pre a column of npair values."""
    if method == 1:
        idx = pickFDC1(ml, ini, 2, 4)
        cr_pre = idx + vmin
    else:
        cr_pre = np.array([_nanmax_first(c) for c in np.asarray(ml).T], dtype=np.float64) + vmin
    res = float(np.mean((cr_pre - np.ravel(cr_0)) ** 2))
    return res, cr_pre


def LHDispPick_field(ml, mlr, vmin, cr_0, res_first: int, res_last: int, start_row: str = "ml"):
    """The field scripts' LHDispPick (method 1) -> (res, cr_pre).

    Start column ini = the first column holding the largest |mm|
    (``[~,ini] = max(max(mlr))``).  ``start_row="mlr"`` is
    fieldexamples/outcrop/core/LHDispPick.m:3-14: pickFDCValue from the peak
    row of mlr(:, ini); ``start_row="ml"`` is
    fieldexamples/QadimahFault/core/LHDispPick.m:4-25: pickFDC1 from ini.
    The misfit is the mean over cr_pre(res_first:end-res_last) (outcrop 15,
    15 at :21; Qadimah 50, 30 at :25)."""
    mlr = np.asarray(mlr, dtype=np.float64)
    colmax = np.array([np.nan if np.all(np.isnan(c)) else np.nanmax(c) for c in mlr.T])
    ini = _nanmax_first(colmax)
    if start_row == "mlr":
        idx = pickFDC1(ml, ini, 2, 4, lpvalue=_nanmax_first(mlr[:, ini - 1]))
    else:
        idx = pickFDC1(ml, ini, 2, 4)
    cr_pre = idx + vmin
    d = (cr_pre - np.ravel(cr_0))[res_first - 1:cr_pre.size - res_last]
    with np.errstate(invalid="ignore"):
        res = float(np.mean(d ** 2)) if d.size else float("nan")
    return res, cr_pre, ini


def norm_trace(data_in) -> np.ndarray:
    """fieldexamples/outcrop/core/norm_trace.m:4-16 (the QadimahFault copy is
    the same): each column divided by its max |x|; MATLAB's max skips NaN;
    a zero column is 0/0 = NaN."""
    d = np.asarray(data_in, dtype=np.float64)
    peak = np.fmax.reduce(np.abs(d), axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        return d / peak[None, :]


# --------------------------------------------------------------------------
# SoftMax adjoint source (ADWDgrad_1.m)
# --------------------------------------------------------------------------

def softmax_cols(X) -> np.ndarray:
    """Column-wise softmax with the max subtracted (exp(x - max) / sum)."""
    X = np.asarray(X, dtype=np.float64)
    e = np.exp(X - X.max(axis=0, keepdims=True))
    return e / e.sum(axis=0, keepdims=True)


def softargmax_grad(ml1, g, beta) -> np.ndarray:
    """ADWDgrad_1.m:21-31: gml(j,i) = 2*beta*g_i * E_j*(j - sum_k k E_k), E = softmax(beta*ml1^(1/40)).

    g == 0 entries are replaced by 1 (ADWDgrad_1.m:26).  The reference sums the
    np x np broadcast (space_M .* E)' .* E over its columns; E_j * sum_k(k E_k)
    is the same sum factored.
    """
    G = np.asarray(ml1, dtype=np.float64) ** (1.0 / 40)
    E = softmax_cols(G * beta)
    npv = E.shape[0]
    j = np.arange(1, npv + 1, dtype=np.float64)[:, None]
    g = np.array(np.ravel(g), dtype=np.float64)
    g[g == 0] = 1
    t3 = E * j - E * np.sum(j * E, axis=0, keepdims=True)
    return t3 * 2 * g[None, :] * beta


def adjoint_side(sfb, g, beta, nt) -> np.ndarray:
    """One branch of ADWDgrad_1.m (lines 14-43 / 48-75): returns graduxt (L x nt)."""
    lf, nf, ccn = sfb["lf"], sfb["nf"], sfb["ccn"]
    if ccn < nt:
        raise ValueError("Index exceeds the number of array elements (ccn < nt)")
    gml = softargmax_grad(sfb["ml1"], g, beta)
    mm = sfb["mm"][:, lf - 1:nf]
    with np.errstate(invalid="ignore", divide="ignore"):
        gmm = mm / np.abs(mm)
    gmm[np.isnan(gmm)] = 0
    gmm = gmm * gml
    graddr = slant_stack_adj(sfb["theta"], gmm, lf, nf, ccn)
    return inv_dft_real(graddr, ccn, nt)


def ADWDgrad_1(nt, ng, ns, npair, is_, w, m, M, SoftArgNorm, grad_outputr, grad_outputl, sfbr, sfbl,
               peak_norm: bool = False):
    """core/ADWDgrad_1.m:1-82 -> a_data_res (nt x ng).  ``peak_norm`` is the
    field copies' extra line 79, a_data_res./max(abs(a_data_res(:)))
    (fieldexamples/outcrop/core/ADWDgrad_1.m, QadimahFault/core/ADWDgrad_1.m)."""
    a = np.zeros((nt, ng))
    if is_ <= ns - math.floor(m / M) - mround(w / M):
        gr = adjoint_side(sfbr, grad_outputr, SoftArgNorm, nt)
        a[:, sfbr["uxtpos"] - 1] = gr.T
    if is_ >= mround(w / M) + math.floor(m / M) + 1:
        gl = adjoint_side(sfbl, grad_outputl, SoftArgNorm, nt)
        pos = sfbl["uxtpos"][:-1]
        a[:, pos - 1] = gl[::-1][:-1].T if gl.shape[0] > 1 else np.zeros((nt, 0))
    if peak_norm:
        with np.errstate(invalid="ignore", divide="ignore"):
            a = a / np.fmax.reduce(np.abs(a), axis=None)
    return a


# --------------------------------------------------------------------------
# Gradient assembly (SWD.m:199-201) and smooth2a.m
# --------------------------------------------------------------------------

def assemble_dk_vs(vs, g_cl, g_cm, g_illum) -> np.ndarray:
    gcl = g_cl / g_illum
    gcm = g_cm / g_illum
    return -4 * vs * gcl + 2 * vs * gcm


def smooth2a(A, Nr: int, Nc: int) -> np.ndarray:
    """core/smooth2a.m:52-76: (2Nr+1) x (2Nc+1) box mean, truncated edges, NaNs ignored."""
    A = np.array(A, dtype=np.float64)
    row, col = A.shape
    eL = np.zeros((row, row))
    for d in range(-Nr, Nr + 1):
        eL += np.eye(row, k=d)
    eR = np.zeros((col, col))
    for d in range(-Nc, Nc + 1):
        eR += np.eye(col, k=d)
    nanm = np.isnan(A)
    A[nanm] = 0
    nrm = eL @ (~nanm).astype(np.float64) @ eR
    nrm[nanm] = np.nan
    return (eL @ A @ eR) / nrm


# --------------------------------------------------------------------------
# Step length and line search of one SWD iteration (SWD.m:203-322)
# --------------------------------------------------------------------------

def swd_step_model(vs, dk, alpha, f, vsmin, vsmax) -> np.ndarray:
    """SWD.m:206-207: vs1 = vs + alpha*f*dk_vs, clipped to [vsmin, vsmax] (NaN passes)."""
    vs1 = vs + (alpha * f) * dk
    vs1[vs1 < vsmin] = vsmin
    vs1[vs1 > vsmax] = vsmax
    return vs1


def swd_line_search(vs, dk, res0, misfit, pur, vsmin, vsmax, f1, f2=float("nan"), res2=float("nan"),
                    alpha=None):
    """SWD.m:203-322 for one iteration.  ``misfit(vs1)`` is the mean
    right-side pick misfit of a forward-only batch of model vs1; ``dk`` is
    the smoothed gradient; f1 / f2 / res2 are the script's carried values.
    Returns (accepted model, dict of alpha, gama, res0, f1, f2, res1, res2,
    res3, forwards)."""
    if alpha is None:
        alpha = np.sqrt(np.sum(vs * vs)) / np.sqrt(np.sum(dk * dk)) * pur          # SWD.m:203-205
    n = 0

    def trial(f):
        nonlocal n
        n += 1
        return misfit(swd_step_model(vs, dk, alpha, f, vsmin, vsmax))

    res1 = trial(f1)                                                               # SWD.m:206-229
    if res1 > res0:
        while res1 > res0 and f1 > 0.0001:                                         # SWD.m:233-256
            f2 = f1
            res2 = res1
            f1 = f1 * 0.5
            res1 = trial(f1)
    else:
        f2 = f1 * 2                                                                # SWD.m:257-278
        res2 = trial(f2)
    with np.errstate(divide="ignore", invalid="ignore"):  # IEEE division as MATLAB: 0/0 = NaN, x/0 = +-Inf
        gama = float(np.float64(f1 ** 2 * (res0 - res2) + f2 ** 2 * (res1 - res0)) /
                     np.float64(2 * res0 * (f1 - f2) + 2 * res1 * f2 - 2 * res2 * f1))
    if np.isinf(gama):                                                             # SWD.m:279-284
        gama = 0
    res3 = trial(gama)                                                             # SWD.m:286-306
    if res3 > res1 or res3 > res2:                                                 # SWD.m:308-322
        if res1 > res2:
            res0 = res2
            gama = f2
        else:
            res0 = res1
            gama = f1
    else:
        res0 = res3
    vs1 = swd_step_model(vs, dk, alpha, gama, vsmin, vsmax)
    return vs1, {"alpha": alpha, "gama": gama, "res0": res0, "f1": f1, "f2": f2, "res1": res1, "res2": res2,
                 "res3": res3, "forwards": n}


# --------------------------------------------------------------------------
# One SWD shot-gradient (the parfor body SWD.m:171-194) for tests
# --------------------------------------------------------------------------

def pick_side(ml, sfb, vmin, cr_0, ini, variant=None):
    """LHDispPick of SWD.m (variant None) or of the field scripts (variant
    dict with auto_start, res_first, res_last) -> (res, cr_pre)."""
    v = variant or {}
    rf, rl = max(int(v.get("res_first", 1)), 1), int(v.get("res_last", 0))
    if v.get("auto_start"):
        res, cr, _ = LHDispPick_field(ml, sfb["ml1"], vmin, cr_0, rf, rl, v.get("start_row", "ml"))
        return res, cr
    res, cr = LHDispPick(ml, vmin, cr_0, 1, ini)
    if (rf, rl) != (1, 0):
        d = (cr - np.ravel(cr_0))[rf - 1:cr.size - rl]
        res = float(np.mean(d ** 2)) if d.size else float("nan")
    return res, cr


def swd_shot(vp, vs, s, nbc, nt, dx, dt, sx, sz, gx, gz, in_wf, fd_order, is_, ns, radon, cr0_r, cr0_l,
             beta, ini, w, precision="d", variant=None):
    """Forward, RTrAD/RTlAD, LHDispPick, ADWDgrad_1, adjoint for one shot.

    ``radon`` = dict(df, fmin, fmax, vmin, vmax, ds, offset, m, M).  Returns a
    dict with seismo_w, picks, residuals, adjoint source and the three images.
    ``variant`` (dict: norm_trace, auto_start, res_first, res_last,
    adj_peak_norm) is the field scripts' parfor body
    (fieldexamples/outcrop/outcrop.m:163-186 with its core/ copies).
    """
    v = variant or {}
    dtx = dt / dx
    su, sw, snaps = staggerfd_eigend(is_, nbc, nt, dtx, dx, dt, sx, sz, gx, gz, s, vp, vs, 1, fd_order, "w", 0,
                                     in_wf, precision)
    r = radon
    ng = np.size(gx)
    np_ = r["vmax"] - r["vmin"] + 1
    out = {"seismo_w": sw}
    sw_rt = norm_trace(sw) if v.get("norm_trace") else sw
    fk = bool(v.get("fk_filter"))
    gr = gl = np.zeros(mround((r["fmax"] - r["fmin"]) / r["df"]) + 1)
    sfbr = sfbl = None
    M, m = r["M"], r["m"]
    if is_ <= ns - math.floor(m / M) - mround(w / M):
        ml, _, sfbr = rt_side(sw_rt, "r", is_, r["df"], dt, np_, r["vmin"], r["vmax"], r["fmin"], r["fmax"], nt, ng,
                              r["ds"], r["offset"], m, M, fk)
        out["res_r"], out["cr_r"] = pick_side(ml, sfbr, r["vmin"], cr0_r, ini, v)
        out["ml_r"] = ml
        gr = out["cr_r"] - np.ravel(cr0_r)
    if is_ >= mround(w / M) + math.floor(m / M) + 1:
        ml, _, sfbl = rt_side(sw_rt, "l", is_, r["df"], dt, np_, r["vmin"], r["vmax"], r["fmin"], r["fmax"], nt, ng,
                              r["ds"], r["offset"], m, M, fk)
        out["res_l"], out["cr_l"] = pick_side(ml, sfbl, r["vmin"], cr0_l, ini, v)
        out["ml_l"] = ml
        gl = out["cr_l"] - np.ravel(cr0_l)
    npair = gr.size
    adj = ADWDgrad_1(nt, ng, ns, npair, is_, w, m, M, beta, gr, gl, sfbr, sfbl, bool(v.get("adj_peak_norm")))
    out["adj"] = adj
    cl_img, cm_img, illum = e2drtm_eigend(snaps, adj, is_, nbc, nt, dtx, dx, dt, gx, gz, s, vp, vs, 1, fd_order, 0,
                                          in_wf, precision)
    out.update(cl_img=cl_img, cm_img=cm_img, illum=illum)
    return out
