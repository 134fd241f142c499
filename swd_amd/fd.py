"""Host mirror of SWD's FD entry points (MEX and .m wrapper signatures).

* ``eigen_staggerfd_single/double`` and ``eigen_e2drtm_single/double`` take
  the MEX argument lists (core/eigen_staggerfd_single.cpp:68,
  core/eigen_e2drtm_single.cpp:67) and call ``swd_staggerfd_f32/f64`` and
  ``swd_e2drtm_f32/f64`` of libswd.
* ``staggerfd_eigen(d)`` and ``e2drtm_eigen(d)`` mirror the MATLAB wrappers
  (core/staggerfd_eigen.m:2-141, core/e2drtm_eigen.m:8-123): model
  preparation through ``swd_prepare_model``, then the MEX call.
All arrays are MATLAB column-major (numpy Fortran order).
"""
from __future__ import annotations

import numpy as np

from ._lib import check, fortran, load, ptr

SOURCE_TYPES = {"p": 1, "s": 2, "s1": 3, "z": 4, "w": 5}


def _staggerfd(dtype, fn, iv, temp, ca, cl, cm, cm1, b, b1, s, snapshots=True, want_u=True):
    iv = np.ascontiguousarray(np.ravel(iv), dtype=dtype)
    if iv.size != 16:
        raise ValueError("input_vector must have 16 entries")
    nt, ng, in_wf, nz, nx = int(iv[0]), int(iv[4]), int(iv[12]), int(iv[13]), int(iv[14])
    cs = [fortran(a, dtype) for a in (temp, ca, cl, cm, cm1, b, b1)]
    s = np.ascontiguousarray(np.ravel(s), dtype=dtype)
    if s.size < nt:
        raise ValueError("source wavelet shorter than nt")
    sw = np.zeros((nt, ng), dtype=dtype, order="F")
    su = np.zeros((nt, ng), dtype=dtype, order="F") if want_u else None
    nrec = nt // in_wf if in_wf > 0 else 0
    sn = [np.zeros((nz, nx, nrec), dtype=dtype, order="F") for _ in range(4)] if snapshots else [None] * 4
    check(fn(ptr(iv), *[ptr(c) for c in cs], ptr(s), ptr(su), ptr(sw), *[ptr(a) for a in sn]))
    S = dict(zip(("fux", "fuz", "bwx", "bwz"), sn)) if snapshots else None
    return su, sw, S


def eigen_staggerfd_single(iv, temp, ca, cl, cm, cm1, b, b1, s, snapshots=True):
    """[seismo_u, seismo_w, S] = eigen_staggerfd_single(...) on the GPU (FP32)."""
    return _staggerfd(np.float32, load().swd_staggerfd_f32, iv, temp, ca, cl, cm, cm1, b, b1, s, snapshots)


def eigen_staggerfd_double(iv, temp, ca, cl, cm, cm1, b, b1, s, snapshots=True):
    """[seismo_u, seismo_w, S] = eigen_staggerfd_double(...) on the GPU (FP64)."""
    return _staggerfd(np.float64, load().swd_staggerfd_f64, iv, temp, ca, cl, cm, cm1, b, b1, s, snapshots)


def _e2drtm(dtype, fn, iv, temp, ca, cl, cm, cm1, b, b1, adj, fux, fuz, bwx, bwz):
    iv = np.ascontiguousarray(np.ravel(iv), dtype=dtype)
    if iv.size != 14:
        raise ValueError("input_vector must have 14 entries")
    nz, nx = int(iv[11]), int(iv[12])
    cs = [fortran(a, dtype) for a in (temp, ca, cl, cm, cm1, b, b1)]
    adj = fortran(adj, dtype)
    sn = [fortran(a, dtype) for a in (fux, fuz, bwx, bwz)]
    out = [np.zeros((nz, nx), dtype=dtype, order="F") for _ in range(3)]
    check(fn(ptr(iv), *[ptr(c) for c in cs], ptr(adj), *[ptr(a) for a in sn], *[ptr(o) for o in out]))
    return tuple(out)


def eigen_e2drtm_single(iv, temp, ca, cl, cm, cm1, b, b1, adj, fux, fuz, bwx, bwz):
    """[cl_img, cm_img, illum_div] = eigen_e2drtm_single(...) on the GPU."""
    return _e2drtm(np.float32, load().swd_e2drtm_f32, iv, temp, ca, cl, cm, cm1, b, b1, adj, fux, fuz, bwx, bwz)


def eigen_e2drtm_double(iv, temp, ca, cl, cm, cm1, b, b1, adj, fux, fuz, bwx, bwz):
    """[cl_img, cm_img, illum_div] = eigen_e2drtm_double(...) on the GPU."""
    return _e2drtm(np.float64, load().swd_e2drtm_f64, iv, temp, ca, cl, cm, cm1, b, b1, adj, fux, fuz, bwx, bwz)


def set_time_block(nt_block: int, rows: int = 0) -> None:
    """Steps per fused FD launch of the MEX entry points (1 = single-step
    kernels) and band rows of the time-blocked kernels (0 = automatic)."""
    check(load().swd_set_fd_time_block(int(nt_block), int(rows)))


def set_fd_path(mode: int) -> int:
    """0 = automatic (fused when possible), 1 = plain kernels only. Returns the old mode."""
    return load().swd_set_fd_path(int(mode))


# --------------------------------------------------------------------------
# wrapper-level preparation (staggerfd_eigen.m:37-107)
# --------------------------------------------------------------------------

def pad_top_of(fd_order: int) -> int:
    return (fd_order - 20) // 2 + 1


def prepare_model(vp, vs, nbc: int, dx: float, dt: float, fd_order: int = 22):
    """temp, ca, cl, cm, cm1, b, b1 (nzbc x nxbc, float64) as the wrapper builds them."""
    vp = fortran(vp, np.float64)
    vs = fortran(vs, np.float64)
    if vp.shape != vs.shape:
        raise ValueError("vp and vs must have the same shape")
    nz, nx = vp.shape
    nzbc, nxbc = nz + pad_top_of(fd_order) + nbc, nx + 2 * nbc
    out = [np.zeros((nzbc, nxbc), order="F") for _ in range(7)]
    check(load().swd_prepare_model(ptr(vp), ptr(vs), nz, nx, nbc, dx, dt, fd_order, *[ptr(o) for o in out]))
    return dict(zip(("temp", "ca", "cl", "cm", "cm1", "b", "b1"), out))


def ricker(f: float, dt: float, nt: int) -> np.ndarray:
    """core/ricker.m with nargin == 3."""
    w = np.zeros(nt)
    check(load().swd_ricker(f, dt, nt, ptr(w)))
    return w


def _fwd_iv(nt, vp_shape, nbc, dtx, gx, gz, sx, sz, source_type, fd_order, in_wf):
    nz, nx = vp_shape
    pt = pad_top_of(fd_order)
    gx = np.ravel(np.asarray(gx, dtype=np.float64))
    gz = np.ravel(np.asarray(gz, dtype=np.float64))
    dg = gx[1] - gx[0] if gx.size > 1 else 1.0
    return np.array([nt, nz + pt + nbc, nx + 2 * nbc, dtx, gx.size, sz + pt, sx + nbc, gz[0] + pt, gx[0] + nbc,
                     dg, SOURCE_TYPES[source_type], fd_order, in_wf, nz, nx, 3], dtype=np.float64)


def _wrapper_fwd(single, is_, nbc, nt, dtx, dx, dt, sx, sz, gx, gz, s, vp, vs, isfs, fd_order, source_type,
                 parameter_type, in_wf, snapshots):
    if not isfs or parameter_type != 0:
        raise ValueError("the wrapper defines b, b1 and cm1 only for isfs=1, parameter_type=0")
    c = prepare_model(vp, vs, nbc, dx, dt, fd_order)
    iv = _fwd_iv(nt, np.shape(vp), nbc, dtx, gx, gz, sx, sz, source_type, fd_order, in_wf)
    args = (c["temp"], c["ca"], c["cl"], c["cm"], c["cm1"], c["b"], c["b1"], s)
    if single:
        return eigen_staggerfd_single(iv, *args, snapshots=snapshots)
    return eigen_staggerfd_double(iv, *args, snapshots=snapshots)


def staggerfd_eigen(is_, nbc, nt, dtx, dx, dt, sx, sz, gx, gz, s, vp, vs, isfs, fd_order, source_type,
                    parameter_type, in_wf, snapshots=True):
    """[seismo_u, seismo_w, www] = staggerfd_eigen(...) (FP32 MEX)."""
    return _wrapper_fwd(True, is_, nbc, nt, dtx, dx, dt, sx, sz, gx, gz, s, vp, vs, isfs, fd_order, source_type,
                        parameter_type, in_wf, snapshots)


def staggerfd_eigend(is_, nbc, nt, dtx, dx, dt, sx, sz, gx, gz, s, vp, vs, isfs, fd_order, source_type,
                     parameter_type, in_wf, snapshots=True):
    """[seismo_u, seismo_w, www] = staggerfd_eigend(...) (FP64 MEX)."""
    return _wrapper_fwd(False, is_, nbc, nt, dtx, dx, dt, sx, sz, gx, gz, s, vp, vs, isfs, fd_order, source_type,
                        parameter_type, in_wf, snapshots)


def _wrapper_adj(single, wavefield_gradient, seismo_w, is_, nbc, nt, dtx, dx, dt, gx, gz, s, vp, vs, isfs, fd_order,
                 parameter_type, in_wf):
    if not isfs or parameter_type != 0:
        raise ValueError("the wrapper defines b, b1 and cm1 only for isfs=1, parameter_type=0")
    c = prepare_model(vp, vs, nbc, dx, dt, fd_order)
    nz, nx = np.shape(vp)
    pt = pad_top_of(fd_order)
    gx = np.ravel(np.asarray(gx, dtype=np.float64))
    gz = np.ravel(np.asarray(gz, dtype=np.float64))
    dg = gx[1] - gx[0] if gx.size > 1 else 1.0
    iv = np.array([nt, nz + pt + nbc, nx + 2 * nbc, dtx, gx.size, nbc, gz[0] + pt, gx[0] + nbc, dg, fd_order,
                   in_wf, nz, nx, dt], dtype=np.float64)
    w = wavefield_gradient
    args = (c["temp"], c["ca"], c["cl"], c["cm"], c["cm1"], c["b"], c["b1"], seismo_w, w["fux"], w["fuz"],
            w["bwx"], w["bwz"])
    if single:
        return eigen_e2drtm_single(iv, *args)
    return eigen_e2drtm_double(iv, *args)


def e2drtm_eigen(wavefield_gradient, seismo_w, is_, nbc, nt, dtx, dx, dt, gx, gz, s, vp, vs, isfs, fd_order,
                 parameter_type, in_wf):
    """[cl_img, cm_img, illum_div] = e2drtm_eigen(...) (FP32 MEX)."""
    return _wrapper_adj(True, wavefield_gradient, seismo_w, is_, nbc, nt, dtx, dx, dt, gx, gz, s, vp, vs, isfs,
                        fd_order, parameter_type, in_wf)


def e2drtm_eigend(wavefield_gradient, seismo_w, is_, nbc, nt, dtx, dx, dt, gx, gz, s, vp, vs, isfs, fd_order,
                  parameter_type, in_wf):
    """[cl_img, cm_img, illum_div] = e2drtm_eigend(...) (FP64 MEX)."""
    return _wrapper_adj(False, wavefield_gradient, seismo_w, is_, nbc, nt, dtx, dx, dt, gx, gz, s, vp, vs, isfs,
                        fd_order, parameter_type, in_wf)
