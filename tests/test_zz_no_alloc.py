"""swd_lhdisppick_noalloc (no device allocation, fixed per-GPU scratch array)
against swd_lhdisppick through the C ABI: picks and misfit are index
arithmetic plus one FP64 mean, so they must be equal bit for bit.  Two calls
in a row, the second with a smaller spectrum (the scratch keeps the first
call's contents), both pick methods, and the size limit of the scratch."""
import ctypes

import numpy as np
import pytest

from swd_amd import _lib

pytestmark = pytest.mark.gpu


def ridge_spectrum(np_, npair, seed):
    """Column-max normalised |mm|-like spectra: one dispersive ridge plus a
    weaker second mode and noise, so the tracking and its windows have work."""
    r = np.random.default_rng(seed)
    j = np.arange(np_)[:, None]
    f = np.arange(npair)[None, :]
    c1 = 0.25 * np_ + 0.5 * np_ * np.exp(-f / (0.4 * npair))
    c2 = 0.6 * np_ + 0.2 * np_ * np.exp(-f / (0.3 * npair))
    ml = np.exp(-((j - c1) / (0.03 * np_ + 2)) ** 2) + 0.5 * np.exp(-((j - c2) / (0.05 * np_ + 2)) ** 2)
    ml += 0.05 * r.random((np_, npair))
    ml /= ml.max(axis=0, keepdims=True)
    return np.asfortranarray(ml, dtype=np.float32)


def call(name, ml, vmin, cr0, method, ini):
    L = _lib.load()
    np_, npair = ml.shape
    res = ctypes.c_double(0.0)
    cr = np.zeros(npair)
    rc = getattr(L, name)(_lib.ptr(ml), np_, npair, vmin, _lib.ptr(cr0), method, ini, ctypes.byref(res),
                          _lib.ptr(cr))
    return rc, res.value, cr


@pytest.mark.parametrize("method", [1, 2])
def test_noalloc_pick_equals_allocating_pick(method):
    cases = [(701, 651, 3), (301, 151, 4)]      # the second call is smaller
    for np_, npair, seed in cases:
        print(f"[no-alloc] method {method}: {np_} x {npair}", flush=True)
        ml = ridge_spectrum(np_, npair, seed)
        cr0 = 100.0 + np.round(np.random.default_rng(seed + 10).uniform(0.2, 0.7, npair) * np_)
        ra, resa, cra = call("swd_lhdisppick", ml, 100.0, cr0, method, 10)
        rb, resb, crb = call("swd_lhdisppick_noalloc", ml, 100.0, cr0, method, 10)
        assert ra == 0 and rb == 0, (ra, rb, _lib.load().swd_last_error())
        np.testing.assert_array_equal(crb, cra)
        assert resb == resa
        assert np.all(np.isfinite(cra)) and np.all((cra >= 100) & (cra <= 100 + np_))


def test_noalloc_pick_rejects_spectra_beyond_its_scratch():
    ml = np.zeros((2048, 1024), dtype=np.float32, order="F")
    rc, _, _ = call("swd_lhdisppick_noalloc", ml, 100.0, np.zeros(1024), 1, 10)
    assert rc == -2
    assert b"no-allocation scratch" in _lib.load().swd_last_error()
