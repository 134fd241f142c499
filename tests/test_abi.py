"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-only routines and argument validation behave (no GPU needed)."""
import os
import re

import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import _lib, fd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "swd_abi.h")).read()
    return sorted(set(re.findall(r"\b(swd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    decl = declared_symbols()
    assert len(decl) >= 10
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) <= set(_lib.SIGNATURES), set(decl) - set(_lib.SIGNATURES)


def test_prepare_model_matches_oracle():
    r = np.random.default_rng(1)
    vp = r.uniform(500, 1500, (13, 17))
    vs = vp / r.uniform(1.6, 2.0, vp.shape)
    for order in (22, 24, 28):
        a = fd.prepare_model(vp, vs, 6, 0.5, 8e-5, order)
        b = O.prep(vp, vs, 6, 0.5, 8e-5, order)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_ricker_matches_oracle():
    np.testing.assert_array_equal(fd.ricker(25, 8e-5, 6000), O.ricker(25, 8e-5, 6000))
    with pytest.raises(_lib.SwdError):
        fd.ricker(25, 8e-5, 10)


def test_mex_entry_rejects_bad_descriptors_before_touching_the_device():
    c = {k: np.ones((12, 14)) for k in ("temp", "ca", "cl", "cm", "cm1", "b", "b1")}
    bad = [
        O.fwd_iv(7, 12, 14, 0.3, 4, 5, 6, 3, 3, 1, 5, 22, 2, 6, 8),     # nt % in_wf
        O.fwd_iv(4, 12, 14, 0.3, 20, 5, 6, 3, 3, 1, 5, 22, 2, 6, 8),    # receivers off grid
        O.fwd_iv(4, 12, 14, 0.3, 4, 5, 6, 3, 3, 1, 5, 23, 2, 6, 8),     # fd_order
    ]
    for iv in bad:
        with pytest.raises(_lib.SwdError) as e:
            fd.eigen_staggerfd_single(iv, *[c[k] for k in O.COEF_ORDER], np.zeros(8))
        assert e.value.code == -2


def test_adwdgrad1_rejects_ccn_below_nt_before_touching_the_device():
    """ADWDgrad_1.m:42-43 reads graduxt(:, 1:nt) of an ifft of length ccn:
    with ccn < nt MATLAB errors; the library returns SWD_ERR_ARG."""
    import ctypes
    L = _lib.load()
    nt, ng, np_, npair = 600, 20, 50, 10
    mm = np.ones((2 * np_, npair), dtype=np.float32, order="F")
    ml = np.ones((np_, npair), dtype=np.float32, order="F")
    g = np.zeros(npair)
    out = np.zeros((nt, ng), dtype=np.float32, order="F")
    P = _lib.ptr
    rc = L.swd_adwdgrad1(nt, ng, 20, npair, 5, 0.0, 0, 1, 10.0, P(g), P(g), P(mm), P(ml), 5, 10, P(mm), P(ml), 5, 5,
                         np_, 100.0, 0.5, 0.5, 11, 500, P(out))
    assert rc == -2
    assert b"ccn" in L.swd_last_error()
    # the oracle's restatement fails the same way
    sfb = {"lf": 11, "nf": 20, "ccn": 500}
    with pytest.raises(ValueError):
        O.adjoint_side(sfb, g, 10.0, nt)
