"""The host code of libswd under AddressSanitizer and UndefinedBehaviorSanitizer
(libswd_san.so, python -m swd_amd.build --sanitize): the host-only entry points
(model preparation, Ricker, argument validation of the MEX, pick, adjoint
source and smooth2a entry points, the no-allocation pick's size check) run
in a child process with gcc's libasan preloaded; any heap overflow, use after
free or undefined behaviour aborts it.  No GPU is needed."""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBASAN = "/usr/lib/x86_64-linux-gnu/libasan.so.6"

CHILD = textwrap.dedent("""
    import ctypes, sys
    import numpy as np
    sys.path.insert(0, ROOT)
    from swd_amd import _lib, fd
    from oracle import swd_oracle as O
    L = _lib.load()
    assert L._name.endswith("libswd_san.so"), L._name
    r = np.random.default_rng(1)
    vp = r.uniform(500, 1500, (13, 17)); vs = vp / 1.8
    for order in (22, 24, 26, 28):
        a = fd.prepare_model(vp, vs, 6, 0.5, 8e-5, order)
        b = O.prep(vp, vs, 6, 0.5, 8e-5, order)
        for k in a:
            assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(fd.ricker(25, 8e-5, 6000), O.ricker(25, 8e-5, 6000))
    c = {k: np.ones((12, 14)) for k in O.COEF_ORDER}
    for iv in (O.fwd_iv(7, 12, 14, 0.3, 4, 5, 6, 3, 3, 1, 5, 22, 2, 6, 8),
               O.fwd_iv(4, 12, 14, 0.3, 20, 5, 6, 3, 3, 1, 5, 22, 2, 6, 8),
               O.fwd_iv(4, 12, 14, 0.3, 4, 5, 6, 3, 3, 1, 5, 23, 2, 6, 8)):
        try:
            fd.eigen_staggerfd_single(iv, *[c[k] for k in O.COEF_ORDER], np.zeros(8))
            raise AssertionError("accepted a bad descriptor")
        except _lib.SwdError as e:
            assert e.code == -2
    P = _lib.ptr
    ml = np.zeros((2048, 1024), dtype=np.float32, order="F")
    res = ctypes.c_double(0); cr = np.zeros(1024)
    assert L.swd_lhdisppick_noalloc(P(ml), 2048, 1024, 100.0, P(np.zeros(1024)), 1, 10, ctypes.byref(res), P(cr)) == -2
    assert L.swd_lhdisppick(P(ml), 4096, 10, 100.0, P(np.zeros(10)), 1, 10, ctypes.byref(res), P(cr)) == -2
    mm = np.ones((100, 10), dtype=np.float32, order="F"); m1 = np.ones((50, 10), dtype=np.float32, order="F")
    g = np.zeros(10); out = np.zeros((600, 20), dtype=np.float32, order="F")
    assert L.swd_adwdgrad1(600, 20, 20, 10, 5, 0.0, 0, 1, 10.0, P(g), P(g), P(mm), P(m1), 5, 10, P(mm), P(m1), 5, 5,
                           50, 100.0, 0.5, 0.5, 11, 500, P(out)) == -2
    n = ctypes.c_int(0)
    prob = dict(nz=30, nx=60, nbc=10, fd_order=22, nt=400, in_wf=5, dx=0.5, dt=1e-4, sx=np.array([1, 30]),
                sz=np.array([1, 1]), ng=60, gx0=1, dg=1, gz=1, df=0.5, fmin=5.0, fmax=80.0, vmin=100.0, vmax=800.0,
                ds=0.5, offset=60, m=0, M=1, ini=10, w=0.0, beta=10.0, ns=0)
    from swd_amd.engine import problem_npair
    assert problem_npair(prob) == 151
    print("HOST_SANITIZED_OK", flush=True)
""")


@pytest.mark.skipif(not os.path.exists(LIBASAN), reason="gcc's libasan is not installed")
def test_host_entry_points_under_asan_and_ubsan():
    from swd_amd import build as B
    lib = B.build_sanitized(verbose=False)
    env = dict(os.environ, SWD_LIBRARY=lib, LD_PRELOAD=LIBASAN,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:protect_shadow_gap=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-c", f"ROOT = {ROOT!r}\n" + CHILD], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0 and "HOST_SANITIZED_OK" in r.stdout, (r.stdout[-3000:], r.stderr[-6000:])
