"""GPU parity of the FD kernels against the FP64 CPU oracle.

FP32 GPU results are judged by relative L2 error against the FP64 oracle
(tolerance 1e-4, the stated bar for gathers and gradients); FP64 GPU results
against the same oracle at 1e-10.  Both the fused march kernels and the plain
per-statement kernels are exercised.
"""
import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import fd

pytestmark = pytest.mark.gpu

TOL32 = 1e-4
TOL64 = 1e-10


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def cfg0():
    """Config 0 geometry (82 x 240 padded, shot at x = 2 m) with nt cut to 1500."""
    nz, nx, nbc, dx, dt, nt = 60, 200, 20, 0.5, 1e-4, 1500
    vs = np.tile(np.linspace(200, 600, nz)[:, None], (1, nx))
    vp = 1.73 * vs
    s = O.ricker(30, dt, nt)
    gx = np.arange(1, 200, dtype=float)
    gz = np.ones(199)
    args = (1, nbc, nt, dt / dx, dx, dt, 5, 1, gx, gz, s, vp, vs, 1, 22, "w", 0, 5)
    ref = O.staggerfd_eigend(*args)
    return dict(args=args, ref=ref, nt=nt, vp=vp, vs=vs, gx=gx, gz=gz, s=s, nbc=nbc, dx=dx, dt=dt)


@pytest.mark.parametrize("path", [0, 1], ids=["fused", "plain"])
def test_forward_fp32_vs_fp64_oracle(cfg0, path):
    old = fd.set_fd_path(path)
    try:
        su, sw, S = fd.staggerfd_eigen(*cfg0["args"])
    finally:
        fd.set_fd_path(old)
    rsu, rsw, rS = cfg0["ref"]
    e = {"seismo_w": rel(sw, rsw), "seismo_u": rel(su, rsu)}
    for k in rS:
        e[k] = rel(S[k], rS[k])
    print("forward fp32 rel-L2 vs fp64:", e)
    assert max(e.values()) < TOL32, e
    np.testing.assert_array_equal(sw[:, 4], np.float32(cfg0["s"]))


@pytest.mark.parametrize("path", [0, 1], ids=["fused", "plain"])
def test_forward_fp64_matches_oracle(cfg0, path):
    old = fd.set_fd_path(path)
    try:
        su, sw, S = fd.staggerfd_eigend(*cfg0["args"])
    finally:
        fd.set_fd_path(old)
    rsu, rsw, rS = cfg0["ref"]
    assert rel(sw, rsw) < TOL64 and rel(su, rsu) < TOL64
    for k in rS:
        assert rel(S[k], rS[k]) < TOL64, k


@pytest.mark.parametrize("path", [0, 1], ids=["fused", "plain"])
def test_adjoint_vs_oracle(cfg0, path):
    c = cfg0
    r = np.random.default_rng(7)
    adj = r.standard_normal((c["nt"], 199)) * 1e-3
    _, _, S = c["ref"]
    a = (S, adj, 1, c["nbc"], c["nt"], c["dt"] / c["dx"], c["dx"], c["dt"], c["gx"], c["gz"], c["s"], c["vp"],
         c["vs"], 1, 22, 0, 5)
    ref = O.e2drtm_eigend(*a)
    old = fd.set_fd_path(path)
    try:
        g32 = fd.e2drtm_eigen(*a)
        g64 = fd.e2drtm_eigend(*a)
    finally:
        fd.set_fd_path(old)
    e32 = [rel(x, y) for x, y in zip(g32, ref)]
    e64 = [rel(x, y) for x, y in zip(g64, ref)]
    print("adjoint fp32 rel-L2 (cl, cm, illum):", e32, "fp64:", e64)
    assert max(e32) < TOL32 and max(e64) < TOL64


@pytest.mark.parametrize("stype", [1, 2, 3, 4])
def test_forward_other_source_types(stype):
    nz, nx, nbc = 20, 30, 8
    r = np.random.default_rng(stype)
    vs = r.uniform(250, 400, (nz, nx))
    vp = 1.8 * vs
    c = O.prep(vp, vs, nbc, 0.5, 1e-4, 22)
    nzbc, nxbc = nz + 2 + nbc, nx + 2 * nbc
    s = r.standard_normal(200)
    iv = O.fwd_iv(200, nzbc, nxbc, 1e-4 / 0.5, 30, 6, 20, 3, 9, 1, stype, 22, 5, nz, nx)
    ref = O.staggerfd(iv, c, s)
    for path in (0, 1):
        old = fd.set_fd_path(path)
        try:
            su, sw, S = fd.eigen_staggerfd_double(iv, *[c[k] for k in O.COEF_ORDER], s)
        finally:
            fd.set_fd_path(old)
        assert rel(sw, ref[1]) < TOL64 and rel(su, ref[0]) < TOL64
        assert rel(S["bwz"], ref[2]["bwz"]) < TOL64


@pytest.mark.parametrize("order", [24, 26, 28])
def test_higher_orders_plain_path(order):
    nz, nx, nbc = 18, 26, 8
    r = np.random.default_rng(order)
    vs = r.uniform(250, 400, (nz, nx))
    vp = 1.8 * vs
    c = O.prep(vp, vs, nbc, 0.5, 5e-5, order)
    pt = (order - 20) // 2 + 1
    nzbc, nxbc = nz + pt + nbc, nx + 2 * nbc
    s = O.ricker(150, 5e-5, 300)
    iv = O.fwd_iv(300, nzbc, nxbc, 1e-4, 26, pt + 1, 20, pt + 1, 9, 1, 5, order, 5, nz, nx)
    ref = O.staggerfd(iv, c, s)
    su, sw, S = fd.eigen_staggerfd_double(iv, *[c[k] for k in O.COEF_ORDER], s)
    assert rel(sw, ref[1]) < TOL64 and rel(S["fux"], ref[2]["fux"]) < TOL64
    r2 = np.random.default_rng(1)
    adj = r2.standard_normal((300, 26))
    ivb = O.adj_iv(300, nzbc, nxbc, 1e-4, 26, nbc, pt + 1, 9, 1, order, 5, nz, nx, 5e-5)
    refb = O.e2drtm(ivb, c, adj, ref[2])
    gb = fd.eigen_e2drtm_double(ivb, *[c[k] for k in O.COEF_ORDER], adj, *[ref[2][k] for k in ("fux", "fuz", "bwx", "bwz")])
    assert max(rel(x, y) for x, y in zip(gb, refb)) < TOL64


def test_source_on_ring_uses_general_path():
    nz, nx, nbc = 16, 20, 6
    r = np.random.default_rng(3)
    vs = r.uniform(250, 400, (nz, nx))
    c = O.prep(1.8 * vs, vs, nbc, 0.5, 1e-4, 22)
    nzbc, nxbc = nz + 2 + nbc, nx + 2 * nbc
    s = r.standard_normal(100)
    iv = O.fwd_iv(100, nzbc, nxbc, 2e-4, 10, 2, 1, 2, 1, 2, 5, 22, 5, nz, nx)   # source on the ring column 0
    ref = O.staggerfd(iv, c, s)
    su, sw, S = fd.eigen_staggerfd_double(iv, *[c[k] for k in O.COEF_ORDER], s)
    assert rel(sw, ref[1]) < TOL64 and rel(S["bwz"], ref[2]["bwz"]) < TOL64
