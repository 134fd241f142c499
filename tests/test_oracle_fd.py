"""Pins the FD part of the CPU oracle (oracle/swd_oracle.c) without the reference.

Closed forms, invariants and physics that a dropped term, a wrong sign, a
swapped index or a mis-staggered difference would break.  Reference behaviour
being pinned: core/eigen_staggerfd_single.cpp:207-357, core/eigen_e2drtm_single.cpp:181-283,
core/staggerfd_eigen.m:37-107, core/damp_circle.m:5-67, core/ricker.m:18-35.
"""
import math

import numpy as np
import pytest

from oracle import swd_oracle as O


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


# ---------------------------------------------------------------- prep
def test_prep_tiny_hand_values():
    # nz=2, nx=3, nbc=3, fd_order=22 -> MATLAB pad_top=2, nzbc=7, nxbc=9
    vp = np.array([[1000.0, 1100.0, 1200.0], [1300.0, 1400.0, 1500.0]])
    vs = vp / 2.0
    dx, dt = 1.0, 1e-4
    c = O.prep(vp, vs, 3, dx, dt, 22)
    assert c["ca"].shape == (7, 9)
    dtx = dt / dx
    # interior (MATLAB rows 3..4, cols 4..6) holds calparam values
    ca0 = vp ** 2
    cm0 = vs ** 2
    np.testing.assert_allclose(c["ca"][2:4, 3:6], ca0)
    np.testing.assert_allclose(c["cm"][2:4, 3:6], cm0)
    np.testing.assert_allclose(c["cl"][2:4, 3:6], ca0 - 2 * cm0)
    # edge replicate: left pad copies column 1, bottom pad copies the last row
    np.testing.assert_allclose(c["cm"][2:4, 0], cm0[:, 0])
    np.testing.assert_allclose(c["cm"][6, 3:6], cm0[-1])
    np.testing.assert_allclose(c["cm"][6, 8], cm0[-1, -1])
    # top pad row 1 (MATLAB) copies the first row; row 2 (= pad_top) is the free surface row
    np.testing.assert_allclose(c["cm"][0, 3:6], cm0[0])
    np.testing.assert_allclose(c["ca"][1, :], 2 * c["cm"][1, :])
    np.testing.assert_allclose(c["cl"][1, :], 0.0)
    np.testing.assert_allclose(c["ca"][0, 3:6], ca0[0])     # row above keeps the padded value
    np.testing.assert_allclose(c["b"][1, :], 2 * dtx)
    np.testing.assert_allclose(c["b"][0, :], dtx)
    np.testing.assert_allclose(c["b"][2:, :], dtx)
    np.testing.assert_array_equal(c["b1"], c["b"])
    np.testing.assert_array_equal(c["cm1"], c["cm"])
    # damp_circle: a = 2, kappa = 3*vmin*ln(1e7)/4, damp1d = kappa*[0, 1/4, 1]
    vmin = 500.0
    kap = 3 * vmin * math.log(1e7) / 4.0
    damp = (1 - c["temp"]) / dt
    np.testing.assert_allclose(damp[2, 0:3], kap * np.array([1.0, 0.25, 0.0]), rtol=1e-12)   # left, reversed
    np.testing.assert_allclose(damp[3, 6:9], kap * np.array([0.0, 0.25, 1.0]), rtol=1e-12)   # right
    np.testing.assert_allclose(damp[4:7, 4], kap * np.array([0.0, 0.25, 1.0]), rtol=1e-12)   # bottom
    np.testing.assert_allclose(damp[0:2, :], 0.0, atol=1e-9)                                  # free-surface top
    np.testing.assert_allclose(damp[2:4, 3:6], 0.0, atol=1e-9)                                # interior
    # lower-left corner: MATLAB (iz, ix) = (5, 3): dist^2 = (3-4)^2 + (5-4)^2 = 2
    np.testing.assert_allclose(damp[4, 2], kap * 2 / 9, rtol=1e-12)
    # (7, 1): dist^2 = 9 + 9
    np.testing.assert_allclose(damp[6, 0], kap * 18 / 9, rtol=1e-12)
    # lower-right (6, 8): dist^2 = (8-3-3)^2 + (6-2-2)^2 = 4 + 4
    np.testing.assert_allclose(damp[5, 7], kap * 8 / 9, rtol=1e-12)


def test_damp_circle_without_free_surface_fills_upper_corners():
    d = O.damp_circle(100.0, 10, 10, 3, 1.0, 0, 3)
    kap = 3 * 100.0 * math.log(1e7) / 4.0
    # upper-left (1,1): dist^2 = (1-4)^2 + (1-4)^2 = 18
    np.testing.assert_allclose(d[0, 0], kap * 18 / 9, rtol=1e-12)
    # top rows of interior columns get the reversed ramp
    np.testing.assert_allclose(d[0:3, 4], kap * np.array([1.0, 0.25, 0.0]), rtol=1e-12)


def test_ricker_closed_form():
    f, dt = 30.0, 1e-4
    w = O.ricker(f, dt, 3000)
    nw = 2 * math.floor(2.2 / f / dt / 2) + 1
    nc = nw // 2
    assert np.all(w[nw:] == 0) and w[nw - 1] != 0
    assert w[nc] == 1.0                                   # alpha = 0 at k = nc+1
    np.testing.assert_allclose(w[:nc][::-1], w[nc + 1:nw], rtol=1e-12)
    k = 7
    beta = (k * f * dt * math.pi) ** 2
    np.testing.assert_allclose(w[nc + k], (1 - 2 * beta) * math.exp(-beta), rtol=1e-12)
    with pytest.raises(ValueError):
        O.ricker(f, dt, nw - 1)


# ---------------------------------------------------------------- forward
def _rand_coefs(nzbc, nxbc, seed=1):
    r = np.random.default_rng(seed)
    c = {k: np.asfortranarray(r.uniform(0.5, 1.5, (nzbc, nxbc))) for k in O.COEF_ORDER}
    return c


def test_forward_two_step_closed_form():
    nzbc, nxbc, nz, nx = 12, 14, 6, 8      # nbc = 3, pad_top (C) = 1
    c = _rand_coefs(nzbc, nxbc)
    dtx = 0.37
    sz, sx = 5, 6                          # 0-based source cell
    s = np.zeros(2)
    s[0] = 1.0
    ng = nxbc - 2
    out = {}
    for row in (sz - 1, sz, sz + 1):
        iv = O.fwd_iv(2, nzbc, nxbc, dtx, ng, sz + 1, sx + 1, row + 1, 2, 1, 5, 22, 1, nz, nx)
        out[row] = O.staggerfd(iv, c, s)
    su, sw, sn = out[sz]
    A = lambda k, i, j: c[k][i, j]
    # step 1: the hard source writes 1 at (sz, sx); traces start at column 1
    assert sw[0, sx - 1] == 1.0 and np.count_nonzero(sw[0]) == 1 and not np.any(su[0])
    # snapshot slot 0 = strains after step 1 (interior origin at (2, 3))
    iz, ix = sz - 2, sx - 3
    np.testing.assert_allclose(sn["bwz"][iz, ix, 0], 1.0)
    np.testing.assert_allclose(sn["bwz"][iz + 1, ix, 0], -1.0)
    np.testing.assert_allclose(sn["bwx"][iz, ix, 0], -1.0)
    np.testing.assert_allclose(sn["bwx"][iz, ix - 1, 0], 1.0)
    assert not np.any(sn["fux"][:, :, 0]) and not np.any(sn["fuz"][:, :, 0])
    assert np.count_nonzero(sn["bwz"][:, :, 0]) == 2 and np.count_nonzero(sn["bwx"][:, :, 0]) == 2
    # step 2 (s[1] = 0): velocities from the step-1 stresses
    u_s = -A("b", sz, sx) * dtx * (A("cl", sz, sx) + A("cm1", sz, sx))
    u_w = A("b", sz, sx - 1) * dtx * (A("cl", sz, sx) + A("cm1", sz, sx - 1))
    w_e = A("b1", sz, sx + 1) * dtx * A("cm1", sz, sx)
    np.testing.assert_allclose(su[1, sx - 1], u_s, rtol=1e-13)
    np.testing.assert_allclose(su[1, sx - 2], u_w, rtol=1e-13)
    np.testing.assert_allclose(sw[1, sx], w_e, rtol=1e-13)
    assert sw[1, sx - 1] == 0.0                                  # overwritten by s[1]
    w_s = A("b1", sz + 1, sx) * dtx * A("ca", sz + 1, sx)
    w_n = A("b1", sz - 1, sx) * dtx * A("ca", sz, sx)
    np.testing.assert_allclose(out[sz + 1][1][1, sx - 1], w_s, rtol=1e-13)
    np.testing.assert_allclose(out[sz - 1][1][1, sx - 1], w_n, rtol=1e-13)
    # u on the row below: uu(sz+1, sx) = b*(xx(sz+1,sx+1)-xx(sz+1,sx) + xz(sz+1,sx)-xz(sz,sx))
    u_sb = A("b", sz + 1, sx) * (dtx * A("cl", sz + 1, sx) + dtx * A("cm1", sz, sx))
    np.testing.assert_allclose(out[sz + 1][0][1, sx - 1], u_sb, rtol=1e-13)


@pytest.fixture(scope="module")
def cfg0():
    nz, nx, nbc, dx, dt, nt = 60, 200, 20, 0.5, 1e-4, 1200
    vs = np.tile(np.linspace(200, 600, nz)[:, None], (1, nx))
    vp = 1.73 * vs
    s = O.ricker(30, dt, nt)
    gx = np.arange(1, 200, dtype=float)
    gz = np.ones(199)
    return dict(nz=nz, nx=nx, nbc=nbc, dx=dx, dt=dt, nt=nt, vs=vs, vp=vp, s=s, gx=gx, gz=gz)


def _fwd(c, s=None, vs=None, sx=5, **kw):
    return O.staggerfd_eigend(1, c["nbc"], c["nt"], c["dt"] / c["dx"], c["dx"], c["dt"], sx, 1, c["gx"], c["gz"],
                              c["s"] if s is None else s, c["vp"], c["vs"] if vs is None else vs, 1, 22, "w", 0, 5,
                              **kw)


def test_forward_receiver_on_source_and_zero_source(cfg0):
    su, sw, sn = _fwd(cfg0)
    np.testing.assert_array_equal(sw[:, 4], cfg0["s"])        # trace 5 sits on the source cell
    su0, sw0, sn0 = _fwd(cfg0, s=np.zeros(cfg0["nt"]))
    assert not np.any(sw0) and not np.any(su0) and not np.any(sn0["fux"])


def test_forward_linear_in_source(cfg0):
    r = np.random.default_rng(3)
    s1, s2 = r.standard_normal(cfg0["nt"]), cfg0["s"]
    _, a, sa = _fwd(cfg0, s=s1)
    _, b, sb = _fwd(cfg0, s=s2)
    _, c, sc = _fwd(cfg0, s=2 * s1 - 3 * s2)
    assert rel(c, 2 * a - 3 * b) < 1e-12
    assert rel(sc["bwz"], 2 * sa["bwz"] - 3 * sb["bwz"]) < 1e-12


def test_forward_density_invariance(cfg0):
    c = cfg0
    nzbc, nxbc = c["nz"] + 2 + c["nbc"], c["nx"] + 2 * c["nbc"]
    iv = O.fwd_iv(c["nt"], nzbc, nxbc, c["dt"] / c["dx"], 199, 3, 25, 3, 21, 1, 5, 22, 5, c["nz"], c["nx"])
    k1 = O.prep(c["vp"], c["vs"], c["nbc"], c["dx"], c["dt"], 22, den=1.0)
    k2 = O.prep(c["vp"], c["vs"], c["nbc"], c["dx"], c["dt"], 22, den=2.5)
    _, a, sa = O.staggerfd(iv, k1, c["s"])
    _, b, sb = O.staggerfd(iv, k2, c["s"])
    assert rel(b, a) < 1e-12 and rel(sb["fux"], sa["fux"]) < 1e-12


def test_forward_fp32_twin_tracks_fp64(cfg0):
    _, a, _ = _fwd(cfg0)
    _, b, _ = _fwd(cfg0, precision="f")
    assert rel(b, a) < 1e-5


def _rayleigh_ratio(kappa):
    # (2 - x^2)^2 = 4 sqrt(1 - x^2/kappa^2) sqrt(1 - x^2), bisection on (0.5, 0.99)
    f = lambda x: (2 - x * x) ** 2 - 4 * math.sqrt(1 - x * x / kappa ** 2) * math.sqrt(1 - x * x)
    lo, hi = 0.5, 0.99
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


@pytest.mark.slow
def test_forward_free_surface_gives_rayleigh_velocity():
    """Homogeneous half-space: the Radon ridge sits at c_R = 0.9193 Vs (free surface + staggering)."""
    assert abs(_rayleigh_ratio(1.73) - 0.91926) < 2e-4
    nz, nx, nbc, dx, dt, nt = 80, 300, 30, 0.5, 1e-4, 3000
    vs = np.full((nz, nx), 300.0)
    vp = 1.73 * vs
    s = O.ricker(25, dt, nt)
    gx = np.arange(1, nx + 1, dtype=float)
    gz = np.ones(nx)
    _, sw, _ = O.staggerfd_eigend(1, nbc, nt, dt / dx, dx, dt, 1, 1, gx, gz, s, vp, vs, 1, 22, "w", 0, 5,
                                  snapshots=False)
    ml, L, _ = O.rt_side(sw, "r", 1, 0.5, dt, 451, 100, 550, 15, 25, nt, nx, dx, 240, 20, 1)
    cr = 100 + np.argmax(ml, axis=0)
    c_r = _rayleigh_ratio(1.73) * 300.0
    # 15-25 Hz is 22-37 cells per wavelength; without the free surface the ridge
    # would sit near Vs = 300 or Vp = 519, not at 0.919 Vs
    assert abs(np.median(cr) - c_r) / c_r < 0.02
    assert np.all(np.abs(cr - c_r) / c_r < 0.035)


# ---------------------------------------------------------------- adjoint
def _adj_setup(nt=6, in_wf=2, seed=5):
    nzbc, nxbc, nz, nx, nbc = 12, 14, 6, 8, 3
    c = _rand_coefs(nzbc, nxbc, seed)
    r = np.random.default_rng(seed)
    nrec = nt // in_wf
    snaps = {k: np.asfortranarray(r.standard_normal((nz, nx, nrec))) for k in ("fux", "fuz", "bwx", "bwz")}
    return c, snaps, dict(nzbc=nzbc, nxbc=nxbc, nz=nz, nx=nx, nbc=nbc, nt=nt, in_wf=in_wf)


def test_adjoint_first_imaging_step_closed_form():
    c, snaps, g = _adj_setup()
    nt, nz, nx = g["nt"], g["nz"], g["nx"]
    for k in snaps:                       # only slot (nt-2)/in_wf = 2 is non-zero
        snaps[k][:, :, :2] = 0
    dtx, dt = 0.41, 0.013
    gz, gx0, ng = 4, 2, 10                # 0-based receiver row; traces at columns 2..11
    A = np.zeros((nt, ng))
    g0, amp = 3, 1.7
    A[nt - 2, g0] = amp
    iv = O.adj_iv(nt, g["nzbc"], g["nxbc"], dtx, ng, g["nbc"], gz + 1, gx0 + 1, 1, 22, g["in_wf"], nz, nx, dt)
    cl, cm, il = O.e2drtm(iv, c, A, snaps)
    xc = gx0 + g0                         # column of the injected receiver
    XX = np.zeros((g["nzbc"], g["nxbc"]))
    ZZ = np.zeros_like(XX)
    XZ = np.zeros_like(XX)
    q = dtx * dt * amp
    ZZ[gz, xc], ZZ[gz + 1, xc] = -q, q
    XZ[gz, xc], XZ[gz, xc + 1] = -q, q
    sl = (slice(2, 2 + nz), slice(g["nbc"], g["nbc"] + nx))
    X, Z, XZi = XX[sl], ZZ[sl], XZ[sl]
    F = {k: snaps[k][:, :, 2] for k in snaps}
    ecl = (F["fux"] + F["bwz"]) * (X + Z)
    ecm = 2 * (X * F["fux"] + Z * F["bwz"] + 0.5 * XZi * (F["fuz"] + F["bwx"]))
    np.testing.assert_allclose(cl, ecl, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(cm, ecm, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(il, (F["fux"] + F["bwz"]) ** 2, rtol=1e-12)


def test_adjoint_illum_linear_and_bilinear():
    c, snaps, g = _adj_setup(nt=40, in_wf=5, seed=9)
    r = np.random.default_rng(11)
    ng = 10
    A1, A2 = r.standard_normal((g["nt"], ng)), r.standard_normal((g["nt"], ng))
    iv = O.adj_iv(g["nt"], g["nzbc"], g["nxbc"], 0.3, ng, g["nbc"], 3, 3, 1, 22, g["in_wf"], g["nz"], g["nx"], 0.01)
    a = O.e2drtm(iv, c, A1, snaps)
    b = O.e2drtm(iv, c, A2, snaps)
    m = O.e2drtm(iv, c, 2 * A1 - 0.5 * A2, snaps)
    for k in range(2):
        assert rel(m[k], 2 * a[k] - 0.5 * b[k]) < 1e-12
    np.testing.assert_allclose(a[2], np.sum((snaps["fux"] + snaps["bwz"]) ** 2, axis=2), rtol=1e-12)
    sc = {k: 3.0 * v for k, v in snaps.items()}
    s3 = O.e2drtm(iv, c, A1, sc)
    assert rel(s3[0], 3 * a[0]) < 1e-12 and rel(s3[1], 3 * a[1]) < 1e-12


def test_adjoint_last_row_never_injected():
    c, snaps, g = _adj_setup(nt=10, in_wf=5)
    A = np.zeros((10, 10))
    A[9, :] = 1.0                         # row nt-1 is never read (loop starts at nt-2)
    iv = O.adj_iv(10, g["nzbc"], g["nxbc"], 0.3, 10, g["nbc"], 3, 3, 1, 22, 5, g["nz"], g["nx"], 0.01)
    cl, cm, _ = O.e2drtm(iv, c, A, snaps)
    assert not np.any(cl) and not np.any(cm)


def test_rejects_undefined_inputs():
    c = _rand_coefs(12, 14)
    with pytest.raises(ValueError):      # nt not a multiple of in_wf
        O.staggerfd(O.fwd_iv(7, 12, 14, 0.3, 4, 5, 6, 3, 3, 1, 5, 22, 2, 6, 8), c, np.zeros(7))
    with pytest.raises(ValueError):      # receivers past the grid
        O.staggerfd(O.fwd_iv(4, 12, 14, 0.3, 20, 5, 6, 3, 3, 1, 5, 22, 2, 6, 8), c, np.zeros(4))
    with pytest.raises(ValueError):      # unknown fd_order
        O.staggerfd(O.fwd_iv(4, 12, 14, 0.3, 4, 5, 6, 3, 3, 1, 5, 23, 2, 6, 8), c, np.zeros(4))


@pytest.mark.parametrize("order", [24, 26, 28])
def test_higher_orders_linear_and_zero(order):
    nzbc, nxbc = 20, 24
    c = _rand_coefs(nzbc, nxbc, 2)
    for k in ("b", "b1"):
        c[k] *= 0.1
    r = np.random.default_rng(order)
    s1, s2 = r.standard_normal(20), r.standard_normal(20)
    pt = (order - 20) // 2
    iv = O.fwd_iv(20, nzbc, nxbc, 0.05, 6, pt + 4, 9, pt + 2, 7, 1, 5, order, 5, 8, 10)
    _, a, sa = O.staggerfd(iv, c, s1)
    _, b, sb = O.staggerfd(iv, c, s2)
    _, m, sm = O.staggerfd(iv, c, s1 + 2 * s2)
    assert rel(m, a + 2 * b) < 1e-12
    _, z, _ = O.staggerfd(iv, c, np.zeros(20))
    assert not np.any(z)


# ---------------------------------------------------------------- adjoint: the velocity half-step
# The imaging condition reads the adjoint stresses back: with unit snapshots in
# one slot and zeros elsewhere, cm_img = 2 XX (fux = 1), 2 ZZ (bwz = 1) or XZ
# (fuz = 1), and cl_img = XX + ZZ (fux = 1), where XX, ZZ, XZ are the adjoint
# stresses of that slot's step over the interior.
S41, S42 = float(np.float32(1.1250)), float(np.float32(-0.0416666667))


def _adj_stress_readout(c, iv, A, nrec, slot, nz, nx):
    out = {}
    for name in ("fux", "bwz", "fuz"):
        sn = {k: np.zeros((nz, nx, nrec), order="F") for k in ("fux", "fuz", "bwx", "bwz")}
        sn[name][:, :, slot] = 1.0
        out[name] = O.e2drtm(iv, c, A, sn)
    X = out["fux"][1] / 2
    Z = out["bwz"][1] / 2
    np.testing.assert_allclose(out["fux"][0], X + Z, rtol=1e-12, atol=1e-300)
    return X, Z, out["fuz"][1]


def _adj_stress_step22(c, XX, ZZ, XZ, UU, WW, dtx, pt):
    """Adjoint stress update (e2drtm cpp:245-248) and free surface (cpp:270), order 22."""
    t = c["temp"]
    B = (slice(1, -1), slice(1, -1))
    xx, zz, xz = XX.copy(), ZZ.copy(), XZ.copy()
    xx[B] = t[B] * XX[B] - dtx * (UU[1:-1, 2:] - UU[1:-1, 1:-1])
    zz[B] = t[B] * ZZ[B] - dtx * (WW[1:-1, 1:-1] - WW[:-2, 1:-1])
    xz[B] = t[B] * XZ[B] - dtx * (UU[2:, 1:-1] - UU[1:-1, 1:-1]) - dtx * (WW[1:-1, 1:-1] - WW[1:-1, :-2])
    zz[pt, :] = 0
    return xx, zz, xz


@pytest.mark.parametrize("gz", [4, 2])
def test_adjoint_two_step_closed_form(gz):
    """Two adjoint steps after one adjoint impulse (eigen_e2drtm_single.cpp:182-187, 238-248, 270).

    Step nt-2 turns the impulse into stresses; step nt-3 forms uu_b / ww_b from
    them with the coefficients inside the differences, hand-derived cell by
    cell below, then the stresses read back through the imaging of slot nt-3.
    gz = 2 is the production receiver row, one below the zeroed zz row pt = 1.
    """
    nzbc, nxbc, nz, nx, nbc, pt = 12, 14, 6, 8, 3, 1
    c = _rand_coefs(nzbc, nxbc, 17)
    nt, in_wf, dtx, dt, amp = 4, 1, 0.41, 0.013, 1.7
    gx0, ng, g0 = 2, 10, 4
    xc = gx0 + g0
    A = np.zeros((nt, ng))
    A[nt - 2, g0] = amp
    iv = O.adj_iv(nt, nzbc, nxbc, dtx, ng, nbc, gz + 1, gx0 + 1, 1, 22, in_wf, nz, nx, dt)
    X, Z, XZr = _adj_stress_readout(c, iv, A, nt, nt - 3, nz, nx)
    k = lambda name, i, j: c[name][i, j]
    p, q = dt * amp, dtx * dt * amp
    # after step nt-2: the impulse in ww_b, and the stresses it makes
    Z1 = np.zeros((nzbc, nxbc))
    XZ1 = np.zeros_like(Z1)
    X1 = np.zeros_like(Z1)
    Z1[gz, xc], Z1[gz + 1, xc] = -q, q
    XZ1[gz, xc], XZ1[gz, xc + 1] = -q, q
    # step nt-3, velocity: uu_b = -b (sum of (coef * stress) differences), by hand
    U2 = np.zeros_like(Z1)
    W2 = np.zeros_like(Z1)
    U2[gz, xc] = k("b", gz, xc) * q * (k("cl", gz, xc) + k("cm1", gz, xc))
    U2[gz, xc + 1] = -k("b", gz, xc + 1) * q * (k("cl", gz, xc) + k("cm1", gz, xc + 1))
    U2[gz + 1, xc] = -k("b", gz + 1, xc) * q * (k("cl", gz + 1, xc) + k("cm1", gz, xc))
    U2[gz + 1, xc + 1] = k("b", gz + 1, xc + 1) * q * (k("cl", gz + 1, xc) + k("cm1", gz, xc + 1))
    W2[gz - 1, xc] = k("b1", gz - 1, xc) * k("ca", gz, xc) * q
    W2[gz, xc] = (k("temp", gz, xc) * p - k("b1", gz, xc) * q *
                  (k("ca", gz, xc) + k("ca", gz + 1, xc) + k("cm1", gz, xc) + k("cm1", gz, xc + 1)))
    W2[gz + 1, xc] = k("b1", gz + 1, xc) * k("ca", gz + 1, xc) * q
    W2[gz, xc - 1] = k("b1", gz, xc - 1) * k("cm1", gz, xc) * q
    W2[gz, xc + 1] = k("b1", gz, xc + 1) * k("cm1", gz, xc + 1) * q
    xx, zz, xz = _adj_stress_step22(c, X1, Z1, XZ1, U2, W2, dtx, pt)
    sl = (slice(pt + 1, pt + 1 + nz), slice(nbc, nbc + nx))
    np.testing.assert_allclose(X, xx[sl], rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(Z, zz[sl], rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(XZr, xz[sl], rtol=1e-12, atol=1e-20)
    assert np.count_nonzero(X) >= 4 and np.count_nonzero(XZr) >= 6


def _scatter_velocity(c, XX, ZZ, XZ, UU, WW, S, pt):
    """Adjoint velocity half-step (e2drtm cpp:182-187 for S = (1,), cpp:188-197
    for S = (S41, S42)) as a scatter: each nonzero coefficient-weighted stress
    adds +-S_q to the velocity cells whose differences read it."""
    n0, n1 = ZZ.shape
    au = np.zeros((n0, n1))
    aw = np.zeros((n0, n1))
    for qq, Sq in enumerate(S, start=1):
        for (r, s) in zip(*np.nonzero(XX)):
            v = c["ca"][r, s] * XX[r, s]            # uu_b: (ca xx)[i, j+q-1] - (ca xx)[i, j-q]
            au[r, s - qq + 1] += Sq * v
            au[r, s + qq] -= Sq * v
            v = c["cl"][r, s] * XX[r, s]            # ww_b: (cl xx)[i+q, j] - (cl xx)[i-q+1, j]
            aw[r - qq, s] += Sq * v
            aw[r + qq - 1, s] -= Sq * v
        for (r, s) in zip(*np.nonzero(ZZ)):
            v = c["cl"][r, s] * ZZ[r, s]            # uu_b: (cl zz)[i, j+q-1] - (cl zz)[i, j-q]
            au[r, s - qq + 1] += Sq * v
            au[r, s + qq] -= Sq * v
            v = c["ca"][r, s] * ZZ[r, s]            # ww_b: (ca zz)[i+q, j] - (ca zz)[i-q+1, j]
            aw[r - qq, s] += Sq * v
            aw[r + qq - 1, s] -= Sq * v
        for (r, s) in zip(*np.nonzero(XZ)):
            v = c["cm1"][r, s] * XZ[r, s]           # uu_b: (cm1 xz)[i+q-1, j] - (cm1 xz)[i-q, j]
            au[r - qq + 1, s] += Sq * v
            au[r + qq, s] -= Sq * v
            aw[r, s - qq] += Sq * v                 # ww_b: (cm1 xz)[i, j+q] - (cm1 xz)[i, j-q+1]
            aw[r, s + qq - 1] -= Sq * v
    B = (slice(pt, n0 - pt), slice(pt, n1 - pt))
    U = np.zeros((n0, n1))
    W = np.zeros((n0, n1))
    U[B] = c["temp"][B] * UU[B] - c["b"][B] * au[B]
    W[B] = c["temp"][B] * WW[B] - c["b1"][B] * aw[B]
    return U, W


def _adj_stress_step(c, XX, ZZ, XZ, U, W, dtx, S, pt):
    """Adjoint stress update on the block (e2drtm cpp:245-268) and zz_b row pt = 0 (cpp:270)."""
    n0, n1 = XX.shape
    r0, r1, c0, c1 = pt, n0 - pt, pt, n1 - pt
    B = (slice(r0, r1), slice(c0, c1))
    sh = lambda a, di, dj: a[r0 + di:r1 + di, c0 + dj:c1 + dj]
    t = c["temp"]
    dx_ = sum(Sq * (sh(U, 0, q) - sh(U, 0, 1 - q)) for q, Sq in enumerate(S, start=1))
    dz_ = sum(Sq * (sh(W, q - 1, 0) - sh(W, -q, 0)) for q, Sq in enumerate(S, start=1))
    du_ = sum(Sq * (sh(U, q, 0) - sh(U, 1 - q, 0)) for q, Sq in enumerate(S, start=1))
    dw_ = sum(Sq * (sh(W, 0, q - 1) - sh(W, 0, -q)) for q, Sq in enumerate(S, start=1))
    xx, zz, xz = XX.copy(), ZZ.copy(), XZ.copy()
    xx[B] = t[B] * XX[B] - dtx * dx_
    zz[B] = t[B] * ZZ[B] - dtx * dz_
    xz[B] = t[B] * XZ[B] - dtx * du_ - dtx * dw_
    zz[pt, :] = 0
    return xx, zz, xz


@pytest.mark.parametrize("order", [22, 24])
def test_adjoint_three_steps_scatter_form(order):
    """Three adjoint steps after one impulse, so that xx_b is nonzero when the
    third velocity half-step reads it (the ca / cl weights of xx_b in uu_b and
    ww_b); velocities in the scatter form, stresses read back through the
    imaging of slot nt-4."""
    pt = (order - 20) // 2
    S = (1.0,) if order == 22 else (S41, S42)
    nzbc, nxbc, nz, nx = 16, 18, 8, 10
    nbc = (nxbc - nx) // 2
    c = _rand_coefs(nzbc, nxbc, 31 + order)
    nt, dtx, dt, amp = 5, 0.33, 0.017, 0.9
    gz, gx0, ng, g0 = pt + 4, pt + 1, 12, 6
    xc = gx0 + g0
    A = np.zeros((nt, ng))
    A[nt - 2, g0] = amp
    iv = O.adj_iv(nt, nzbc, nxbc, dtx, ng, nbc, gz + 1, gx0 + 1, 1, order, 1, nz, nx, dt)
    X, Z, XZr = _adj_stress_readout(c, iv, A, nt, nt - 4, nz, nx)
    zero = np.zeros((nzbc, nxbc))
    U, W = zero.copy(), zero.copy()
    XX, ZZ, XZ = zero.copy(), zero.copy(), zero.copy()
    for step in range(3):
        U, W = _scatter_velocity(c, XX, ZZ, XZ, U, W, S, pt)
        if step == 0:
            W[gz, xc] += dt * amp
        XX, ZZ, XZ = _adj_stress_step(c, XX, ZZ, XZ, U, W, dtx, S, pt)
    sl = (slice(pt + 1, pt + 1 + nz), slice(nbc, nbc + nx))
    np.testing.assert_allclose(X, XX[sl], rtol=1e-11, atol=1e-22)
    np.testing.assert_allclose(Z, ZZ[sl], rtol=1e-11, atol=1e-22)
    np.testing.assert_allclose(XZr, XZ[sl], rtol=1e-11, atol=1e-22)
    assert np.count_nonzero(X) >= 12


def test_adjoint_order24_closed_forms():
    """Order 24 in the adjoint loop: one step by hand (stress stencil with S41/S42,
    e2drtm cpp:249-254), then a second step whose velocity comes from the
    scatter form of cpp:188-197."""
    nzbc, nxbc, nz, nx, pt = 14, 16, 6, 8, 2
    nbc = (nxbc - nx) // 2
    c = _rand_coefs(nzbc, nxbc, 23)
    dtx, dt, amp = 0.37, 0.011, 1.3
    gz, gx0, ng, g0 = 5, 3, 10, 4
    xc = gx0 + g0
    p, q = dt * amp, dtx * dt * amp
    # one step: nt = 3, slot nt-2 = 1
    A = np.zeros((3, ng))
    A[1, g0] = amp
    iv = O.adj_iv(3, nzbc, nxbc, dtx, ng, nbc, gz + 1, gx0 + 1, 1, 24, 1, nz, nx, dt)
    X, Z, XZr = _adj_stress_readout(c, iv, A, 3, 1, nz, nx)
    Z1 = np.zeros((nzbc, nxbc))
    XZ1 = np.zeros_like(Z1)
    Z1[gz, xc], Z1[gz + 1, xc], Z1[gz - 1, xc], Z1[gz + 2, xc] = -S41 * q, S41 * q, -S42 * q, S42 * q
    XZ1[gz, xc], XZ1[gz, xc + 1], XZ1[gz, xc - 1], XZ1[gz, xc + 2] = -S41 * q, S41 * q, -S42 * q, S42 * q
    sl = (slice(pt + 1, pt + 1 + nz), slice(nbc, nbc + nx))
    assert not np.any(X)
    np.testing.assert_allclose(Z, Z1[sl], rtol=1e-13, atol=0)
    np.testing.assert_allclose(XZr, XZ1[sl], rtol=1e-13, atol=0)
    # two steps: nt = 4, slot nt-3 = 1
    A = np.zeros((4, ng))
    A[2, g0] = amp
    iv = O.adj_iv(4, nzbc, nxbc, dtx, ng, nbc, gz + 1, gx0 + 1, 1, 24, 1, nz, nx, dt)
    X, Z, XZr = _adj_stress_readout(c, iv, A, 4, 1, nz, nx)
    W1 = np.zeros_like(Z1)
    W1[gz, xc] = p
    U2, W2 = _scatter_velocity(c, np.zeros_like(Z1), Z1, XZ1, np.zeros_like(Z1), W1, (S41, S42), pt)
    t = c["temp"]
    B = (slice(pt, nzbc - pt), slice(pt, nxbc - pt))
    r0, r1, c0, c1 = pt, nzbc - pt, pt, nxbc - pt
    sh = lambda a, di, dj: a[r0 + di:r1 + di, c0 + dj:c1 + dj]
    xx, zz, xz = np.zeros_like(Z1), Z1.copy(), XZ1.copy()
    xx[B] = -dtx * (S41 * (sh(U2, 0, 1) - sh(U2, 0, 0)) + S42 * (sh(U2, 0, 2) - sh(U2, 0, -1)))
    zz[B] = t[B] * Z1[B] - dtx * (S41 * (sh(W2, 0, 0) - sh(W2, -1, 0)) + S42 * (sh(W2, 1, 0) - sh(W2, -2, 0)))
    xz[B] = (t[B] * XZ1[B] - dtx * (S41 * (sh(U2, 1, 0) - sh(U2, 0, 0)) + S42 * (sh(U2, 2, 0) - sh(U2, -1, 0)))
             - dtx * (S41 * (sh(W2, 0, 0) - sh(W2, 0, -1)) + S42 * (sh(W2, 0, 1) - sh(W2, 0, -2))))
    zz[pt, :] = 0
    np.testing.assert_allclose(X, xx[sl], rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(Z, zz[sl], rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(XZr, xz[sl], rtol=1e-12, atol=1e-20)
    assert np.count_nonzero(X) >= 8


def test_forward_order24_one_step_closed_form():
    """Order 24 in the forward loop (eigen_staggerfd_single.cpp:229-235, 307-311):
    the strains of the hard source after step 0 (snapshot slot 0) and the
    step-1 velocities they make, by hand."""
    nzbc, nxbc, nz, nx, pt = 16, 18, 8, 10, 2
    nbc = (nxbc - nx) // 2
    c = _rand_coefs(nzbc, nxbc, 29)
    dtx = 0.29
    sz, sx = 7, 8                              # 0-based source cell, away from the free-surface row
    s = np.zeros(2)
    s[0] = 1.0
    ng, gx0 = nxbc - 2 * pt, pt
    out = {}
    for row in (sz, sz + 1):
        iv = O.fwd_iv(2, nzbc, nxbc, dtx, ng, sz + 1, sx + 1, row + 1, gx0 + 1, 1, 5, 24, 1, nz, nx)
        out[row] = O.staggerfd(iv, c, s)
    su, sw, sn = out[sz]
    # strains of slot 0 (interior origin (pt+1, nbc))
    iz, ix = sz - pt - 1, sx - nbc
    cw = {0: S41, 1: -S41, -1: S42, 2: -S42}  # bwx along x and bwz along z around the impulse
    ebwx = np.zeros((nz, nx))
    ebwz = np.zeros((nz, nx))
    for d, v in cw.items():
        ebwx[iz, ix + d] = v
        ebwz[iz + d, ix] = v
    np.testing.assert_allclose(sn["bwx"][:, :, 0], ebwx, rtol=1e-13, atol=0)
    np.testing.assert_allclose(sn["bwz"][:, :, 0], ebwz, rtol=1e-13, atol=0)
    assert not np.any(sn["fux"][:, :, 0]) and not np.any(sn["fuz"][:, :, 0])
    # step 1: uu on the source row from xx = dtx cl bwz and xz = dtx cm1 bwx
    k = lambda name, i, j: c[name][i, j]
    eu = np.zeros(ng)
    for d, v in cw.items():
        j = sx + d
        eu[j - gx0] = k("b", sz, j) * dtx * S41 * v * (k("cl", sz, sx) + k("cm1", sz, j))
    np.testing.assert_allclose(su[1], eu, rtol=1e-12, atol=1e-300)
    assert sw[1, sx - gx0] == 0.0              # overwritten by s[1] = 0
    # ww one row below the source, from zz = dtx ca bwz in the source column
    ew = k("b1", sz + 1, sx) * dtx * S41 * (S41 * k("ca", sz + 1, sx) - S42 * k("ca", sz + 2, sx)
                                           - S42 * k("ca", sz, sx))
    np.testing.assert_allclose(out[sz + 1][1][1, sx - gx0], ew, rtol=1e-12)
