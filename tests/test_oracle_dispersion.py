"""Pins the dispersion part of the CPU oracle (Radon, picking, SoftMax adjoint source).

Reference behaviour pinned: core/RTr.m:5-32, core/RTl.m:5-32, core/RTrAD.m:28-34,
core/pickFDC1.m:19-131, core/LHDispPick.m:4-12, core/ADWDgrad_1.m:13-77,
core/smooth2a.m:52-76.  MATLAB's findpeaks/interp1/softmax/fft are third-party
(toolboxes of R2023b); their published definitions are restated and checked on
hand-built inputs.
"""
import math

import numpy as np
import pytest

from oracle import swd_oracle as O


def test_ccn_and_frequency_indices():
    assert O.ccn_of(0.5, 8e-5) == 24999
    assert O.ccn_of(0.5, 1e-4) == 20000
    assert O.ccn_of(0.1, 8e-5) == 124999
    assert O.mround(5 / 0.5) + 1 == 11 and O.mround(80 / 0.5) + 1 == 161
    assert O.mround(2.5) == 3 and O.mround(-2.5) == -3


def test_windows_right_and_left():
    # is=3, M=3, m=0 -> source trace 7; b=10 traces; offset 5
    assert O.rt_window("r", 3, 10, 5, 0, 3) == (7, 10)          # nend = 11 clipped to b
    assert O.rt_window("l", 3, 10, 5, 0, 3) == (3, 7)           # nend = 2
    assert O.rt_window("l", 1, 10, 5, 0, 3) == (1, 1)           # nend < 1 -> 0
    assert O.rt_window("r", 2, 100, 5, 1, 3) == (5, 9)          # m = 1 shifts the start


def _pw(nt, ng, dt, dx, c0, f0=30.0, shift=0.05):
    """Plane wave w(t - x/c0) built in the frequency domain (exact delays)."""
    n = 4 * nt
    f = np.fft.rfftfreq(n, dt)
    W = (f / f0) ** 2 * np.exp(-(f / f0) ** 2)
    x = np.arange(ng) * dx
    D = W[:, None] * np.exp(-2j * np.pi * f[:, None] * (shift + x[None, :] / c0))
    return np.fft.irfft(D, n=n, axis=0)[:nt]


def test_plane_wave_peaks_at_its_velocity():
    nt, ng, dt, dx, c0 = 4000, 120, 1e-4, 0.5, 350.0
    d = _pw(nt, ng, dt, dx, c0)
    ml, L, sfb = O.rt_side(d, "r", 1, 0.5, dt, 701, 100, 800, 20, 60, nt, ng, dx, ng, 0, 1)
    assert L == ng and ml.shape == (701, 81)
    v = 100 + np.argmax(ml, axis=0)
    assert np.all(np.abs(v - c0) <= 1)
    np.testing.assert_allclose(ml.max(axis=0), 1.0)
    # the left side reads the same traces flipped: for a wave travelling right the
    # left window sees it arriving from the far trace first (negative moveout)
    dl = d[:, ::-1].copy()
    mll, Ll, _ = O.rt_side(dl, "l", ng, 0.5, dt, 701, 100, 800, 20, 60, nt, ng, dx, ng, 0, 1)
    np.testing.assert_allclose(mll, ml, rtol=1e-10, atol=1e-12)


def test_single_trace_spectrum_is_flat_in_velocity():
    r = np.random.default_rng(0)
    d = r.standard_normal((500, 8))
    ml, L, sfb = O.rt_side(d, "r", 8, 0.5, 1e-3, 51, 100, 150, 10, 40, 500, 8, 1.0, 1, 0, 1)
    assert L == 1 and np.allclose(sfb["x"], 0.0)
    np.testing.assert_allclose(ml, 1.0, rtol=1e-12)
    # |mm| equals |fft| of the trace at the bin
    D = np.fft.fft(d[:, 7], 2000)
    np.testing.assert_allclose(sfb["ml1"][0], np.abs(D[sfb["lf"] - 1:sfb["nf"]]), rtol=1e-10)


def test_radon_linearity_and_truncation():
    r = np.random.default_rng(1)
    a, b = r.standard_normal((600, 20)), r.standard_normal((600, 20))
    args = (1, 0.5, 2e-3, 31, 100, 130, 5, 40, 600, 20, 1.0, 20, 0, 1)
    _, _, sa = O.rt_side(a, "r", *args)
    _, _, sb = O.rt_side(b, "r", *args)
    _, _, sc = O.rt_side(2 * a - b, "r", *args)
    np.testing.assert_allclose(sc["mm"], 2 * sa["mm"] - sb["mm"], rtol=1e-9, atol=1e-9)
    # ccn = 1000 > nt pads; with dt = 4e-3, ccn = 500 < nt = 600 truncates the trace
    _, _, st = O.rt_side(a, "r", 1, 0.5, 4e-3, 31, 100, 130, 5, 40, 600, 20, 1.0, 20, 0, 1)
    assert st["ccn"] == 500
    _, _, s2 = O.rt_side(a[:500], "r", 1, 0.5, 4e-3, 31, 100, 130, 5, 40, 500, 20, 1.0, 20, 0, 1)
    np.testing.assert_allclose(st["mm"], s2["mm"])


def test_slant_stack_adjoint_identity():
    r = np.random.default_rng(2)
    theta = (2 * np.pi * 0.5) * np.outer(1.0 / np.arange(100, 131), np.arange(12) * 0.5)
    lf, nf, ccn = 4, 20, 64
    d = r.standard_normal((12, ccn)) + 1j * r.standard_normal((12, ccn))
    y = r.standard_normal((31, nf - lf + 1)) + 1j * r.standard_normal((31, nf - lf + 1))
    mm = O.slant_stack(theta, d, lf, nf)[:, lf - 1:nf]
    g = O.slant_stack_adj(theta, y, lf, nf, ccn)
    lhs = np.vdot(y, mm)
    rhs = np.vdot(g[:, lf - 1:nf], d[:, lf - 1:nf])
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)
    assert not np.any(g[:, :lf - 1]) and not np.any(g[:, nf:])


def test_inverse_dft_one_sided_closed_form():
    ccn, nt, k = 50, 40, 7
    z = 0.3 - 1.1j
    g = np.zeros((1, ccn), dtype=complex)
    g[0, k] = z
    a = O.inv_dft_real(g, ccn, nt)[0]
    t = np.arange(nt)
    np.testing.assert_allclose(a, np.real(z * np.exp(2j * np.pi * k * t / ccn)) / ccn, rtol=1e-12, atol=1e-15)
    with pytest.raises(ValueError):
        O.inv_dft_real(g, ccn, ccn + 1)


# ---------------------------------------------------------------- picking
@pytest.mark.parametrize("y,locs", [
    ([1, 3, 2], [2]),
    ([1, 3, 3, 2], [2]),           # flat peak -> first index
    ([1, 3, 3, 4, 2], [4]),        # shoulder is not a peak
    ([3, 1, 2], []),               # endpoints never count
    ([1, 2, 2], []),
    ([0, 2, 1, 5, 5, 5, 1, 0, 1], [2, 4]),
    ([1, 2, np.nan, 3, 1], []),    # NaN breaks the rise/fall pattern
    ([1, 3, 1, np.nan, 1, 4, 2], [2, 6]),
])
def test_findpeaks_rules(y, locs):
    assert list(O.findpeaks_locs(np.array(y, dtype=float))) == locs


def _ridge_ml(a, b, centres, width=6.0):
    j = np.arange(1, a + 1)[:, None]
    return np.exp(-0.5 * ((j - np.asarray(centres)[None, :]) / width) ** 2) + 0.05


def test_pick_follows_single_ridge():
    centres = np.linspace(80, 40, 30)
    ml = _ridge_ml(120, 30, centres)
    idx = O.pickFDC1(ml, 10, 2, 4)
    np.testing.assert_array_equal(idx, np.round(centres))


def _three_ridges(a, centres, width=5.0):
    j = np.arange(1, a + 1)[:, None]
    out = 0.05 + np.zeros((a, len(centres)))
    for c in (20.0, 100.0):
        out += np.exp(-0.5 * ((j - c) / width) ** 2)
    return out + np.exp(-0.5 * ((j - np.asarray(centres)[None, :]) / width) ** 2)


def test_pick_rejects_width_jump_and_interpolates():
    """Tracking right of lp: a column whose local-minimum window is > fg x wider is
    rejected (chazhi = 0), carries the previous pick, and is re-filled by linear
    interpolation between the kept neighbours (pickFDC1.m:79-131)."""
    a = 120
    centres = [60.0] * 8 + [0.0] + [64.0] * 3
    ml = _three_ridges(a, centres)
    j = np.arange(1, a + 1)
    ml[:, 8] = 0.05 + np.exp(-0.5 * ((j - 30.0) / 5.0) ** 2)   # single ridge: window = whole column
    idx = O.pickFDC1(ml, 3, 2, 4)
    np.testing.assert_array_equal(idx[:8], 60.0)
    np.testing.assert_array_equal(idx[9:], 64.0)
    assert idx[8] == pytest.approx(62.0)                          # interpolated, not 30


def test_pick_left_of_lp_and_nearest_extrapolation():
    """Left of lp the width test uses lfg; rejected end columns are filled by
    nearest extrapolation (the 'nearest','extrap' pass of pickFDC1.m:131)."""
    a = 120
    centres = [0.0, 58.0, 60.0, 60.0, 60.0]
    ml = _three_ridges(a, centres)
    j = np.arange(1, a + 1)
    ml[:, 0] = 0.05 + np.exp(-0.5 * ((j - 30.0) / 5.0) ** 2)
    idx = O.pickFDC1(ml, 4, 2, 4)
    # window width 119 vs ~40 * lfg = 160: not rejected with lfg = 4 -> argmax (30)
    assert idx[0] == 30.0 and idx[1] == 58.0
    idx2 = O.pickFDC1(ml, 4, 2, 2)
    # with lfg = 2 column 1 is rejected; it is outside the kept range -> nearest (58)
    assert idx2[0] == 58.0 and idx2[1] == 58.0


def test_pick_ties_take_first_index_and_lhdisppick():
    ml = np.ones((30, 4))
    ml[10:15, :] = 2.0                                # flat top: first index wins
    idx = O.pickFDC1(ml, 2, 2, 4)
    np.testing.assert_array_equal(idx, 11.0)
    res, cr = O.LHDispPick(ml, 100, np.full(4, 112.0), 1, 2)
    np.testing.assert_array_equal(cr, 111.0)          # cr_pre = idx + vmin (not vmin + idx - 1)
    assert res == 1.0


# ---------------------------------------------------------------- soft-argmax
def test_softargmax_grad_is_derivative_of_expected_index():
    r = np.random.default_rng(4)
    npv, beta = 25, 7.0
    G = 1.0 + 0.05 * r.standard_normal((npv, 3))
    g = np.array([0.7, -2.0, 0.0])
    gml = O.softargmax_grad(G ** 40, g, beta)
    j = np.arange(1, npv + 1)

    def kappa(col):
        e = np.exp(beta * col - np.max(beta * col))
        e /= e.sum()
        return np.sum(j * e)

    geff = np.where(g == 0, 1.0, g)
    for c in range(3):
        num = np.zeros(npv)
        for q in range(npv):
            h = 1e-6
            up, dn = G[:, c].copy(), G[:, c].copy()
            up[q] += h
            dn[q] -= h
            num[q] = (kappa(up) - kappa(dn)) / (2 * h)
        np.testing.assert_allclose(gml[:, c], 2 * geff[c] * num, rtol=1e-5, atol=1e-8)


def test_softmax_stable_at_reference_beta():
    G = np.array([[1.0], [1.4], [1.39999]])
    E = O.softmax_cols(1e4 * G)
    assert np.all(np.isfinite(E)) and abs(E.sum() - 1) < 1e-15 and E[1, 0] > E[2, 0] > E[0, 0]


def test_adwdgrad_branch_layout():
    """Right branch fills the window traces; the left branch drops the source trace and un-flips."""
    nt, ng, dt = 300, 9, 2e-3
    r = np.random.default_rng(6)
    d = r.standard_normal((nt, ng))
    args = dict(df=0.5, dt=dt, np_=21, vmin=100, vmax=120, fmin=5, fmax=20, a=nt, b=ng, ds=1.0, offset=9, m=0, M=2)
    is_ = 3                                             # source trace 5
    _, Lr, sr = O.rt_side(d, "r", is_, *args.values())
    _, Ll, sl = O.rt_side(d, "l", is_, *args.values())
    assert (Lr, Ll) == (5, 5)
    npair = sr["nf"] - sr["lf"] + 1
    g = np.linspace(-3, 3, npair)
    a = O.ADWDgrad_1(nt, ng, 3, npair, is_, 0, 0, 2, 10.0, g, g, sr, sl)
    gr = O.adjoint_side(sr, g, 10.0, nt)
    gl = O.adjoint_side(sl, g, 10.0, nt)
    np.testing.assert_array_equal(a[:, 4:9], gr.T)
    np.testing.assert_array_equal(a[:, 0:4], gl[::-1][:-1].T)
    # guards: with ns=3, w=0, M=2 the right branch needs is <= 3, the left is >= 1
    a2 = O.ADWDgrad_1(nt, ng, 2, npair, is_, 0, 0, 2, 10.0, g, g, sr, sl)
    assert not np.any(a2[:, 4:9])


def test_smooth2a_box_mean_truncated_edges():
    r = np.random.default_rng(7)
    A = r.standard_normal((9, 11))
    A[3, 4] = np.nan
    out = O.smooth2a(A, 1, 2)
    for i in range(9):
        for k in range(11):
            blk = A[max(i - 1, 0):i + 2, max(k - 2, 0):k + 3]
            if np.isnan(A[i, k]):
                assert np.isnan(out[i, k])
            else:
                np.testing.assert_allclose(out[i, k], np.nanmean(blk), rtol=1e-12)


def test_assembly_formula():
    vs = np.full((2, 2), 300.0)
    dk = O.assemble_dk_vs(vs, np.full((2, 2), 2.0), np.full((2, 2), 3.0), np.full((2, 2), 4.0))
    np.testing.assert_allclose(dk, -4 * 300 * 0.5 + 2 * 300 * 0.75)
