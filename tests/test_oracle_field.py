"""The oracle's field-script variants pinned on closed forms and hand-built
spectra (norm_trace, the field LHDispPick, the peak-normalised ADWDgrad_1),
and the host side of the field-data inputs (variants, manual-pick slicing,
the outcrop geometry).  CPU only."""
import os

import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import field as F

OUTCROP_DISP = "/root/reference/fieldexamples/outcrop/outcropDisp.mat"


def test_norm_trace_closed_form():
    r = np.random.default_rng(1)
    d = r.standard_normal((50, 7))
    gains = 10.0 ** r.uniform(-3, 3, 7)
    out = O.norm_trace(d * gains)
    assert np.allclose(np.max(np.abs(out), axis=0), 1.0, rtol=0, atol=1e-15)
    assert np.array_equal(np.sign(out), np.sign(d))
    assert np.allclose(out, O.norm_trace(d), rtol=1e-13, atol=0)       # positive column gains drop out
    k = np.argmax(np.abs(d[:, 2]))
    assert out[k, 2] == np.sign(d[k, 2])


def test_norm_trace_nan_and_zero_columns():
    d = np.array([[1.0, 0.0, np.nan], [-4.0, 0.0, 2.0], [np.nan, 0.0, -8.0]])
    out = O.norm_trace(d)
    assert np.allclose(out[:2, 0], [0.25, -1.0]) and np.isnan(out[2, 0])   # max skips NaN
    assert np.all(np.isnan(out[:, 1]))                                     # 0/0
    assert np.isnan(out[0, 2]) and np.allclose(out[1:, 2], [0.25, -1.0])


def _spectrum(np_=40, npair=14, strong=9, seed=0):
    """Unit-peak Gaussian columns along a smooth ridge, |mm| column gains
    with the largest at column ``strong`` (0-based)."""
    r = np.random.default_rng(seed)
    j = np.arange(1, np_ + 1)[:, None]
    ridge = np.round(12 + 10 * np.linspace(0, 1, npair) ** 1.5).astype(int)
    ml = np.exp(-0.5 * ((j - ridge[None, :]) / 2.5) ** 2) + 0.01 * r.random((np_, npair))
    ml = ml / ml.max(axis=0, keepdims=True)
    gain = np.linspace(1.0, 3.0, npair)
    gain[strong] = 10.0
    return ml, ml * gain[None, :], ridge


def test_field_pick_start_and_window():
    ml, mlr, ridge = _spectrum()
    cr0 = np.zeros(ml.shape[1])
    res, cr, ini = O.LHDispPick_field(ml, mlr, 100.0, cr0, 1, 0, "ml")
    assert ini == 10                                                   # 1-based column of the largest |mm|
    assert np.array_equal(cr - 100.0, ridge)                           # the ridge, picked from there
    res2, cr2, _ = O.LHDispPick_field(ml, mlr, 100.0, cr0, 1, 0, "mlr")
    assert np.array_equal(cr, cr2)                                     # pickFDCValue from the argmax row
    assert np.array_equal(O.pickFDC1(ml, 10, 2, 4), O.pickFDC1(ml, 10, 2, 4, lpvalue=ridge[9]))
    d = np.arange(ml.shape[1], dtype=float) - 3.0
    res3, _, _ = O.LHDispPick_field(ml, mlr, 100.0, cr - d, 3, 4, "ml")
    assert res3 == pytest.approx(np.mean(d[2:ml.shape[1] - 4] ** 2), rel=1e-15)
    assert res == pytest.approx(np.mean(cr ** 2))
    # the misfit window of SWD.m's picker path through pick_side
    r4, c4 = O.pick_side(ml, {"ml1": mlr}, 100.0, cr - d, 10, {"res_first": 3, "res_last": 4})
    assert np.array_equal(c4, cr) and r4 == pytest.approx(res3)


def test_adwdgrad_peak_norm_closed_form():
    r = np.random.default_rng(5)
    nt, ng, np_, L = 64, 9, 30, 6
    seismo = r.standard_normal((nt, ng))
    radon = dict(df=10.0, dt=1e-3, np_=np_, vmin=100.0, vmax=100.0 + np_ - 1, fmin=100.0, fmax=300.0)
    ml, _, sfb = O.rt_side(seismo, "r", 1, radon["df"], radon["dt"], np_, radon["vmin"], radon["vmax"],
                           radon["fmin"], radon["fmax"], nt, ng, 0.5, L, 0, 1)
    g = np.round(r.normal(0, 3, ml.shape[1]))
    a = O.ADWDgrad_1(nt, ng, 5, ml.shape[1], 1, 3, 0, 1, 10.0, g, g, sfb, None)        # right branch only
    an = O.ADWDgrad_1(nt, ng, 5, ml.shape[1], 1, 3, 0, 1, 10.0, g, g, sfb, None, peak_norm=True)
    assert np.max(np.abs(an)) == pytest.approx(1.0, abs=1e-15)
    assert np.allclose(an, a / np.max(np.abs(a)), rtol=1e-15, atol=0)


def test_variant_structs():
    v = F.variant_struct("outcrop")
    assert (v.norm_trace, v.auto_start, v.res_first, v.res_last, v.adj_peak_norm) == (1, 1, 15, 15, 1)
    assert v.vp_ratio == pytest.approx(1.732)
    q = F.variant_struct("qadimah")
    assert (q.res_first, q.res_last) == (50, 30)
    z = F.variant_struct("swd")
    assert (z.norm_trace, z.auto_start, z.res_first, z.res_last, z.adj_peak_norm, z.vp_ratio, z.fk_filter) == (0,) * 7
    assert F.variant_struct({"fk_filter": 1}).fk_filter == 1
    with pytest.raises(ValueError):
        F.variant_struct({"norm": 1})


def test_manual_pick_slicing_rules():
    nf, ns, no = 991, 4, 20
    cube = np.arange(nf * ns * no, dtype=float).reshape((nf, ns, no), order="F")
    cr = F.manual_picks(cube, 20.0, 85.0, 0.1, 50, 5)          # outcrop.m:148-157: rows 191:841, offset 10
    assert cr.shape == (651, ns)
    assert np.array_equal(cr, cube[190:841, :, 9])
    cr = F.manual_picks(cube, 10.0, 60.0, 0.1, 39, 3)          # QademahSWD3.m:149-153: rows 91:591, offset 13
    assert cr.shape == (501, ns)
    assert np.array_equal(cr, cube[90:591, :, 12])
    with pytest.raises(ValueError):
        F.manual_picks(cube, 20.0, 85.0, 0.1, 52, 5)           # 52/5 is not an offset index
    with pytest.raises(ValueError):
        F.manual_picks(cube, 20.0, 120.0, 0.1, 50, 5)          # past the cube's 100 Hz


def test_outcrop_geometry():
    prob, mod = F.outcrop()
    assert prob["nt"] == 7450 and prob["nt"] % prob["in_wf"] == 0
    vpmax = 671 * 1.6 * 1.732
    assert vpmax * prob["dt"] / prob["dx"] == pytest.approx(0.5)
    assert np.size(prob["sx"]) == 60 and prob["ng"] == 60 and prob["M"] == 1
    assert prob["fmin"] == 20.0 and F.outcrop_fmin(5) == 25.0
    from swd_amd.engine import problem_npair
    assert problem_npair(prob) == 651
    assert mod["vp"][0, 0] == pytest.approx(150 * 1.732)


@pytest.mark.skipif(not os.path.exists(OUTCROP_DISP), reason="the outcrop pick cube is not present here")
def test_outcrop_manual_cube_fits_the_outcrop_problem():
    crr, crl = F.load_manual_cubes(OUTCROP_DISP)
    assert crr.shape == (991, 60, 10) and crl.shape == (991, 60, 10)
    prob, _ = F.outcrop(offset=50)
    from swd_amd.engine import problem_npair
    for cube in (crr, crl):
        cr = F.manual_picks(cube, prob["fmin"], prob["fmax"], prob["df"], prob["offset"], 5)
        assert cr.shape == (problem_npair(prob), 60)
        assert np.all(np.isfinite(cr)) and cr.min() >= 1 and cr.max() <= 1000


GOLDEN_PICKS = os.path.join(os.path.dirname(__file__), "golden", "outcrop_manual_picks_offset50.npz")


def test_outcrop_manual_picks_fixture():
    """The committed fixture is the outcrop script's slice of its manual cubes
    (regenerated and compared where outcropDisp.mat is present)."""
    d = np.load(GOLDEN_PICKS)
    prob, _ = F.outcrop(offset=int(d["offset"]))
    from swd_amd.engine import problem_npair
    assert d["cr0_r"].shape == d["cr0_l"].shape == (problem_npair(prob), d["shots"].size)
    assert np.all(np.isfinite(d["cr0_r"])) and np.all(np.isfinite(d["cr0_l"]))
    if os.path.exists(OUTCROP_DISP):
        import importlib.util
        spec = importlib.util.spec_from_file_location(
            "make_outcrop_manual_picks", os.path.join(os.path.dirname(GOLDEN_PICKS), "make_outcrop_manual_picks.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        fresh = mod.make(OUTCROP_DISP)
        for k in d.files:
            np.testing.assert_array_equal(d[k], fresh[k])


QADIMAH_MANUAL = "/root/reference/fieldexamples/QadimahFault/manualDispfile3of3.mat"
GOLDEN_QPICKS = os.path.join(os.path.dirname(__file__), "golden", "qadimah_manual_picks_offset39.npz")


def test_qadimah_survey_and_manual_picks_fixture():
    """QademahSWD3.m:75-136 geometry, and the committed fixture as the script's
    smoothed slice of its manual cubes (regenerated and compared where
    manualDispfile3of3.mat is present)."""
    from swd_amd.engine import problem_npair
    prob, mod = F.qadimah()
    assert prob["sx"].size == prob["ns"] == 120 and prob["ng"] == 240 and prob["M"] == 2
    assert prob["dt"] == pytest.approx(0.25 / (1200 * 1.732) * 0.5, rel=1e-15) and prob["nt"] % prob["in_wf"] == 0
    assert problem_npair(prob) == 501 and mod["vs"][0, 0] == 200 and mod["vs"][-1, 7] == 1100
    d = np.load(GOLDEN_QPICKS)
    assert d["cr0_r"].shape == d["cr0_l"].shape == (501, d["shots"].size)
    assert np.all(np.isfinite(d["cr0_r"])) and np.all(np.isfinite(d["cr0_l"]))
    if os.path.exists(QADIMAH_MANUAL):
        import importlib.util
        spec = importlib.util.spec_from_file_location(
            "make_qadimah_manual_picks", os.path.join(os.path.dirname(GOLDEN_QPICKS), "make_qadimah_manual_picks.py"))
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        fresh = m.make(QADIMAH_MANUAL)
        for k in d.files:
            np.testing.assert_array_equal(d[k], fresh[k])


def _fk_brute(d):
    """fk_filter.m's steps with explicit sums and explicit shift indices (1-based as in MATLAB)."""
    ns, ntr = d.shape
    nf, nk = 2 * ns, 2 * ntr
    P = np.zeros((nf, nk), dtype=complex)
    P[:ns, :ntr] = d
    F2 = np.zeros((nf, nk), dtype=complex)
    for u in range(nf):
        for v in range(nk):
            F2[u, v] = sum(P[t, x] * np.exp(-2j * np.pi * (u * t / nf + v * x / nk)) for t in range(ns) for x in range(ntr))
    # fftshift of even sizes: shifted(i, j) = F2((i + nf/2) mod nf, (j + nk/2) mod nk)
    S = np.array([[F2[(i + nf // 2) % nf, (j + nk // 2) % nk] for j in range(nk)] for i in range(nf)])
    for i in range(1, ns + 1):                 # for i = 1:nfhalf, for j = 1:nkhalf
        for j in range(1, ntr + 1):
            S[i - 1, j - 1] = 0
            S[i + ns - 1, j + ntr - 1] = 0
    B = np.array([[S[(i + nf // 2) % nf, (j + nk // 2) % nk] for j in range(nk)] for i in range(nf)])
    out = np.zeros((ns, ntr))
    for t in range(ns):
        for x in range(ntr):
            out[t, x] = np.real(sum(B[u, v] * np.exp(2j * np.pi * (u * t / nf + v * x / nk))
                                    for u in range(nf) for v in range(nk))) / (nf * nk)
    return out


@pytest.mark.parametrize("shape", [(5, 3), (6, 4), (4, 1)])
def test_fk_filter_brute_force(shape):
    d = np.random.default_rng(shape[0] * 7 + shape[1]).standard_normal(shape)
    assert np.allclose(O.fk_filter(d), _fk_brute(d), rtol=0, atol=1e-12)
    d2 = np.random.default_rng(99).standard_normal(shape)
    assert np.allclose(O.fk_filter(2 * d - 3 * d2), 2 * O.fk_filter(d) - 3 * O.fk_filter(d2), atol=1e-12)


@pytest.mark.parametrize("nt,L,ccn,c0,npair", [(40, 7, 64, 3, 9), (41, 8, 64, 0, 12), (50, 6, 30, 2, 7), (33, 1, 40, 1, 5)])
def test_fk_spectrum_factorisation(nt, L, ccn, c0, npair):
    """The form the device evaluates: with W the DFT at the bins (t < min(nt, ccn)),
    N = 2nt, M = 2L, E / O the even / odd-t sums of W, Xe / Xo the even / odd-t
    sums of each trace and H = U' X with U'[k, s] = -(1/N) sum_{t-s odd} W[t, k] cot(pi (t-s)/N):
      W^T fk_filter(X) [:, l] = D[:, l]/2 - (1/M) sum_{l-l' odd} [(2/N)(O Xe[l'] + E Xo[l'])
                                                                  + 2 cot(pi (l-l')/M) H[:, l']]."""
    r = np.random.default_rng(nt + L)
    X = r.standard_normal((nt, L))
    nte = min(nt, ccn)
    W = np.exp(-2j * np.pi * np.outer(np.arange(nte), c0 + np.arange(npair)) / ccn)
    ref = W.T @ O.fk_filter(X)[:nte]
    N, M = 2 * nt, 2 * L
    D = W.T @ X[:nte]
    t = np.arange(nte)
    E, Od = W[t % 2 == 0].sum(0), W[t % 2 == 1].sum(0)
    Xe, Xo = X[0::2].sum(0), X[1::2].sum(0)
    tau = t[:, None] - np.arange(nt)[None, :]
    K = np.zeros(tau.shape)
    odd = np.abs(tau) % 2 == 1
    K[odd] = -1.0 / np.tan(np.pi * tau[odd] / N) / N
    H = (W.T @ K) @ X
    Df = np.zeros((npair, L), dtype=complex)
    for l in range(L):
        acc = np.zeros(npair, dtype=complex)
        for lp in range(1 - l % 2, L, 2):
            acc += (2 / N) * (Od * Xe[lp] + E * Xo[lp]) + 2 / np.tan(np.pi * (l - lp) / M) * H[:, lp]
        Df[:, l] = D[:, l] / 2 - acc / M
    assert np.max(np.abs(Df - ref)) <= 1e-13 * np.max(np.abs(ref))
