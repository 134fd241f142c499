"""GPU parity of the dispersion path (Radon, picking, SoftMax adjoint source)
against the FP64 CPU oracle.  FP32 results: relative L2 <= 1e-4 (stated bar);
picks: within one velocity bin end to end, identical on identical inputs."""
import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import dispersion as Dp

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def gather():
    """Config 0 forward gather (FP64 oracle), shot on trace 5 (is = 5, M = 1)."""
    nz, nx, nbc, dx, dt, nt = 60, 200, 20, 0.5, 1e-4, 3000
    vs = np.tile(np.linspace(200, 600, nz)[:, None], (1, nx))
    vp = 1.73 * vs
    s = O.ricker(30, dt, nt)
    gx = np.arange(1, 200, dtype=float)
    _, sw, _ = O.staggerfd_eigend(1, nbc, nt, dt / dx, dx, dt, 5, 1, gx, np.ones(199), s, vp, vs, 1, 22, "w", 0, 5,
                                  snapshots=False)
    radon = (5, 0.5, dt, 701, 100, 800, 5, 80, nt, 199, 0.5, 199, 0, 1)
    return sw, radon, nt


@pytest.mark.parametrize("side", ["r", "l"])
def test_radon_vs_oracle(gather, side):
    sw, radon, nt = gather
    ml_o, L_o, sfb = O.rt_side(sw, side, *radon)
    fn = Dp.RTrAD if side == "r" else Dp.RTlAD
    ml, L, g = fn(sw, *radon)
    key = "mlr" if side == "r" else "mll"
    e = {"ml": rel(ml, ml_o), "mlr": rel(g[key], sfb["ml1"]),
         "mm": rel(g["mm"][:, sfb["lf"] - 1:], sfb["mm"][:, sfb["lf"] - 1:])}
    print(side, "radon fp32 rel-L2 vs fp64:", e)
    assert L == L_o and g["ccn"] == sfb["ccn"] == 20000
    assert max(e.values()) < 1e-4, e
    assert float(np.max(np.abs(ml - ml_o))) < 1e-4
    # single-precision gathers (SWD_single.m) take the f32 entry point
    ml32, _ = (Dp.RTr if side == "r" else Dp.RTl)(sw.astype(np.float32), *radon)
    assert rel(ml32, ml_o) < 1e-4


def test_pick_identical_on_identical_input_and_within_one_bin(gather):
    sw, radon, nt = gather
    for side in ("r", "l"):
        ml_o, _, _ = O.rt_side(sw, side, *radon)
        ml32 = ml_o.astype(np.float32)
        cr0 = np.full(ml32.shape[1], 250.0)
        res_o, cr_o = O.LHDispPick(ml32.astype(np.float64), 100, cr0, 1, 10)
        res_g, cr_g = Dp.LHDispPick(ml32, 100, cr0, 1, 10)
        np.testing.assert_array_equal(cr_g, cr_o)
        assert res_g == pytest.approx(res_o, rel=1e-12)
        # end to end: GPU Radon + GPU pick vs FP64 Radon + pick
        ml_g, _ = (Dp.RTr if side == "r" else Dp.RTl)(sw, *radon)
        _, cr_e = Dp.LHDispPick(ml_g, 100, cr0, 1, 10)
        _, cr_f = O.LHDispPick(ml_o, 100, cr0, 1, 10)
        assert np.max(np.abs(cr_e - cr_f)) <= 1.0
        # method 2 (column argmax)
        _, c2 = Dp.LHDispPick(ml32, 100, cr0, 2, 10)
        np.testing.assert_array_equal(c2, np.argmax(ml32, axis=0) + 101.0)


def test_pick_hand_built_columns():
    j = np.arange(1, 121)[:, None]
    def ridges(centres):
        out = 0.05 + np.zeros((120, len(centres)))
        for c in (20.0, 100.0):
            out = out + np.exp(-0.5 * ((j - c) / 5.0) ** 2)
        return out + np.exp(-0.5 * ((j - np.asarray(centres)[None, :]) / 5.0) ** 2)
    ml = ridges([60.0] * 8 + [0.0] + [64.0] * 3)
    ml[:, 8] = 0.05 + np.exp(-0.5 * ((np.arange(1, 121) - 30.0) / 5.0) ** 2)
    ml = ml.astype(np.float32)
    for ini in (3, 10):
        _, cg = Dp.LHDispPick(ml, 0, np.zeros(12), 1, ini)
        _, co = O.LHDispPick(ml.astype(np.float64), 0, np.zeros(12), 1, ini)
        np.testing.assert_array_equal(cg, co)
    flat = np.ones((30, 4), dtype=np.float32)
    flat[10:15] = 2.0
    _, c = Dp.LHDispPick(flat, 100, np.zeros(4), 1, 2)
    np.testing.assert_array_equal(c, 111.0)


@pytest.mark.parametrize("beta", [10.0, 1e4])
def test_adwdgrad_vs_oracle(gather, beta):
    sw, radon, nt = gather
    _, _, sr = O.rt_side(sw, "r", *radon)
    _, _, sl = O.rt_side(sw, "l", *radon)
    npair = sr["nf"] - sr["lf"] + 1
    r = np.random.default_rng(3)
    gr = np.round(r.normal(0, 5, npair))
    gl = np.round(r.normal(0, 5, npair))
    # identical (float32-rounded) spectra into both implementations
    for s in (sr, sl):
        s["mm"] = s["mm"].astype(np.complex64).astype(np.complex128)
        s["ml1"] = s["ml1"].astype(np.float32).astype(np.float64)
    ref = O.ADWDgrad_1(nt, 199, 5, npair, 5, 0, 0, 1, beta, gr, gl, sr, sl)
    space_M = np.tile(np.arange(1, 702)[:, None], (1, npair))
    def sfb(s, side):
        d = {"mm": s["mm"], "lf": s["lf"], "nf": s["nf"], "ccn": s["ccn"], "np": 701, "vmin": 100.0, "df": 0.5,
             "ds": 0.5, "m": 0}
        if side == "r":
            d["mlr"] = s["ml1"]; d["uxtposr"] = s["uxtpos"]
        else:
            d["mll"] = s["ml1"]; d["uxtposl"] = s["uxtpos"]
        return d
    a = Dp.ADWDgrad_1(nt, 199, 5, npair, 5, 0, 0, 1, beta, gr, gl, space_M, sfb(sr, "r"), sfb(sl, "l"))
    e = rel(a, ref)
    print(f"ADWDgrad_1 beta={beta} fp32 rel-L2 vs fp64: {e:.3e}")
    assert e < 1e-4


@pytest.mark.parametrize("side", ["r", "l"])
def test_fp64_spectra_are_picked_in_fp64(gather, side):
    """swd_lhdisppick_f64 (SWD.m passes double spectra): on the oracle's own
    FP64 spectrum the picks are those of the oracle's pickFDC1, both methods,
    with no rounding of the spectrum to FP32 in between."""
    sw, radon, nt = gather
    ml_o, _, _ = O.rt_side(sw, side, *radon)
    assert ml_o.dtype == np.float64
    cr0 = np.full(ml_o.shape[1], 300.0)
    for method in (1, 2):
        rg, cg = Dp.LHDispPick(ml_o, 100, cr0, method, 10)
        ro, co = O.LHDispPick(ml_o, 100, cr0, method, 10)
        np.testing.assert_array_equal(cg, co)
        assert rg == pytest.approx(ro, rel=1e-15)


@pytest.mark.parametrize("np_", [97, 701, 768, 769, 1500, 2048])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pick_random_ridges_plateaus_and_nans(np_, dtype):
    """pickFDC1 through LHDispPick on spectra built to hit every branch of the
    pick kernel's local-minimum search: several drifting ridges (width jumps
    and rejections), values quantised to 1/64 (runs of equal values, where
    findpeaks' run rule decides), a few NaNs, and velocity counts on both
    sides of the kernel's 768-velocity register ring.  Picks must equal the
    oracle's exactly (same spectrum into both)."""
    r = np.random.default_rng(np_ * 7 + (1 if dtype == np.float64 else 0))
    npair = 61
    j = np.arange(1, np_ + 1)[:, None]
    ml = 0.02 * r.random((np_, npair))
    for _ in range(4):
        c0, drift, w = r.uniform(0.1, 0.9) * np_, r.uniform(-0.3, 0.3) * np_ / npair, r.uniform(0.01, 0.05) * np_
        centres = c0 + drift * np.arange(npair)
        ml += r.uniform(0.3, 1.0) * np.exp(-0.5 * ((j - centres[None, :]) / w) ** 2)
    ml = np.round(ml * 64) / 64                     # plateaus of equal values
    ml[r.integers(0, np_, 5), r.integers(0, npair, 5)] = np.nan
    ml = ml.astype(dtype)
    for ini in (1, 30, npair):
        cr0 = np.full(npair, 0.5 * np_)
        res_g, cr_g = Dp.LHDispPick(ml, 100, cr0, 1, ini)
        res_o, cr_o = O.LHDispPick(ml.astype(np.float64), 100, cr0, 1, ini)
        np.testing.assert_array_equal(cr_g, cr_o)
        assert res_g == pytest.approx(res_o, rel=1e-12)
