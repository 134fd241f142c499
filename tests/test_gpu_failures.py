"""Failure paths on the GPU: inputs the reference would turn into NaN, an
error or an unstable run, and the library's answer to each.

- A Radon column of zero amplitude (mm = 0 in a whole frequency column): its
  phase mm ./ abs(mm) is 0/0, which ADWDgrad_1.m:32-33 sets to 0, so the
  column contributes nothing; the adjoint source stays finite and equals the
  FP64 oracle's.
- A model whose velocities break the staggered-grid stability bound
  (vmax dt / dx > 1/sqrt(2)) is rejected with SWD_ERR_CFL by the engine and,
  request by request, by the server, which goes on serving the others."""
import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import dispersion as Dp
from swd_amd import problems as P
from swd_amd._lib import SwdError
from swd_amd.engine import Engine

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_zero_amplitude_radon_columns_contribute_nothing():
    nz, nx, nbc, dx, dt, nt = 40, 160, 15, 0.5, 1e-4, 1500
    vs = np.tile(np.linspace(200, 500, nz)[:, None], (1, nx))
    _, sw, _ = O.staggerfd_eigend(1, nbc, nt, dt / dx, dx, dt, 5, 1, np.arange(1, 160, dtype=float), np.ones(159),
                                  O.ricker(30, dt, nt), 1.73 * vs, vs, 1, 22, "w", 0, 5, snapshots=False)
    radon = (5, 0.5, dt, 701, 100, 800, 5, 80, nt, 159, 0.5, 159, 0, 1)
    _, _, sr = O.rt_side(sw, "r", *radon)
    _, _, sl = O.rt_side(sw, "l", *radon)
    npair = sr["nf"] - sr["lf"] + 1
    dead = [0, 17, npair - 1]             # frequency columns with no energy at all
    for s in (sr, sl):
        s["mm"] = s["mm"].astype(np.complex64).astype(np.complex128)
        s["ml1"] = s["ml1"].astype(np.float32).astype(np.float64)
        for k in dead:
            s["mm"][:, s["lf"] - 1 + k] = 0
            s["ml1"][:, k] = 0
    r = np.random.default_rng(8)
    gr = np.round(r.normal(0, 5, npair))
    gl = np.round(r.normal(0, 5, npair))
    ref = O.ADWDgrad_1(nt, 159, 5, npair, 5, 0, 0, 1, 10.0, gr, gl, sr, sl)
    assert np.all(np.isfinite(ref))

    def sfb(s, side):
        d = {"mm": s["mm"], "lf": s["lf"], "nf": s["nf"], "ccn": s["ccn"], "np": 701, "vmin": 100.0, "df": 0.5,
             "ds": 0.5, "m": 0}
        d["mlr" if side == "r" else "mll"] = s["ml1"]
        d["uxtposr" if side == "r" else "uxtposl"] = s["uxtpos"]
        return d
    a = Dp.ADWDgrad_1(nt, 159, 5, npair, 5, 0, 0, 1, 10.0, gr, gl, None, sfb(sr, "r"), sfb(sl, "l"))
    assert np.all(np.isfinite(a))
    e = rel(a, ref)
    print(f"ADWDgrad_1 with {len(dead)} zero-amplitude columns: fp32 rel-L2 vs fp64 {e:.2e}")
    assert e < 1e-4
    # the dead columns carry no gradient: zeroing their misfit changes nothing
    g0r, g0l = gr.copy(), gl.copy()
    g0r[dead] = 0
    g0l[dead] = 0
    b = Dp.ADWDgrad_1(nt, 159, 5, npair, 5, 0, 0, 1, 10.0, g0r, g0l, None, sfb(sr, "r"), sfb(sl, "l"))
    np.testing.assert_array_equal(a, b)


def _small():
    prob, mod = P.plain_single_shot(nt=1000)
    return prob, mod


def test_engine_rejects_cfl_violation():
    prob, mod = _small()
    e = Engine(prob)
    try:
        fast = 2.0 / np.sqrt(2.0) * prob["dx"] / prob["dt"]      # vmax dt / dx = 1.41
        with pytest.raises(SwdError) as ex:
            e.set_model(np.full_like(mod["vp"], fast), np.full_like(mod["vs"], fast / 1.73))
        assert ex.value.code == -3
        e.set_model(mod["vp"], mod["vs"])                        # the handle stays usable
        e.set_wavelet(mod["s"])
        e.set_observed_picks(np.full((e.npair, 1), 300.0), np.full((e.npair, 1), 300.0))
        g = e.gradient()
        assert np.isfinite(g["res_r"]).all()
    finally:
        e.close()


def test_server_fails_an_unstable_request_alone():
    from swd_amd.serve import Server
    prob, mod = _small()
    from swd_amd.engine import problem_npair
    npair = problem_npair(prob)
    cr0 = np.full((npair, 1), 300.0)
    fast = 1.5 * prob["dx"] / prob["dt"]
    s = Server(hbm_bytes=4 << 30, max_active=2, nt_block=(4, 4))
    try:
        bad = s.submit(prob, np.full_like(mod["vp"], fast), np.full_like(mod["vs"], fast / 1.73), mod["s"], cr0, cr0)
        good = s.submit(prob, mod["vp"], mod["vs"], mod["s"], cr0, cr0)
        with pytest.raises(SwdError) as ex:
            s.wait(bad)
        assert ex.value.code == -3
        r = s.wait(good)
    finally:
        s.close()
    e = Engine(prob)
    try:
        e.set_wavelet(mod["s"])
        e.set_observed_picks(cr0, cr0)
        e.set_model(mod["vp"], mod["vs"])
        np.testing.assert_array_equal(r["dk_vs"], e.gradient()["dk_vs"])
    finally:
        e.close()
