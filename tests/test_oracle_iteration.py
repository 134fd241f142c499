"""The oracle's restatement of the SWD step length and line search
(SWD.m:203-322), pinned on misfits whose behaviour in the step factor is
known in closed form."""
import numpy as np

from oracle import swd_oracle as O


def _setup(n=50, seed=0):
    r = np.random.default_rng(seed)
    vs = r.uniform(200.0, 400.0, (5, n))
    dk = r.standard_normal((5, n))
    return vs, dk


def _factor(vs, vs1, dk, alpha):
    """Recover the step factor f from vs1 = vs + alpha f dk (unclipped cells)."""
    d = dk.ravel()
    k = int(np.argmax(np.abs(d)))
    return (vs1.ravel()[k] - vs.ravel()[k]) / (alpha * d[k])


def test_alpha_is_norm_ratio_times_pur():
    vs, dk = _setup()
    _, rec = O.swd_line_search(vs, dk, 1.0, lambda m: 1.0, -0.05, 0.0, 1e9, 0.5)
    assert np.isclose(rec["alpha"], np.linalg.norm(vs) / np.linalg.norm(dk) * -0.05, rtol=1e-14)
    # a flat misfit makes the parabola 0/0: NaN as in MATLAB (not an error)
    assert np.isnan(rec["gama"]) and rec["forwards"] == 3


def test_parabola_step_hits_the_minimum_of_a_quadratic_misfit():
    vs, dk = _setup()
    alpha = 2.0
    for fstar in (0.3, 0.9, 1.7):
        mis = lambda m, fs=fstar: (_factor(vs, m, dk, alpha) - fs) ** 2 + 0.25
        res0 = mis(vs.copy())
        vs1, rec = O.swd_line_search(vs, dk, res0, mis, -0.05, 0.0, 1e9, 0.5, alpha=alpha)
        assert np.isclose(rec["gama"], fstar, rtol=1e-9), (fstar, rec)
        assert np.allclose(vs1, vs + (alpha * fstar) * dk, rtol=1e-12)
        assert np.isclose(rec["res0"], 0.25, rtol=1e-9)


def test_doubling_when_the_first_trial_improves():
    vs, dk = _setup()
    mis = lambda m: (_factor(vs, m, dk, 1.0) - 3.0) ** 2
    _, rec = O.swd_line_search(vs, dk, 9.0, mis, -0.05, 0.0, 1e9, 0.5, alpha=1.0)
    assert rec["f1"] == 0.5 and rec["f2"] == 1.0      # f2 = 2 f1, one extra trial
    assert rec["forwards"] == 3                       # f1, f2, parabola


def test_halving_while_the_misfit_grows():
    vs, dk = _setup()
    # misfit rises for f > 0.05: trials at 0.5, 0.25, 0.125, 0.0625 fail, 0.03125 succeeds
    mis = lambda m: 1.0 + (_factor(vs, m, dk, 1.0) - 0.04) * 10.0
    res0 = 1.0
    _, rec = O.swd_line_search(vs, dk, res0, mis, -0.05, 0.0, 1e9, 0.5, alpha=1.0)
    assert np.isclose(rec["f1"], 0.03125) and np.isclose(rec["f2"], 0.0625)
    assert rec["forwards"] == 6                       # 5 halving trials + the parabola trial


def test_fallback_to_the_better_trial():
    vs, dk = _setup()
    # misfit with a kink: the parabola step lands where the misfit is high
    def mis(m):
        f = _factor(vs, m, dk, 1.0)
        return 10.0 if 0.6 < f < 0.9 else (f - 1.0) ** 2
    _, rec = O.swd_line_search(vs, dk, 1.0, mis, -0.05, 0.0, 1e9, 0.5, alpha=1.0)
    assert rec["res3"] > min(rec["res1"], rec["res2"])
    assert rec["gama"] in (rec["f1"], rec["f2"])
    assert rec["res0"] == min(rec["res1"], rec["res2"])


def test_step_model_clips_and_keeps_nan():
    vs = np.array([[100.0, 300.0, 500.0, np.nan]])
    dk = np.array([[-1e3, 0.0, 1e3, 1.0]])
    v = O.swd_step_model(vs, dk, 1.0, 1.0, 150.0, 450.0)
    assert v[0, 0] == 150.0 and v[0, 1] == 300.0 and v[0, 2] == 450.0 and np.isnan(v[0, 3])
