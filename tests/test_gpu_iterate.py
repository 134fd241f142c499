"""One SWD iteration on the device (swd_engine_iterate) against the oracle's
restatement of SWD.m:199-322 driving a second engine through the host path,
plus smooth2a on the GPU against the oracle."""
import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import problems as P
from swd_amd.engine import Engine
from swd_amd.inversion import invert, smooth2a

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nr,nc", [(1, 2), (0, 0), (3, 1)])
def test_smooth2a_vs_oracle(nr, nc):
    r = np.random.default_rng(nr * 10 + nc)
    A = r.standard_normal((37, 53))
    A[r.random(A.shape) < 0.05] = np.nan
    g = smooth2a(A, nr, nc)
    ref = O.smooth2a(A, nr, nc)
    assert np.array_equal(np.isnan(g), np.isnan(ref))
    m = ~np.isnan(ref)
    assert np.allclose(g[m], ref[m], rtol=1e-13, atol=0)


def _engine(prob, mod, cr0=None):
    e = Engine(prob)
    e.set_wavelet(mod["s"])
    if cr0 is None:
        e.set_model(mod["vp_true"], mod["vs_true"])
        cr0 = e.observe()
    else:
        e.set_observed_picks(*cr0)
    e.set_model(mod["vp"], mod["vs"])
    return e, cr0


def test_device_iteration_equals_host_restatement():
    prob, mod = P.plain_single_shot(nt=3000)
    vsmin, vsmax = float(mod["vs_true"].min()), float(mod["vs_true"].max())
    A, cr0 = _engine(prob, mod)
    r = A.iterate(vsmin, vsmax, pur=-0.05, smooth_z=1, smooth_x=2, state={"f1": 0.5})
    vs_a = A.model()
    dk_a = A.read("smoothed_gradient")
    assert np.isfinite(dk_a).all()

    # the same iteration on the host path of a second engine
    B, _ = _engine(prob, mod, cr0)
    g = B.gradient()
    res0 = float(g["res_r"][0])                    # one shot, right branch enabled
    assert res0 == r["residual"]
    ref = O.smooth2a(g["dk_vs"], 1, 2)
    assert np.allclose(dk_a, ref, rtol=1e-12, atol=0)
    vs0 = np.asfortranarray(mod["vs"], dtype=np.float64)
    alpha = np.sqrt(np.sum(vs0 * vs0)) / np.sqrt(np.sum(dk_a * dk_a)) * -0.05
    assert abs(r["alpha"] - alpha) <= 1e-13 * abs(alpha)

    def misfit(vs1):
        B.set_model(mod["vp"], vs1)               # host coefficient preparation
        return float(B.forward_picks()["res_r"][0])

    vs_b, rec = O.swd_line_search(vs0, dk_a, res0, misfit, -0.05, vsmin, vsmax, 0.5, alpha=r["alpha"])
    print("device:", {k: r[k] for k in ("residual", "res0", "alpha", "gama", "forwards")}, "host:", rec)
    assert rec["forwards"] == r["forwards"]
    assert rec["gama"] == r["gama"] and rec["res0"] == r["res0"] and rec["f1"] == r["f1"]
    np.testing.assert_array_equal(vs_a, vs_b)

    # device-built coefficient planes equal the host's: identical forward gathers
    A.forward_picks()
    B.set_model(mod["vp"], vs_a)
    B.forward_picks()
    np.testing.assert_array_equal(A.read("gathers"), B.read("gathers"))
    A.close()
    B.close()


def test_inversion_reduces_the_misfit():
    prob, mod = P.plain_single_shot(nt=3000)
    out = invert(prob, mod, iterations=3)
    res = out["residual"]
    print("residual history:", res, "records:", [(x["res0"], x["gama"], x["forwards"]) for x in out["records"]])
    assert np.all(np.isfinite(out["vs"]))
    assert out["records"][-1]["res0"] <= res[0]
