"""The continuous-batching gradient server: every request's result is the
result of the same solve alone (bitwise), while requests of different sizes
overlap on the device; admission respects the HBM budget."""
import numpy as np
import pytest

from swd_amd import problems as P
from swd_amd._lib import SwdError
from swd_amd.engine import Engine
from swd_amd.serve import Server, poisson_trace, request_templates, run_trace, subset

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def templates():
    models = P.uneven_models(seed=3, shots=4, sizes=(60, 90, 120), nt_range=(600, 800), f0=40.0)
    return request_templates(models, shots=6)


def _alone(tpl, k):
    q, mod, cr_r, cr_l = tpl
    e = Engine(subset(q, k))
    e.set_wavelet(mod["s"])
    e.set_observed_picks(cr_r[:, :k], cr_l[:, :k])
    e.set_model(mod["vp"], mod["vs"])
    g = e.gradient()
    foot = e.info()["device_bytes"]
    e.close()
    return g, foot


def _check_same(templates, trace, out):
    for (ta, ti, k), r in zip(trace, out["results"]):
        g, _ = _alone(templates[ti], k)
        np.testing.assert_array_equal(r["dk_vs"], g["dk_vs"])
        np.testing.assert_array_equal(r["res_r"], g["res_r"])
        np.testing.assert_array_equal(r["res_l"], g["res_l"])


def test_mixed_requests_equal_alone(templates):
    trace = poisson_trace(templates, rate=5000.0, n=10, seed=1, max_shots=6)
    s = Server(hbm_bytes=8 << 30, max_active=4, nt_block=(5, 4))
    out = run_trace(s, templates, trace)
    st = s.stats()
    s.close()
    print({k: v for k, v in out.items() if k != "results"}, st)
    assert st["completed"] == 10 and st["failed"] == 0
    assert st["peak_active"] >= 2            # requests overlapped on the device
    assert st["bytes_in_use"] == 0           # every request's buffers went back to the arena
    _check_same(templates, trace, out)


def test_hbm_budget_limits_admission(templates):
    trace = poisson_trace(templates, rate=5000.0, n=6, seed=2, max_shots=6)
    foots = [_alone(templates[ti], k)[1] for _, ti, k in trace]
    budget = int(1.2 * max(foots))
    s = Server(hbm_bytes=budget, max_active=8, nt_block=(4, 4))
    out = run_trace(s, templates, trace)
    st = s.stats()
    assert st["completed"] == 6 and st["failed"] == 0
    assert st["peak_bytes"] <= st["budget"] <= budget + 256
    # a request larger than the whole budget fails alone; the server goes on
    q, mod, cr_r, cr_l = templates[int(np.argmax([t[0]["nx"] for t in templates]))]
    big = dict(q)
    big["nt"] = q["nt"] * 8
    t = s.submit(big, mod["vp"], mod["vs"], np.tile(mod["s"], 8), cr_r, cr_l)
    with pytest.raises(SwdError):
        s.wait(t)
    tr2 = trace[:2]
    out2 = run_trace(s, templates, tr2)
    s.close()
    _check_same(templates, trace, out)
    _check_same(templates, tr2, out2)


def test_one_request_at_a_time(templates):
    trace = poisson_trace(templates, rate=5000.0, n=4, seed=3, max_shots=6)
    s = Server(hbm_bytes=8 << 30, max_active=1)
    out = run_trace(s, templates, trace)
    st = s.stats()
    s.close()
    assert st["peak_active"] == 1 and st["completed"] == 4
    _check_same(templates, trace, out)


def test_paused_requests_resume_exactly(templates):
    """With room for one request's shots per launch the scheduler pauses the
    others (least remaining work runs first, long-paused ones rejoin); paused
    and resumed wavefields give the same results as uninterrupted solves."""
    trace = poisson_trace(templates, rate=5000.0, n=6, seed=4, max_shots=6)
    s = Server(hbm_bytes=8 << 30, max_active=6, nt_block=(3, 5))
    s.set_wave_slots(1)
    out = run_trace(s, templates, trace)
    st = s.stats()
    s.close()
    assert st["completed"] == 6 and st["peak_active"] >= 2
    assert st["rebuilds"] > 6                # selections changed while requests were in flight
    _check_same(templates, trace, out)
