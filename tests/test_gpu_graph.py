"""FD phases replayed as captured HIP graphs (swd_engine_set_graph) give the
bitwise results of the eager launches: on the first replay (capture), on
later replays with eager work and host synchronisations in between, after a
model change (same buffers, same graph) and after a launch-shape change (new
capture)."""
import numpy as np
import pytest

from swd_amd import problems as P
from swd_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _same(g, ref, gath, gath_ref):
    np.testing.assert_array_equal(gath, gath_ref)
    for k in ("dk_vs", "cr_r", "cr_l", "res_r", "res_l"):
        np.testing.assert_array_equal(g[k], ref[k])


def test_graph_replays_match_eager():
    prob, mod = P.production_shots(nshots=3, nt=1500)
    e = Engine(prob)
    e.set_wavelet(mod["s"])
    e.set_model(mod["vp_true"], mod["vs_true"])
    e.observe()
    vs2 = mod["vs"] * np.linspace(0.97, 1.03, prob["nx"])[None, :]
    e.set_time_block(5, 0)
    e.set_model(mod["vp"], mod["vs"])
    ref = e.gradient()
    gref = e.read("gathers")
    e.set_model(1.73 * vs2, vs2)
    ref2 = e.gradient()
    gref2 = e.read("gathers")
    e.set_time_block(3, 24)
    ref3 = e.gradient()
    e.set_time_block(5, 0)

    e.set_model(mod["vp"], mod["vs"])
    e.gradient()
    t_eager = e.timings()
    assert e.set_graph(True) is False
    for rep in range(3):
        g = e.gradient()
        _same(g, ref, e.read("gathers"), gref)
        t = e.timings()
        assert t["fwd_steps_per_launch"] == 5 and t["adj_steps_per_launch"] == 5
        # launch durations sampled by event records inside the graph
        for k in ("fwd_kernel_us", "adj_kernel_us"):
            assert 0.5 * t_eager[k] < t[k] < 2.0 * t_eager[k], (k, t[k], t_eager[k])
        # eager work between replays
        assert e.set_graph(False) is True
        _same(e.gradient(), ref, e.read("gathers"), gref)
        e.set_graph(True)
    # a model change rewrites the coefficient planes in place: the captured graphs replay
    e.set_model(1.73 * vs2, vs2)
    for rep in range(2):
        _same(e.gradient(), ref2, e.read("gathers"), gref2)
    # a new launch shape captures new graphs
    e.set_time_block(3, 24)
    for rep in range(2):
        g = e.gradient()
        np.testing.assert_array_equal(g["dk_vs"], ref3["dk_vs"])
        assert e.timings()["fwd_steps_per_launch"] == 3
    e.set_time_block(5, 0)
    e.set_model(mod["vp"], mod["vs"])
    _same(e.gradient(), ref, e.read("gathers"), gref)
    e.close()


def test_batch_graph_replays_match_eager():
    """swd_batch_gradient over uneven models (different grids and nt) with
    the first engine's graph setting on and off: every model's gradient
    bitwise equal across repeated batches, with graph-replayed single-engine
    gradients of the first engine between them and after one model's change
    (time-blocked batches run the overlapped schedule, eagerly; the
    single-engine graphs stay valid around it)."""
    from swd_amd.engine import batch_gradient
    models = P.uneven_models(shots=2, sizes=(200, 300, 450), nt_range=(1500, 2500))
    engs = []
    for prob, mod in models:
        e = Engine(prob)
        e.set_wavelet(mod["s"])
        e.set_model(mod["vp_true"], mod["vs_true"])
        e.observe()
        e.set_model(mod["vp"], mod["vs"])
        e.set_time_block(5, 0)
        engs.append(e)
    ref, _ = batch_gradient(engs)
    ref0 = engs[0].gradient()["dk_vs"]
    prob1, mod1 = models[1]
    vs1 = mod1["vs"] * 1.02
    engs[1].set_model(1.73 * vs1, vs1)
    ref_b, _ = batch_gradient(engs)
    engs[1].set_model(mod1["vp"], mod1["vs"])

    assert engs[0].set_graph(True) is False
    for rep in range(3):
        out, us = batch_gradient(engs)
        for a, b in zip(out, ref):
            np.testing.assert_array_equal(a, b)
        assert us[0] > 0 and us[1] > 0
        # a single-engine gradient of the first engine between batch replays
        np.testing.assert_array_equal(engs[0].gradient()["dk_vs"], ref0)
        engs[0].set_graph(False)
        out, _ = batch_gradient(engs)
        for a, b in zip(out, ref):
            np.testing.assert_array_equal(a, b)
        engs[0].set_graph(True)
    engs[1].set_model(1.73 * vs1, vs1)
    for rep in range(2):
        out, _ = batch_gradient(engs)
        for a, b in zip(out, ref_b):
            np.testing.assert_array_equal(a, b)
    for e in engs:
        e.close()
