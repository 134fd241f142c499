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


def test_served_request_vs_oracle():
    """One served request (3 shots of a 30 x 120 model, nt 1500) against the
    FP64 oracle's restatement of SWD.m:171-201.  The served result is bitwise
    the request solved alone; where the FP32 picks equal the FP64 picks its
    dk_vs is within 1e-4 of the oracle's, otherwise the request's chain after
    the picks is judged from the oracle's picks (a pick one bin away changes
    the gradient by construction)."""
    from oracle import swd_oracle as O
    nz, nx, dt, nt = 30, 120, 1e-4, 1500
    vs = np.tile(np.linspace(220, 420, nz)[:, None], (1, nx))
    vs_true = vs * (1 + 0.04 * P.checkerboard(nz, nx, 15, 1.0, 1.0))
    q = dict(nz=nz, nx=nx, nbc=10, fd_order=22, nt=nt, in_wf=5, dx=0.5, dt=dt, sx=np.array([21, 61, 101]),
             sz=np.array([1, 1, 1]), ng=120, gx0=1, dg=1, gz=1, df=0.5, fmin=5.0, fmax=80.0, vmin=100.0,
             vmax=800.0, ds=0.5, offset=120, m=0, M=1, ini=10, w=0.0, beta=10.0, ns=120,
             is_index=np.array([21, 61, 101]))
    s = O.ricker(30, dt, nt)
    gx = np.arange(1, 121, dtype=float)
    gz = np.ones(120)
    radon_t = (q["df"], dt, 701, q["vmin"], q["vmax"], q["fmin"], q["fmax"], nt, 120, q["ds"], q["offset"], 0, 1)
    radon = dict(df=q["df"], fmin=q["fmin"], fmax=q["fmax"], vmin=q["vmin"], vmax=q["vmax"], ds=q["ds"],
                 offset=q["offset"], m=0, M=1)
    cr0_r, cr0_l, shots, tot = [], [], [], 0
    for k, sx in enumerate(q["sx"]):
        _, sw, _ = O.staggerfd_eigend(int(sx), 10, nt, dt / 0.5, 0.5, dt, int(sx), 1, gx, gz, s, 1.73 * vs_true,
                                      vs_true, 1, 22, "w", 0, 5, snapshots=False)
        cr0_r.append(O.pickFDC1(O.rt_side(sw, "r", int(sx), *radon_t)[0], q["ini"], 2, 4) + q["vmin"])
        cr0_l.append(O.pickFDC1(O.rt_side(sw, "l", int(sx), *radon_t)[0], q["ini"], 2, 4) + q["vmin"])
        o = O.swd_shot(1.73 * vs, vs, s, 10, nt, 0.5, dt, int(sx), 1, gx, gz, 5, 22, int(sx), 120, radon,
                       cr0_r[-1], cr0_l[-1], q["beta"], q["ini"], 0.0)
        shots.append(o)
        tot = tot + np.array([o["cl_img"], o["cm_img"], o["illum"]])
    cr0_r, cr0_l = np.array(cr0_r).T, np.array(cr0_l).T
    dk_ref = O.assemble_dk_vs(vs, *tot)
    srv = Server(hbm_bytes=4 << 30, max_active=4, nt_block=(5, 5))
    try:
        t = srv.submit(q, 1.73 * vs, vs, s, cr0_r, cr0_l)
        r = srv.wait(t)
    finally:
        srv.close()
    e = Engine(q)
    try:
        e.set_wavelet(s)
        e.set_observed_picks(cr0_r, cr0_l)
        e.set_model(1.73 * vs, vs)
        g = e.gradient()
        np.testing.assert_array_equal(r["dk_vs"], g["dk_vs"])
        np.testing.assert_array_equal(r["res_r"], g["res_r"])
        ocr_r = np.array([o["cr_r"] for o in shots]).T
        ocr_l = np.array([o["cr_l"] for o in shots]).T
        assert np.max(np.abs(g["cr_r"] - ocr_r)) <= 1 and np.max(np.abs(g["cr_l"] - ocr_l)) <= 1
        lit = tot[2] > 1e-60
        same = np.array_equal(g["cr_r"], ocr_r) and np.array_equal(g["cr_l"], ocr_l)
        if same:
            dk = r["dk_vs"]
            np.testing.assert_allclose(r["res_r"], [o["res_r"] for o in shots], rtol=1e-9)
        else:
            e.run("forward")
            e.run("dispersion")
            e.set_picks(ocr_r, ocr_l)
            e.run("adjoint_source")
            e.run("adjoint")
            e.run("assemble")
            dk = e.read("gradient")[0]
    finally:
        e.close()
    err = float(np.linalg.norm(dk[lit] - dk_ref[lit]) / np.linalg.norm(dk_ref[lit]))
    print(f"served request: picks equal to FP64: {same}; dk_vs rel-L2 vs FP64 over lit cells {err:.2e}"
          + ("" if same else " (from the oracle's picks)"))
    assert np.isfinite(dk[lit]).all() and err < 1e-4


def test_cancelled_requests_leave_the_others_unchanged(templates):
    """A long request cancelled while on the GPU and one cancelled while
    queued: both end with SWD_ERR_BUSY, every other result is bitwise its
    result alone, and every buffer returns to the arena."""
    import time
    trace = poisson_trace(templates, rate=5000.0, n=6, seed=5, max_shots=6)
    s = Server(hbm_bytes=8 << 30, max_active=2, nt_block=(4, 4))
    q0, mod0, cr_r0, cr_l0 = templates[0]
    big = dict(q0)
    big["nt"] = q0["nt"] * 8      # runs long enough to be cancelled on the GPU
    tickets = [s.submit(big, mod0["vp"], mod0["vs"], np.tile(mod0["s"], 8), cr_r0, cr_l0)]
    for ta, ti, k in trace[1:]:
        q, mod, cr_r, cr_l = templates[ti]
        tickets.append(s.submit(subset(q, k), mod["vp"], mod["vs"], mod["s"], cr_r[:, :k], cr_l[:, :k]))
    s.cancel(tickets[5])           # still queued behind two slots
    time.sleep(0.03)
    s.cancel(tickets[0])           # on the GPU
    results = {}
    for i, t in enumerate(tickets):
        try:
            results[i] = s.wait(t)
        except SwdError as ex:
            results[i] = ex
    st = s.stats()
    s.close()
    assert isinstance(results[5], SwdError) and results[5].code == -8, results[5]
    assert isinstance(results[0], SwdError) and results[0].code == -8, results[0]
    assert st["bytes_in_use"] == 0 and st["active"] == 0
    for i in range(1, 5):
        r = results[i]
        assert not isinstance(r, Exception), r
        g, _ = _alone(templates[trace[i][1]], trace[i][2])
        np.testing.assert_array_equal(r["dk_vs"], g["dk_vs"])
        np.testing.assert_array_equal(r["res_r"], g["res_r"])


def test_destroy_with_requests_queued_and_in_admission(templates):
    """Shutting the server down while requests are queued, being admitted and
    running: queued ones fail, the process carries on, a new server works."""
    trace = poisson_trace(templates, rate=5000.0, n=8, seed=6, max_shots=6)
    for delay in (0.0, 0.005, 0.05):
        s = Server(hbm_bytes=8 << 30, max_active=2, nt_block=(4, 4))
        for ta, ti, k in trace:
            q, mod, cr_r, cr_l = templates[ti]
            s.submit(subset(q, k), mod["vp"], mod["vs"], mod["s"], cr_r[:, :k], cr_l[:, :k])
        if delay:
            import time
            time.sleep(delay)
        s.close()
    s = Server(hbm_bytes=8 << 30, max_active=2, nt_block=(4, 4))
    out = run_trace(s, templates, trace[:2])
    s.close()
    _check_same(templates, trace[:2], out)


def test_failed_launch_table_fails_the_admitted_requests_only(templates, monkeypatch):
    """A launch-table rebuild that fails (test hook SWD_SERVER_FAIL_REBUILD)
    fails the requests on the GPU and in the ready list, never one still being
    admitted; later requests are served normally."""
    monkeypatch.setenv("SWD_SERVER_FAIL_REBUILD", "3")
    s = Server(hbm_bytes=8 << 30, max_active=3, nt_block=(4, 4))
    monkeypatch.delenv("SWD_SERVER_FAIL_REBUILD")
    trace = poisson_trace(templates, rate=5000.0, n=6, seed=7, max_shots=6)
    tickets = []
    for ta, ti, k in trace:
        q, mod, cr_r, cr_l = templates[ti]
        tickets.append(s.submit(subset(q, k), mod["vp"], mod["vs"], mod["s"], cr_r[:, :k], cr_l[:, :k]))
    ok, failed = [], 0
    for i, t in enumerate(tickets):
        try:
            ok.append((i, s.wait(t)))
        except SwdError as ex:
            assert ex.code == -2, ex
            failed += 1
    st = s.stats()
    s.close()
    assert failed >= 1 and st["bytes_in_use"] == 0 and st["active"] == 0
    for i, r in ok:
        g, _ = _alone(templates[trace[i][1]], trace[i][2])
        np.testing.assert_array_equal(r["dk_vs"], g["dk_vs"])
