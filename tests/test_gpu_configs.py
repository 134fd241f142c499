"""The benchmark workloads at their stated sizes, checked against the FP64
oracle where the oracle finishes in seconds and through size-independent
invariants elsewhere.

- Production shots (60 shots, 100 x 600 cells, 132 x 660 padded, nt 6000,
  600 receivers, Radon 151 x 701): the whole gradient of SWD.m:171-201.  The
  middle shot's observed picks (SWD.m:150-162), picks, adjoint source and
  per-shot images against the oracle's FP64 restatement of that shot; the
  shot's images alone equal its images inside the 60-shot batch, bitwise; a
  second launch shape gives the same images for all 60 shots, bitwise.
- Uneven batch (8 faulted layered models, nx 200..1200, nz 40..150, nt
  2000..8000, 16 shots each, autotuned launch shape): every model's gradient
  is bitwise its gradient alone, and shot 1 of every model is within 1e-4 of
  the FP64 oracle's gather.
- Serving (Poisson arrivals at 10 requests/s, requests of 1-32 shots of those
  models): every served result is bitwise the result of the request alone.
- Dispersion path alone (60 chirp gathers of 6000 x 600): picks within one
  velocity bin of the FP64 oracle's, adjoint sources at beta 10 within 1e-4.
- The reference's CPU-runnable single shot (60 x 200 cells, 82 x 240 padded,
  nt 3000, 199 receivers): observed picks, gather, picks, misfit, adjoint
  source and dk_vs against the oracle (the chain after the picks from the
  oracle's picks wherever an FP32 pick lands one bin away).

FP32 tolerance: relative L2 <= 1e-4 against FP64 for gathers, adjoint
sources, images and gradients; picks within one velocity bin (1 m/s)."""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import problems as P
from swd_amd.dispersion import DispersionBatch
from swd_amd.engine import Engine, batch_autotune, batch_gradient
from swd_amd.serve import Server, poisson_trace, request_templates, run_trace, subset

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def rel(a, b):
    a, b = np.ravel(a).astype(np.float64), np.ravel(b).astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _radon_tuple(q, np_):
    return (q["df"], q["dt"], np_, q["vmin"], q["vmax"], q["fmin"], q["fmax"], q["nt"], q["ng"], q["ds"],
            q["offset"], q["m"], q["M"])


def _oracle_observed_shot(q, mod, k):
    """FP64 observed picks of shot k (SWD.m:150-162) from the true model."""
    is_ = int(q["is_index"][k]) if q.get("is_index") is not None else k + 1
    gx = q["gx0"] + q["dg"] * np.arange(q["ng"], dtype=float)
    _, sw, _ = O.staggerfd_eigend(is_, q["nbc"], q["nt"], q["dt"] / q["dx"], q["dx"], q["dt"], int(q["sx"][k]),
                                  int(q["sz"][k]), gx, np.full(q["ng"], float(q["gz"])), mod["s"], mod["vp_true"],
                                  mod["vs_true"], 1, 22, "w", 0, q["in_wf"], snapshots=False)
    np_ = int(q["vmax"] - q["vmin"] + 1)
    out = []
    for side in ("r", "l"):
        ml, _, _ = O.rt_side(sw, side, is_, *_radon_tuple(q, np_))
        out.append(O.pickFDC1(ml, q["ini"], 2, 4) + q["vmin"])
    return out


def _oracle_shot(q, mod, k, cr0_r, cr0_l):
    is_ = int(q["is_index"][k]) if q.get("is_index") is not None else k + 1
    ns = q["ns"] if q["ns"] else int(np.size(q["sx"]))
    gx = q["gx0"] + q["dg"] * np.arange(q["ng"], dtype=float)
    radon = dict(df=q["df"], fmin=q["fmin"], fmax=q["fmax"], vmin=q["vmin"], vmax=q["vmax"], ds=q["ds"],
                 offset=q["offset"], m=q["m"], M=q["M"])
    return O.swd_shot(mod["vp"], mod["vs"], mod["s"], q["nbc"], q["nt"], q["dx"], q["dt"], int(q["sx"][k]),
                      int(q["sz"][k]), gx, np.full(q["ng"], float(q["gz"])), q["in_wf"], 22, is_, ns, radon, cr0_r,
                      cr0_l, q["beta"], q["ini"], q["w"])


def _stages(e, cr_r, cr_l):
    e.run("forward")
    e.run("dispersion")
    e.set_picks(cr_r, cr_l)
    e.run("adjoint_source")
    e.run("adjoint")
    e.run("assemble")


# ------------------------------------------------------------ production shots
SHOT = 30          # sx = 301, the middle of the line


@pytest.fixture(scope="module")
def production():
    prob, mod = P.production_shots()
    assert prob["sx"].size == 60 and prob["nt"] == 6000 and prob["ng"] == 600
    print(f"[production] 60 shots x nt {prob['nt']}: observed picks on the GPU", flush=True)
    e = Engine(prob)
    e.set_wavelet(mod["s"])
    e.set_model(mod["vp_true"], mod["vs_true"])
    cr0_r, cr0_l = e.observe()
    print("[production] FP64 oracle: observed picks and the gradient of shot", SHOT, flush=True)
    obs_r, obs_l = _oracle_observed_shot(prob, mod, SHOT)
    gpu_obs = (cr0_r[:, SHOT].copy(), cr0_l[:, SHOT].copy())
    cr0_r[:, SHOT], cr0_l[:, SHOT] = obs_r, obs_l
    ref = _oracle_shot(prob, mod, SHOT, obs_r, obs_l)
    e.set_observed_picks(cr0_r, cr0_l)
    e.set_model(mod["vp"], mod["vs"])
    print("[production] GPU gradient", flush=True)
    g = e.gradient()
    yield prob, mod, e, g, ref, (obs_r, obs_l), gpu_obs, (cr0_r, cr0_l)
    e.close()


def test_production_observed_picks_vs_oracle(production):
    prob, mod, e, g, ref, obs, gpu_obs, _ = production
    for side, a, b in (("r", gpu_obs[0], obs[0]), ("l", gpu_obs[1], obs[1])):
        d = np.abs(a - b)
        print(f"observed picks shot {SHOT + 1} side {side}: {int(np.count_nonzero(d))} of {d.size} differ, "
              f"max {d.max():g} m/s")
        assert d.max() <= 1.0


def test_production_gradient_shot_vs_oracle(production):
    """The middle shot of the full gradient: picks within one bin; given the
    oracle's picks, adjoint source and per-shot images within 1e-4."""
    prob, mod, e, g, ref, obs, _, _ = production
    dr = np.abs(g["cr_r"][:, SHOT] - ref["cr_r"])
    dl = np.abs(g["cr_l"][:, SHOT] - ref["cr_l"])
    print(f"picks shot {SHOT + 1}: right {int(np.count_nonzero(dr))} / left {int(np.count_nonzero(dl))} of "
          f"{dr.size} differ, max {max(dr.max(), dl.max()):g} m/s")
    assert dr.max() <= 1.0 and dl.max() <= 1.0
    if not dr.any():
        assert g["res_r"][SHOT] == pytest.approx(ref["res_r"], rel=1e-9)
    cr_r, cr_l = g["cr_r"].copy(), g["cr_l"].copy()
    cr_r[:, SHOT], cr_l[:, SHOT] = ref["cr_r"], ref["cr_l"]
    _stages(e, cr_r, cr_l)
    gath = e.read("gathers")[SHOT]
    adj = e.read("adjoint_sources")[SHOT]
    img = e.read("images")[SHOT]
    err = {"gather": rel(gath, ref["seismo_w"]), "adjoint_source": rel(adj, ref["adj"]),
           "cl_img": rel(img[0], ref["cl_img"]), "cm_img": rel(img[1], ref["cm_img"]),
           "illum": rel(img[2], ref["illum"])}
    print(f"production shot {SHOT + 1}, FP32 rel-L2 vs FP64 given the oracle's picks:",
          {k: f"{v:.2e}" for k, v in err.items()})
    assert max(err.values()) < 1e-4, err


def test_production_images_alone_equal_in_batch_and_launch_shape(production):
    prob, mod, e, g, ref, obs, _, (cr0_r, cr0_l) = production
    cr_r, cr_l = g["cr_r"].copy(), g["cr_l"].copy()
    cr_r[:, SHOT], cr_l[:, SHOT] = ref["cr_r"], ref["cr_l"]
    e.set_time_block(5, 26)
    _stages(e, cr_r, cr_l)
    imgs = e.read("images")
    one = dict(prob)
    one.update(sx=prob["sx"][SHOT:SHOT + 1].copy(), sz=prob["sz"][SHOT:SHOT + 1].copy(),
               is_index=np.array([SHOT + 1]), ns=60)
    a = Engine(one)
    try:
        a.set_wavelet(mod["s"])
        a.set_observed_picks(cr0_r[:, SHOT:SHOT + 1], cr0_l[:, SHOT:SHOT + 1])
        a.set_model(mod["vp"], mod["vs"])
        _stages(a, cr_r[:, SHOT:SHOT + 1], cr_l[:, SHOT:SHOT + 1])
        np.testing.assert_array_equal(a.read("images")[0], imgs[SHOT])
    finally:
        a.close()
    e.set_time_block(4, 33)
    _stages(e, cr_r, cr_l)
    np.testing.assert_array_equal(e.read("images"), imgs)
    e.set_time_block(5, 26)


# ------------------------------------------------------------ uneven batch
def test_uneven_batch_full_size():
    models = P.uneven_models()
    assert len(models) == 8 and all(np.size(p["sx"]) == 16 for p, _ in models)
    print("[uneven] 8 models:", [(p["nz"], p["nx"], p["nt"]) for p, _ in models], flush=True)
    with cf.ThreadPoolExecutor(THREADS) as ex:
        futs = []
        for prob, mod in models:
            gx = prob["gx0"] + prob["dg"] * np.arange(prob["ng"], dtype=float)
            futs.append(ex.submit(O.staggerfd_eigend, 1, prob["nbc"], prob["nt"], prob["dt"] / prob["dx"],
                                  prob["dx"], prob["dt"], int(prob["sx"][0]), 1, gx, np.ones(prob["ng"]), mod["s"],
                                  mod["vp"], mod["vs"], 1, 22, "w", 0, prob["in_wf"], snapshots=False))
        engs = []
        try:
            for prob, mod in models:
                e = Engine(prob)
                engs.append(e)
                e.set_wavelet(mod["s"])
                e.set_model(mod["vp_true"], mod["vs_true"])
                e.observe()
                e.set_model(mod["vp"], mod["vs"])
            tune = batch_autotune(engs, steps=60)
            print("[uneven] autotuned", tune["fwd"], tune["adj"], flush=True)
            outs, us = batch_gradient(engs)
            gath = [e.read("gathers")[0] for e in engs]
            for (prob, mod), e, m in zip(models, engs, outs):
                np.testing.assert_array_equal(m, e.gradient()["dk_vs"])
        finally:
            for e in engs:
                e.close()
        refs = [f.result()[1] for f in futs]
    for (prob, _), gg, sw in zip(models, gath, refs):
        err = rel(gg, sw)
        print(f"[uneven] {prob['nz']} x {prob['nx']} x nt {prob['nt']}: shot 1 gather rel-L2 vs FP64 {err:.2e}")
        assert err < 1e-4


# ------------------------------------------------------------ serving
def test_serving_trace_full_size():
    """32 requests over the uneven models' size range at 10 requests/s."""
    templates = request_templates(P.uneven_models(), 32)
    trace = poisson_trace(templates, rate=10.0, n=32, seed=10)
    s = Server(max_active=16, nt_block=(5, 5))
    try:
        out = run_trace(s, templates, trace)
        st = s.stats()
    finally:
        s.close()
    print("[serving]", {k: round(v, 2) for k, v in out.items() if k != "results"}, st, flush=True)
    assert st["completed"] == 32 and st["failed"] == 0 and st["peak_active"] >= 2
    for (ta, ti, k), r in zip(trace, out["results"]):
        q, mod, cr_r, cr_l = templates[ti]
        e = Engine(subset(q, k))
        try:
            e.set_wavelet(mod["s"])
            e.set_observed_picks(cr_r[:, :k], cr_l[:, :k])
            e.set_model(mod["vp"], mod["vs"])
            g = e.gradient()
        finally:
            e.close()
        np.testing.assert_array_equal(r["dk_vs"], g["dk_vs"])
        np.testing.assert_array_equal(r["res_r"], g["res_r"])
        np.testing.assert_array_equal(r["res_l"], g["res_l"])


# ------------------------------------------------------------ dispersion path alone
def test_dispersion_60_shots_vs_oracle():
    prob, _ = P.production_shots(nshots=60)
    q = prob
    g = P.chirp_gathers(nshots=60, nt=q["nt"], ng=q["ng"], dt=q["dt"], dx=q["ds"], sx=q["sx"])
    D = DispersionBatch(q)
    try:
        D.set_gathers(g)
        D.set_observed_picks(np.ones((D.npair, 60)), np.ones((D.npair, 60)))
        picks = D.pick()
        r = np.random.default_rng(60)
        gr = np.round(r.normal(0, 5, (D.npair, 60)))
        gl = np.round(r.normal(0, 5, (D.npair, 60)))
        D.set_observed_picks(picks["cr_r"] - gr, picks["cr_l"] - gl)
        adj, ms = D.adjoint(10.0)
        np_ = D.np
        npair = D.npair
    finally:
        D.close()

    def oracle(s):
        sides = {}
        for side in ("r", "l"):
            ml, _, sfb = O.rt_side(g[s].astype(np.float64), side, s + 1, *_radon_tuple(q, np_))
            _, cr = O.LHDispPick(ml, q["vmin"], np.ones(npair), 1, q["ini"])
            sides[side] = (cr, sfb)
        a = O.ADWDgrad_1(q["nt"], q["ng"], 60, npair, s + 1, q["w"], q["m"], q["M"], 10.0, gr[:, s], gl[:, s],
                         sides["r"][1], sides["l"][1])
        return sides, a

    print("[dispersion] FP64 oracle on 120 windows", flush=True)
    with cf.ThreadPoolExecutor(THREADS) as ex:
        refs = list(ex.map(oracle, range(60)))
    worst_adj, n_cols, n_within = 0.0, 0, 0
    for s, (sides, a) in enumerate(refs):
        for side, key in (("r", "cr_r"), ("l", "cr_l")):
            cr, sfb = sides[side]
            if len(sfb["uxtpos"]) <= 1:
                continue  # a one-trace window has a flat spectrum
            d = np.abs(picks[key][:, s] - cr)
            n_cols += d.size
            n_within += int(np.count_nonzero(d <= 1.0))
            assert d.max() <= 1.0, (s, side, len(sfb["uxtpos"]), d.max())
        worst_adj = max(worst_adj, rel(adj[s], a))
    print(f"[dispersion] 60 shots: {n_within} of {n_cols} picks within one bin; worst adjoint-source rel-L2 "
          f"{worst_adj:.2e}; adjoint batch {ms:.2f} ms")
    assert worst_adj < 1e-4


def test_single_shot_workload_full_size_vs_oracle():
    """The single-shot workload at its stated size (nt 3000): the device's
    observed picks (SWD.m:150-162) within one bin of the oracle's; the
    gather, the adjoint source and dk_vs within 1e-4 of FP64; picks within
    one bin, and where any pick differs the stages after the picks are judged
    from the oracle's picks."""
    prob, mod = P.plain_single_shot(nt=3000)
    q = prob
    gx = np.arange(q["gx0"], q["gx0"] + q["ng"], dtype=float)
    gz = np.full(q["ng"], float(q["gz"]))
    radon = (q["df"], q["dt"], 701, q["vmin"], q["vmax"], q["fmin"], q["fmax"], q["nt"], q["ng"], q["ds"],
             q["offset"], q["m"], q["M"])
    _, obs, _ = O.staggerfd_eigend(int(q["is_index"][0]), q["nbc"], q["nt"], q["dt"] / q["dx"], q["dx"], q["dt"],
                                   int(q["sx"][0]), int(q["sz"][0]), gx, gz, mod["s"], mod["vp_true"],
                                   mod["vs_true"], 1, 22, "w", 0, q["in_wf"], snapshots=False)
    cr0_r = O.pickFDC1(O.rt_side(obs, "r", 5, *radon)[0], q["ini"], 2, 4) + q["vmin"]
    cr0_l = O.pickFDC1(O.rt_side(obs, "l", 5, *radon)[0], q["ini"], 2, 4) + q["vmin"]
    rad = dict(df=q["df"], fmin=q["fmin"], fmax=q["fmax"], vmin=q["vmin"], vmax=q["vmax"], ds=q["ds"],
               offset=q["offset"], m=q["m"], M=q["M"])
    o = O.swd_shot(mod["vp"], mod["vs"], mod["s"], q["nbc"], q["nt"], q["dx"], q["dt"], int(q["sx"][0]),
                   int(q["sz"][0]), gx, gz, q["in_wf"], 22, 5, 5, rad, cr0_r, cr0_l, q["beta"], q["ini"], q["w"])
    dk_ref = O.assemble_dk_vs(mod["vs"], o["cl_img"], o["cm_img"], o["illum"])

    e = Engine(prob)
    e.set_wavelet(mod["s"])
    e.set_model(mod["vp_true"], mod["vs_true"])
    a, b = e.observe()
    d_obs = max(np.max(np.abs(a[:, 0] - cr0_r)), np.max(np.abs(b[:, 0] - cr0_l)))
    e.set_model(mod["vp"], mod["vs"])
    e.set_observed_picks(cr0_r[:, None], cr0_l[:, None])
    g = e.gradient()
    err = {"gather": rel(e.read("gathers")[0], o["seismo_w"])}
    dr = np.max(np.abs(g["cr_r"][:, 0] - o["cr_r"]))
    dl = np.max(np.abs(g["cr_l"][:, 0] - o["cr_l"]))
    if dr == 0 and dl == 0:
        err["adjoint_source"] = rel(e.read("adjoint_sources")[0], o["adj"])
        err["dk_vs"] = rel(g["dk_vs"], dk_ref)
        assert g["res_r"][0] == pytest.approx(o["res_r"], rel=1e-9)
    else:
        e.run("forward")
        e.run("dispersion")
        e.set_picks(o["cr_r"][:, None], o["cr_l"][:, None])
        e.run("adjoint_source")
        e.run("adjoint")
        e.run("assemble")
        err["adjoint_source_given_oracle_picks"] = rel(e.read("adjoint_sources")[0], o["adj"])
        err["dk_vs_given_oracle_picks"] = rel(e.read("gradient")[0], dk_ref)
    e.close()
    print(f"single shot nt 3000: observed picks max diff {d_obs:g} m/s, picks {dr:g} / {dl:g} m/s, rel-L2 {err}")
    assert d_obs <= 1.0 and dr <= 1 and dl <= 1
    assert max(err.values()) < 1e-4, err
