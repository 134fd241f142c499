"""GPU parity of the batched engine (full SWD shot-gradient) against the FP64
oracle, plus the batching invariants (a shot's result is the same alone and in
a batch of other models).  Gradient tolerance: relative L2 <= 1e-4 when the
FP32 picks equal the FP64 picks; picks within one velocity bin."""
import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import problems as P
from swd_amd.engine import Engine, batch_gradient

pytestmark = pytest.mark.gpu


def rel(a, b):
    """Relative L2 over finite entries; unilluminated cells are 0/0 in both
    (g_cl./g_illum, SWD.m:199) and must be NaN in the same places."""
    a, b = np.ravel(a), np.ravel(b)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = np.isfinite(b)
    return float(np.linalg.norm(a[m] - b[m]) / max(np.linalg.norm(b[m]), 1e-300))


def oracle_observed(prob, mod, is_list):
    """FP64 observed picks (SWD.m:150-162) from the true model."""
    q = prob
    gx = q["gx0"] + q["dg"] * np.arange(q["ng"], dtype=float)
    gz = np.full(q["ng"], float(q["gz"]))
    radon = (q["df"], q["dt"], int(q["vmax"] - q["vmin"] + 1), q["vmin"], q["vmax"], q["fmin"], q["fmax"], q["nt"],
             q["ng"], q["ds"], q["offset"], q["m"], q["M"])
    crr, crl = [], []
    for k, is_ in enumerate(is_list):
        _, sw, _ = O.staggerfd_eigend(is_, q["nbc"], q["nt"], q["dt"] / q["dx"], q["dx"], q["dt"], q["sx"][k],
                                      q["sz"][k], gx, gz, mod["s"], mod["vp_true"], mod["vs_true"], 1, 22, "w", 0,
                                      q["in_wf"], snapshots=False)
        ml, _, _ = O.rt_side(sw, "r", is_, *radon)
        crr.append(O.pickFDC1(ml, q["ini"], 2, 4) + q["vmin"])
        if is_ > 1 or q["M"] == 1:
            ml, _, _ = O.rt_side(sw, "l", is_, *radon)
            crl.append(O.pickFDC1(ml, q["ini"], 2, 4) + q["vmin"])
        else:
            crl.append(np.ones_like(crr[-1]))
    return np.array(crr).T, np.array(crl).T


def oracle_gradient(prob, mod, is_list, cr0_r, cr0_l):
    q = prob
    gx = q["gx0"] + q["dg"] * np.arange(q["ng"], dtype=float)
    gz = np.full(q["ng"], float(q["gz"]))
    radon = dict(df=q["df"], fmin=q["fmin"], fmax=q["fmax"], vmin=q["vmin"], vmax=q["vmax"], ds=q["ds"],
                 offset=q["offset"], m=q["m"], M=q["M"])
    ns = q["ns"] if q["ns"] else len(is_list)
    tot = None
    shots = []
    for k, is_ in enumerate(is_list):
        o = O.swd_shot(mod["vp"], mod["vs"], mod["s"], q["nbc"], q["nt"], q["dx"], q["dt"], q["sx"][k], q["sz"][k],
                       gx, gz, q["in_wf"], 22, is_, ns, radon, cr0_r[:, k], cr0_l[:, k], q["beta"], q["ini"], q["w"])
        shots.append(o)
        imgs = np.array([o["cl_img"], o["cm_img"], o["illum"]])
        tot = imgs if tot is None else tot + imgs
    return O.assemble_dk_vs(mod["vs"], *tot), shots, tot


@pytest.fixture(scope="module")
def single():
    prob, mod = P.plain_single_shot(nt=1500)
    cr0_r, cr0_l = oracle_observed(prob, mod, [5])
    dk, shots, _ = oracle_gradient(prob, mod, [5], cr0_r, cr0_l)
    return prob, mod, cr0_r, cr0_l, dk, shots


def _engine(prob, mod, cr0_r, cr0_l):
    e = Engine(prob)
    e.set_model(mod["vp"], mod["vs"])
    e.set_wavelet(mod["s"])
    e.set_observed_picks(cr0_r, cr0_l)
    return e


def test_engine_single_shot_gradient_vs_fp64_oracle(single):
    prob, mod, cr0_r, cr0_l, dk_ref, shots = single
    e = _engine(prob, mod, cr0_r, cr0_l)
    g = e.gradient()
    o = shots[0]
    gath = e.read("gathers")[0]
    err = {"gather": rel(gath, o["seismo_w"])}
    dr = np.max(np.abs(g["cr_r"][:, 0] - o["cr_r"]))
    dl = np.max(np.abs(g["cr_l"][:, 0] - o["cr_l"]))
    assert dr <= 1 and dl <= 1, (dr, dl)
    if dr == 0 and dl == 0:
        err["adjoint_source"] = rel(e.read("adjoint_sources")[0], o["adj"])
        err["dk_vs"] = rel(g["dk_vs"], dk_ref)
        assert g["res_r"][0] == pytest.approx(o["res_r"], rel=1e-9)
    else:
        # a pick one bin away changes the gradient by construction: the chain
        # after the picks is then judged from the oracle's picks
        e.run("forward")
        e.run("dispersion")
        e.set_picks(o["cr_r"][:, None], o["cr_l"][:, None])
        e.run("adjoint_source")
        e.run("adjoint")
        e.run("assemble")
        err["adjoint_source_given_oracle_picks"] = rel(e.read("adjoint_sources")[0], o["adj"])
        err["dk_vs_given_oracle_picks"] = rel(e.read("gradient")[0], dk_ref)
    print("engine fp32 rel-L2 vs fp64 oracle:", err, "timings:", e.timings())
    assert max(err.values()) < 1e-4, err
    e.close()


def test_engine_observe_vs_oracle(single):
    """Observed picks of SWD.m:150-162 computed on the device (swd_engine_observe)
    from the true model, against the FP64 oracle's, within one velocity bin."""
    prob, mod, cr0_r, cr0_l, _, _ = single
    e = Engine(prob)
    e.set_wavelet(mod["s"])
    e.set_model(mod["vp_true"], mod["vs_true"])
    a, b = e.observe()
    e.close()
    d = np.concatenate([np.abs(a[:, 0] - cr0_r[:, 0]), np.abs(b[:, 0] - cr0_l[:, 0])])
    print(f"observe(): {int(np.count_nonzero(d))} of {d.size} picks differ from FP64, max {d.max():g} m/s")
    assert d.max() <= 1.0


def test_engine_stages_given_oracle_picks(single):
    """Adjoint source, adjoint FD and dk_vs on the GPU from the oracle's picks:
    the FP32 chain downstream of the (integer) picks against FP64."""
    prob, mod, cr0_r, cr0_l, dk_ref, shots = single
    o = shots[0]
    e = _engine(prob, mod, cr0_r, cr0_l)
    e.run("forward")
    e.run("dispersion")
    e.set_picks(o["cr_r"][:, None], o["cr_l"][:, None])
    e.run("adjoint_source")
    e.run("adjoint")
    e.run("assemble")
    err = {"adjoint_source": rel(e.read("adjoint_sources")[0], o["adj"]),
           "dk_vs": rel(e.read("gradient")[0], dk_ref)}
    print("engine stages given oracle picks, fp32 rel-L2 vs fp64:", err)
    assert max(err.values()) < 1e-4, err
    e.close()


def test_weakly_lit_cells_match_fp64():
    """A short record (nt 1000) leaves cells the wavefield barely reaches.
    Where the strains are FP32 numbers (illum above 1e-60 in FP64) the image
    sums keep even their squares (sub-FP32 sums spill into FP64), so dk_vs =
    f(cl_img ./ illum, ...) is finite and within 1e-4 of the oracle there.
    Cells whose FP64 wavefield is below FP32's range (as in the reference's
    _single MEX) read zero strain: 0/0 there, and nowhere else."""
    prob, mod = P.plain_single_shot(nt=1000)
    cr0_r, cr0_l = oracle_observed(prob, mod, [5])
    dk_ref, shots, tot = oracle_gradient(prob, mod, [5], cr0_r, cr0_l)
    o = shots[0]
    e = _engine(prob, mod, cr0_r, cr0_l)
    e.run("forward")
    e.run("dispersion")
    e.set_picks(o["cr_r"][:, None], o["cr_l"][:, None])
    e.run("adjoint_source")
    e.run("adjoint")
    e.run("assemble")
    dk = e.read("gradient")[0]
    e.close()
    il = tot[2]
    lit = il > 1e-60
    nonfin = ~np.isfinite(dk)
    sub = (il > 0) & (il < 1e-38)
    err = float(np.linalg.norm(dk[lit] - dk_ref[lit]) / np.linalg.norm(dk_ref[lit]))
    print(f"cells: {il.size}, FP64 illum below FP32 range {int(sub.sum())}, above 1e-60 {int(lit.sum())}, "
          f"non-finite on the GPU {int(nonfin.sum())} (largest FP64 illum among them "
          f"{(il[nonfin].max() if nonfin.any() else 0):.1e}); dk_vs rel-L2 over cells above 1e-60 {err:.2e}")
    assert not nonfin[lit].any()
    assert err < 1e-4


def test_engine_three_shots_vs_oracle():
    nz, nx = 30, 120
    vs = np.tile(np.linspace(220, 420, nz)[:, None], (1, nx))
    vs_true = vs * (1 + 0.08 * P.checkerboard(nz, nx, 15, 1.0, 1.0) / 2)
    dt, nt = 1e-4, 1500
    prob = dict(nz=nz, nx=nx, nbc=10, fd_order=22, nt=nt, in_wf=5, dx=0.5, dt=dt, sx=np.array([1, 41, 81]),
                sz=np.array([1, 1, 1]), ng=120, gx0=1, dg=1, gz=1, df=0.5, fmin=5.0, fmax=80.0, vmin=100.0,
                vmax=800.0, ds=0.5, offset=120, m=0, M=40, ini=10, w=3.0, beta=10.0, ns=0)
    mod = {"vs": vs, "vp": 1.73 * vs, "vs_true": vs_true, "vp_true": 1.73 * vs_true, "s": O.ricker(30, dt, nt)}
    cr0_r, cr0_l = oracle_observed(prob, mod, [1, 2, 3])
    dk_ref, shots, tot = oracle_gradient(prob, mod, [1, 2, 3], cr0_r, cr0_l)
    # the device's observed picks of the same shots (SWD.m:150-162)
    ob = Engine(prob)
    ob.set_wavelet(mod["s"])
    ob.set_model(mod["vp_true"], mod["vs_true"])
    gr0, gl0 = ob.observe()
    ob.close()
    assert np.max(np.abs(gr0 - cr0_r)) <= 1
    assert np.max(np.abs(gl0[:, 1:] - cr0_l[:, 1:])) <= 1     # shot 1 has no left window (M = 40)
    e = _engine(prob, mod, cr0_r, cr0_l)
    g = e.gradient()
    same = True
    for k, o in enumerate(shots):
        assert np.max(np.abs(g["cr_r"][:, k] - o["cr_r"])) <= 1
        same &= np.array_equal(g["cr_r"][:, k], o["cr_r"])
        if k > 0:
            assert np.max(np.abs(g["cr_l"][:, k] - o["cr_l"])) <= 1
        if "cr_l" in o:   # shot 1's one-trace left window: a flat spectrum, picks not compared
            same &= np.array_equal(g["cr_l"][:, k], o["cr_l"])
    gath = e.read("gathers")
    for k, o in enumerate(shots):
        assert rel(gath[k], o["seismo_w"]) < 1e-4
    if same:
        err = rel(g["dk_vs"], dk_ref)
    else:
        e.run("forward")
        e.run("dispersion")
        e.set_picks(np.array([o["cr_r"] for o in shots]).T,
                    np.array([o.get("cr_l", np.ones_like(o["cr_r"])) for o in shots]).T)
        e.run("adjoint_source")
        e.run("adjoint")
        e.run("assemble")
        err = rel(e.read("gradient")[0], dk_ref)
    print("3-shot dk_vs rel-L2:", err, "(from the oracle's picks)" if not same else "")
    assert err < 1e-4
    e.close()


def test_batch_is_the_same_alone_and_in_mix(single):
    prob, mod, cr0_r, cr0_l, _, _ = single
    a = _engine(prob, mod, cr0_r, cr0_l)
    ga = a.gradient()["dk_vs"]
    # a second, differently sized model in the same launches
    nz, nx = 24, 90
    vs = np.tile(np.linspace(250, 400, nz)[:, None], (1, nx))
    prob2 = dict(nz=nz, nx=nx, nbc=12, fd_order=22, nt=1000, in_wf=5, dx=0.5, dt=1e-4, sx=np.array([1, 46]),
                 sz=np.array([1, 1]), ng=90, gx0=1, dg=1, gz=1, df=0.5, fmin=5.0, fmax=80.0, vmin=100.0, vmax=800.0,
                 ds=0.5, offset=90, m=0, M=45, ini=10, w=3.0, beta=10.0, ns=0)
    b = Engine(prob2)
    b.set_model(1.73 * vs, vs)
    b.set_wavelet(O.ricker(30, 1e-4, 1000))
    b.observe()
    b.set_model(1.73 * vs * 1.02, vs * 1.02)
    gb = b.gradient()["dk_vs"]
    (ma, mb), us = batch_gradient([a, b])
    np.testing.assert_array_equal(ma, ga)
    np.testing.assert_array_equal(mb, gb)
    a.close()
    b.close()


def test_plain_engine_path_matches_fused(single):
    prob, mod, cr0_r, cr0_l, _, _ = single
    e = _engine(prob, mod, cr0_r, cr0_l)
    gf = e.gradient()
    e.set_fd_path(1)
    gp = e.gradient()
    assert np.array_equal(gf["cr_r"], gp["cr_r"])
    assert rel(gp["dk_vs"], gf["dk_vs"]) < 1e-5
    e.close()


def test_uneven_batch_autotuned_is_the_same_alone_and_in_mix():
    """Three faulted layered models of different grid and nt in shared launches
    with the autotuned launch shape: each gradient is bitwise the gradient of
    the model alone (with the same settings), and the FD gathers of one shot
    per model agree with the FP64 oracle."""
    from swd_amd import problems as Pm
    from swd_amd.engine import batch_autotune
    models = Pm.uneven_models(seed=7, shots=3, sizes=(60, 90, 130), nt_range=(600, 900), f0=40.0)
    engs = []
    for prob, mod in models:
        e = Engine(prob)
        e.set_wavelet(mod["s"])
        e.set_model(mod["vp_true"], mod["vs_true"])
        e.observe()
        e.set_model(mod["vp"], mod["vs"])
        engs.append(e)
    tune = batch_autotune(engs, steps=60)
    # a batch tunes the time block only: one single-step candidate and time
    # blocks 2..6 at band heights chosen from occupancy (rows 0)
    tab = tune["table"]
    assert tune["candidates"] == tab.shape[0] == 6
    assert list(tab[:, 0]) == [1, 2, 3, 4, 5, 6] and np.all(tab[1:, 1] == 0)
    for ph in ("fwd", "adj"):
        assert tune[ph][1] == 0 or tune[ph][0] == 1
    outs, us = batch_gradient(engs)
    assert us[0] > 0 and us[1] > 0
    for (prob, mod), e, m in zip(models, engs, outs):
        g = e.gradient()["dk_vs"]
        np.testing.assert_array_equal(m, g)
        gx = prob["gx0"] + prob["dg"] * np.arange(prob["ng"])
        _, sw, _ = O.staggerfd_eigend(1, prob["nbc"], prob["nt"], prob["dt"] / prob["dx"], prob["dx"], prob["dt"],
                                      int(prob["sx"][0]), 1, gx, np.ones(prob["ng"]), mod["s"], mod["vp"], mod["vs"],
                                      1, 22, "w", 0, prob["in_wf"], snapshots=False)
        err = rel(e.read("gathers")[0], sw)
        print(prob["nz"], prob["nx"], prob["nt"], "gather rel-L2 vs FP64:", err)
        assert err < 1e-4
    for e in engs:
        e.close()


def test_overlapped_batch_equals_phase_by_phase_batch_and_alone():
    """The overlapped batch schedule (forward and adjoint launch sequences on
    two streams, each model's adjoint joining as soon as its forward and its
    dispersion stage have run) against the phase-by-phase batch (single-step
    launches, all forwards then all adjoints) and against each model alone:
    bitwise the same gradients, on three models of different grid and nt
    whose forwards end at different launches."""
    from swd_amd.engine import batch_gradient
    models = P.uneven_models(shots=2, sizes=(200, 300, 450), nt_range=(1500, 2500))
    engs = []
    for prob, mod in models:
        e = Engine(prob)
        e.set_wavelet(mod["s"])
        e.set_model(mod["vp_true"], mod["vs_true"])
        e.observe()
        e.set_model(mod["vp"], mod["vs"])
        engs.append(e)
    assert len({p["nt"] for p, _ in models}) == 3
    for e in engs:
        e.set_time_block(5, 0)
    ovl, us = batch_gradient(engs)
    assert us[0] > 0 and us[1] > 0
    alone = [e.gradient()["dk_vs"] for e in engs]
    for e in engs:
        e.set_time_block(1, 0)
    seq, _ = batch_gradient(engs)
    for a, b, c in zip(ovl, seq, alone):
        assert np.isfinite(a).mean() > 0.5
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(a, c)
    for e in engs:
        e.close()
