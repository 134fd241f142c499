"""The time-blocked FD kernels (NT steps per launch) are bitwise equal to the
single-step kernels, for every supported NT and band height, in the MEX
entry points and in the batched engine."""
import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import fd, problems as P
from swd_amd.engine import Engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    nz, nx, nbc, dx, dt, nt = 40, 150, 15, 0.5, 1e-4, 600
    r = np.random.default_rng(11)
    vs = np.tile(np.linspace(200, 450, nz)[:, None], (1, nx)) * r.uniform(0.95, 1.05, (nz, nx))
    vp = 1.73 * vs
    s = O.ricker(40, dt, nt)
    gx = np.arange(1, nx, dtype=float)
    args = (1, nbc, nt, dt / dx, dx, dt, 17, 1, gx, np.ones(nx - 1), s, vp, vs, 1, 22, "w", 0, 5)
    return args


def _fwd(args, nt_block, rows):
    fd.set_time_block(nt_block, rows)
    try:
        return fd.staggerfd_eigen(*args)
    finally:
        fd.set_time_block(4, 0)


@pytest.mark.parametrize("nt_block,rows", [(2, 32), (3, 9), (4, 32), (4, 7), (4, 0), (5, 26), (6, 40)])
def test_mex_forward_and_adjoint_bitwise(case, nt_block, rows):
    su1, sw1, S1 = _fwd(case, 1, 32)
    su, sw, S = _fwd(case, nt_block, rows)
    np.testing.assert_array_equal(sw, sw1)
    np.testing.assert_array_equal(su, su1)
    for k in S1:
        np.testing.assert_array_equal(S[k], S1[k])
    a = case
    adj = np.random.default_rng(2).standard_normal((a[2], a[8].size)).astype(np.float32)
    b = (S1, adj) + a[:1] + (a[1], a[2], a[3], a[4], a[5], a[8], a[9], a[10], a[11], a[12], 1, 22, 0, 5)
    fd.set_time_block(1, 32)
    g1 = fd.e2drtm_eigen(*b)
    fd.set_time_block(nt_block, rows)
    try:
        g = fd.e2drtm_eigen(*b)
    finally:
        fd.set_time_block(4, 0)
    for x, y in zip(g, g1):
        np.testing.assert_array_equal(x, y)


def test_engine_gradient_bitwise_across_time_blocks():
    prob, mod = P.plain_single_shot(nt=800)
    e = Engine(prob)
    e.set_wavelet(mod["s"])
    e.set_model(mod["vp_true"], mod["vs_true"])
    e.observe()
    e.set_model(mod["vp"], mod["vs"])
    e.set_time_block(1, 32)
    g1 = e.gradient()
    gath1 = e.read("gathers")
    for nt_block, rows in ((4, 32), (5, 16), (3, 50), (6, 0)):
        e.set_time_block(nt_block, rows)
        g = e.gradient()
        t = e.timings()
        assert t["fwd_steps_per_launch"] == nt_block and t["adj_steps_per_launch"] == nt_block
        assert t["fwd_band_rows"] >= 8 if rows == 0 else t["fwd_band_rows"] == rows
        np.testing.assert_array_equal(e.read("gathers"), gath1)
        np.testing.assert_array_equal(g["cr_r"], g1["cr_r"])
        np.testing.assert_array_equal(g["dk_vs"], g1["dk_vs"])
    # separate forward / adjoint settings
    e.set_time_block(2, 20, phase=0)
    e.set_time_block(5, 0, phase=1)
    g = e.gradient()
    t = e.timings()
    assert (t["fwd_steps_per_launch"], t["fwd_band_rows"], t["adj_steps_per_launch"]) == (2, 20, 5)
    np.testing.assert_array_equal(g["dk_vs"], g1["dk_vs"])
    # the autotuner installs the fastest settings it timed; results unchanged
    tune = e.autotune(steps=60)
    tab = tune["table"]
    assert tune["candidates"] == tab.shape[0] > 10
    assert np.all(tab[:, 2] > 0) and np.all(tab[:, 3] > 0)
    i_f, i_a = int(np.argmin(tab[:, 2])), int(np.argmin(tab[:, 3]))
    assert tune["fwd"][0] == tab[i_f, 0] and tune["adj"][0] == tab[i_a, 0]
    g = e.gradient()
    t = e.timings()
    assert t["fwd_steps_per_launch"] == tune["fwd"][0] and t["adj_steps_per_launch"] == tune["adj"][0]
    # the snapshot buffer it kept is the fastest placement it timed; the
    # gradient on it is bitwise the same
    pl, kept = t["snapshot_placement_ms"], t["snapshot_placement_kept"]
    assert 1 <= len(pl) <= 3 and min(pl) > 0 and pl[kept] == min(pl)
    np.testing.assert_array_equal(e.read("gathers"), gath1)
    np.testing.assert_array_equal(g["cr_r"], g1["cr_r"])
    np.testing.assert_array_equal(g["dk_vs"], g1["dk_vs"])
    e.close()


def test_mex_separate_b1_plane(case):
    """The MEX entry points take b1 separately from b.  With b1 != b the
    time-blocked kernels read the b1 plane: time blocks stay bitwise equal to
    the single-step kernels, the result follows the FP64 oracle with the
    same planes, and it differs from the b1 = b result."""
    a = case
    is_, nbc, nt, dtx, dx, dt, sx, sz, gx, gz, s, vp, vs = a[:13]
    c = dict(fd.prepare_model(vp, vs, nbc, dx, dt, 22))
    c_b1 = dict(c)
    c_b1["b1"] = c["b1"] * np.linspace(0.95, 1.05, c["b1"].shape[1])[None, :]
    iv = fd._fwd_iv(nt, vp.shape, nbc, dtx, gx, gz, sx, sz, "w", 22, 5)

    def run(cc, nt_block):
        fd.set_time_block(nt_block, 0)
        try:
            return fd.eigen_staggerfd_single(iv, cc["temp"], cc["ca"], cc["cl"], cc["cm"], cc["cm1"], cc["b"],
                                             cc["b1"], s)
        finally:
            fd.set_time_block(4, 0)

    ref1 = run(c_b1, 1)
    for nt_block in (4, 5):
        out = run(c_b1, nt_block)
        np.testing.assert_array_equal(out[1], ref1[1])
        for k in ref1[2]:
            np.testing.assert_array_equal(out[2][k], ref1[2][k])
    same = run(c, 4)
    _, sw64, _ = O.staggerfd(iv, c_b1, s, "d", snapshots=False)
    e = float(np.linalg.norm(ref1[1] - sw64) / np.linalg.norm(sw64))
    d = float(np.linalg.norm(same[1] - ref1[1]) / np.linalg.norm(ref1[1]))
    print(f"separate b1: rel-L2 vs FP64 {e:.2e}; differs from b1 = b by {d:.2e}")
    assert e < 1e-4 and d > 1e-3
    # adjoint + imaging with the same planes: time blocks bitwise equal to single steps
    nz, nx = vp.shape
    iva = np.array([nt, nz + 2 + nbc, nx + 2 * nbc, dtx, gx.size, nbc, 1 + 2, gx[0] + nbc, 1.0, 22, 5, nz, nx, dt])
    adj = np.random.default_rng(4).standard_normal((nt, gx.size)).astype(np.float32)
    S = ref1[2]

    def run_adj(nt_block):
        fd.set_time_block(nt_block, 0)
        try:
            return fd.eigen_e2drtm_single(iva, c_b1["temp"], c_b1["ca"], c_b1["cl"], c_b1["cm"], c_b1["cm1"],
                                          c_b1["b"], c_b1["b1"], adj, S["fux"], S["fuz"], S["bwx"], S["bwz"])
        finally:
            fd.set_time_block(4, 0)

    g1 = run_adj(1)
    for nt_block in (4, 5):
        for x, y in zip(run_adj(nt_block), g1):
            np.testing.assert_array_equal(x, y)


def test_engine_gradient_bitwise_with_several_imaging_levels_per_launch():
    """Snapshots every 2 steps: a 5-step launch holds 2 or 3 snapshot and
    imaging levels, so the kernels take their general (inline) record and
    imaging paths instead of the one-level ones."""
    prob, mod = P.plain_single_shot(nt=800)
    prob = dict(prob, in_wf=2)
    e = Engine(prob)
    try:
        e.set_wavelet(mod["s"])
        e.set_model(mod["vp_true"], mod["vs_true"])
        e.observe()
        e.set_model(mod["vp"], mod["vs"])
        e.set_time_block(1, 32)
        g1 = e.gradient()
        gath1 = e.read("gathers")
        # cells the short record never reaches have zero illumination (0/0 there, as in the reference)
        lit = np.isfinite(g1["dk_vs"])
        assert lit.mean() > 0.3 and np.any(g1["dk_vs"][lit] != 0)
        for nt_block, rows in ((5, 16), (3, 0), (2, 24)):
            e.set_time_block(nt_block, rows)
            g = e.gradient()
            assert e.timings()["adj_steps_per_launch"] == nt_block
            np.testing.assert_array_equal(e.read("gathers"), gath1)
            np.testing.assert_array_equal(g["dk_vs"], g1["dk_vs"])
    finally:
        e.close()
