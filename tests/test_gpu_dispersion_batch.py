"""The batched dispersion path alone (swd_dispersion_*) on dispersive chirp
gathers of the Radon configuration's size (nt 6000, 600 traces, dt 8e-5,
0.5 m traces, 5-80 Hz at 0.5 Hz, 100-800 m/s): picks against the FP64
oracle and against the chirp's own phase velocity, and the adjoint sources
against the oracle's ADWDgrad_1 for beta 5, 10 and 20.  FP32 results are
judged by relative L2 <= 1e-4; picks within one velocity bin."""
import numpy as np
import pytest

from oracle import swd_oracle as O
from swd_amd import problems as P
from swd_amd.dispersion import DispersionBatch

pytestmark = pytest.mark.gpu

NSHOT = 2


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def chirp():
    prob, _ = P.production_shots(nshots=NSHOT)
    g = P.chirp_gathers(nshots=NSHOT, nt=prob["nt"], ng=prob["ng"], dt=prob["dt"], dx=prob["ds"], sx=prob["sx"])
    D = DispersionBatch(prob)
    D.set_gathers(g)
    D.set_observed_picks(np.ones((D.npair, NSHOT)), np.ones((D.npair, NSHOT)))
    picks = D.pick()
    q = prob
    radon = lambda s: (s + 1, q["df"], q["dt"], D.np, q["vmin"], q["vmax"], q["fmin"], q["fmax"], q["nt"], q["ng"],
                       q["ds"], q["offset"], q["m"], q["M"])
    ref = []
    for s in range(NSHOT):
        sides = {}
        for side in ("r", "l"):
            ml, _, sfb = O.rt_side(g[s].astype(np.float64), side, *radon(s))
            _, cr = O.LHDispPick(ml, q["vmin"], np.ones(D.npair), 1, q["ini"])
            sides[side] = (ml, cr, sfb)
        ref.append(sides)
    yield prob, g, D, picks, ref
    D.close()


def test_batch_picks_vs_oracle_and_chirp(chirp):
    prob, g, D, picks, ref = chirp
    f = prob["fmin"] + prob["df"] * np.arange(D.npair)
    c = P.chirp_phase_velocity(f)
    band = (f >= 15) & (f <= 70)
    for s in range(NSHOT):
        for side, key in (("r", "cr_r"), ("l", "cr_l")):
            ml, cr, sfb = ref[s][side]
            L = len(sfb["uxtpos"])
            gpu = picks[key][:, s]
            if L > 1:  # a one-trace window (the first shot's left side) has a flat spectrum
                assert np.max(np.abs(gpu - cr)) <= 1.0, (s, side)
            if L < 200:
                continue  # too short an aperture to resolve the chirp's velocity
            # cr_pre = index + vmin is one bin above the axis velocity
            err = (gpu - 1.0) - c
            frac = float(np.mean(np.abs(err[band]) <= 1.0))
            print(f"shot {s + 1} side {side} ({L} traces): picks within 1 bin of c(f) over 15-70 Hz: {frac:.3f}")
            assert frac >= 0.9


@pytest.mark.parametrize("beta", [5.0, 10.0, 20.0])
def test_batch_adjoint_vs_oracle(chirp, beta):
    prob, g, D, picks, ref = chirp
    r = np.random.default_rng(int(beta))
    gr = np.round(r.normal(0, 5, (D.npair, NSHOT)))
    gl = np.round(r.normal(0, 5, (D.npair, NSHOT)))
    # observed picks chosen so that cr_pre - cr_0 = gr / gl exactly
    D.set_observed_picks(picks["cr_r"] - gr, picks["cr_l"] - gl)
    adj, ms = D.adjoint(beta)
    q = prob
    for s in range(NSHOT):
        a_ref = O.ADWDgrad_1(q["nt"], q["ng"], NSHOT, D.npair, s + 1, q["w"], q["m"], q["M"], beta, gr[:, s],
                             gl[:, s], ref[s]["r"][2], ref[s]["l"][2])
        e = rel(adj[s], a_ref)
        print(f"beta {beta} shot {s + 1}: adjoint source rel-L2 vs FP64 {e:.3e} ({ms:.2f} ms for the batch)")
        assert e < 1e-4


def test_bf16x3_radon_matches_fp32_and_oracle(chirp):
    """The BF16 split-precision DFT / inverse-DFT GEMMs (swd_set_radon_precision(1)):
    picks within one bin of the FP32 path, adjoint sources within 1e-4 of the oracle."""
    from swd_amd.dispersion import set_radon_precision
    prob, g, D, picks, ref = chirp
    old = set_radon_precision(1)
    try:
        D.set_observed_picks(np.ones((D.npair, NSHOT)), np.ones((D.npair, NSHOT)))
        p16 = D.pick()
        for key, side in (("cr_r", "r"), ("cr_l", "l")):
            for s in range(NSHOT):
                if len(ref[s][side][2]["uxtpos"]) < 200:
                    continue  # short windows have near-flat spectra: picks there move with any rounding
                assert np.max(np.abs(p16[key][:, s] - picks[key][:, s])) <= 1.0
        r = np.random.default_rng(3)
        gr = np.round(r.normal(0, 5, (D.npair, NSHOT)))
        gl = np.round(r.normal(0, 5, (D.npair, NSHOT)))
        D.set_observed_picks(p16["cr_r"] - gr, p16["cr_l"] - gl)
        adj, ms = D.adjoint(10.0)
    finally:
        set_radon_precision(old)
    q = prob
    for s in range(NSHOT):
        # the oracle with the picks the bf16 path made (identical g by construction)
        a_ref = O.ADWDgrad_1(q["nt"], q["ng"], NSHOT, D.npair, s + 1, q["w"], q["m"], q["M"], 10.0, gr[:, s],
                             gl[:, s], ref[s]["r"][2], ref[s]["l"][2])
        e = rel(adj[s], a_ref)
        print(f"bf16x3 shot {s + 1}: adjoint source rel-L2 vs FP64 {e:.3e}")
        assert e < 1e-4
