"""CPU checks of the benchmark's contract and byte models: `--gpus N > 1`
exits with a message (one GPU only), and the compulsory bytes per
time-blocked launch that `roofline.achieved` / `roofline.frac` divide by are
the figures the kernels' data layout fixes at the production shape."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from swd_amd import problems as P  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def _production():
    prob, _ = P.production_shots()
    return prob


def test_compulsory_bytes_at_the_production_shape():
    """60 shots, 132 x 660 padded (100 x 600 interior), 5 steps per launch
    with one imaging step, 600 receivers.  Per shot: 5 state fields of 4 B
    read once and written once (87120 cells x 40 B = 3484800 B), the 3-field
    snapshot of the imaging step (60000 x 12 B = 720000 B), 600 receivers x 5
    steps x 4 B (12000 B); per model the 5 coefficient planes once (87120 x
    20 B).  The adjoint adds the FP32 partial-image read + write (60000 x
    24 B) and 1/32 of the FP64 flush (60000 x 48 B / 32)."""
    prob = _production()
    state, snap, rec, coef = 87120 * 40, 60000 * 12, 600 * 5 * 4, 87120 * 20
    fwd = 60 * (state + snap + rec) + coef
    adj = fwd + 60 * (60000 * 24 + 60000 * 48 // 32)
    assert fwd == 254_750_400 and adj == 346_550_400
    assert bench.fd_compulsory_bytes_per_launch(prob, 60, False, 5) == pytest.approx(fwd, rel=0, abs=0.5)
    assert bench.fd_compulsory_bytes_per_launch(prob, 60, True, 5) == pytest.approx(adj, rel=0, abs=0.5)


def test_nominal_model_exceeds_the_compulsory_bytes():
    """The unblocked per-step model (k x the per-step bytes) moves the state
    once per step, so at k = 5 it is several times the compulsory bytes of a
    time-blocked launch: the reason it is not the roofline's divisor."""
    prob = _production()
    for adj in (False, True):
        nominal = 5 * bench.fd_bytes_per_launch(prob, 60, adj)
        compulsory = bench.fd_compulsory_bytes_per_launch(prob, 60, adj, 5)
        assert 4.0 < nominal / compulsory < 7.0
