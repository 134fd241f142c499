"""SWD hot-path benchmark on one MI355X.

Workload (one "step"): one SWD gradient of the single-model production-shots
configuration - 100 x 600 cells at 0.5 m with 30-cell borders (132 x 660
padded), 60 Ricker 25 Hz shots 5 m apart, 600 receivers, dt = 8e-5 s,
nt = 6000, snapshots every 5 steps - all 60 shots in one batch:
forward FD + snapshots, RTrAD/RTlAD (151 frequencies x 701 velocities),
LHDispPick, ADWDgrad_1 (beta = 10), adjoint FD + imaging, shot reduction and
dk_vs.  Synthetic inputs: checkerboard "true" model for the observed picks
(computed before timing), 300 m/s homogeneous current model.

Prints one JSON line (the last line of stdout).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def fd_bytes_per_launch(prob, nshots, adjoint):
    """Nominal HBM bytes of one FD time step over the batch (FP32), the
    per-step model of an unblocked stencil: 60 B per padded cell-step (5 state
    fields read + written, 5 coefficients read) plus, averaged over the
    in_wf-step snapshot cycle, 16 B per interior cell per snapshot step
    (forward: strain written; adjoint: strain read) and, for the adjoint, the
    image update: an FP32 read-modify-write of the 3 partial images (24 B)
    plus, every 32nd imaging step, the FP64 read-modify-write of the 3 images
    (48 B / 32).  A time-blocked launch of k steps never moves the
    intermediate steps' bytes, so k x this exceeds what it moves."""
    nzbc = prob["nz"] + 2 + prob["nbc"]
    nxbc = prob["nx"] + 2 * prob["nbc"]
    cells = nzbc * nxbc
    inner = prob["nz"] * prob["nx"]
    per = 60.0 * cells + 16.0 * inner / prob["in_wf"]
    if adjoint:
        per += (24.0 + 48.0 / 32) * inner / prob["in_wf"]
    return per * nshots


def fd_compulsory_bytes_per_launch(prob, nshots, adjoint, k):
    """Bytes a time-blocked launch of k steps must move to or from HBM, as the
    kernels lay the data out: each shot's state (5 fields) read once and
    written once (40 B per padded cell); the coefficient planes (temp, b, ca,
    cl, cm1; b1 = b) read once per model (20 B per padded cell); per imaging
    step (k / in_wf per launch on average) and interior cell the 3-field
    snapshot (12 B: fux, bwz, fuz + bwx; written forward, read adjoint) and,
    adjoint only, the FP32 partial-image read-modify-write (24 B) and the
    FP64 flush every 32nd imaging step (3 x 8 B read + written, 48 B / 32);
    receivers or adjoint-source samples (4 B per receiver and step)."""
    nzbc = prob["nz"] + 2 + prob["nbc"]
    nxbc = prob["nx"] + 2 * prob["nbc"]
    cells = nzbc * nxbc
    inner = prob["nz"] * prob["nx"]
    imaging = k / prob["in_wf"]
    per_shot = 40.0 * cells + imaging * 12.0 * inner + 4.0 * k * prob["ng"]
    if adjoint:
        per_shot += imaging * (24.0 + 48.0 / 32) * inner
    return per_shot * nshots + 20.0 * cells


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(prob, mod, max_threads=16):
    """The FP32 statement-structure C port of the MEX time loops (oracle/):
    forward + adjoint of production-grid shots at the full nt, one shot per
    thread (a parpool worker of the reference is single-threaded), on the
    host cores this job may use, at most `max_threads`: a one-GPU job gets a
    16-core share of the host whatever nproc reports.  Also times one shot
    alone on one core."""
    from oracle import swd_oracle as O
    L = O.lib()
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    threads = max(1, min(max_threads, avail))
    nz, nx, nbc = prob["nz"], prob["nx"], prob["nbc"]
    nzbc, nxbc = nz + 2 + nbc, nx + 2 * nbc
    nt = prob["nt"]
    c = O.prep(mod["vp"], mod["vs"], nbc, prob["dx"], prob["dt"], 22)
    f32 = {k: np.asfortranarray(v, dtype=np.float32) for k, v in c.items()}
    dtx = prob["dt"] / prob["dx"]
    ivf = np.array([nt, nzbc, nxbc, dtx, prob["ng"], 3, nbc + 1, 3, nbc + 1, 1, 5, 22, prob["in_wf"], nz, nx, 3],
                   dtype=np.float32)
    ivb = np.array([nt, nzbc, nxbc, dtx, prob["ng"], nbc, 3, nbc + 1, 1, 22, prob["in_wf"], nz, nx, prob["dt"]],
                   dtype=np.float32)
    s = np.ascontiguousarray(mod["s"][:nt], dtype=np.float32)
    p = O._p

    def run(nsh, nthreads):
        sx = np.ascontiguousarray((prob["sx"][:nsh] + nbc) % nxbc, dtype=np.int32)
        rng = np.random.default_rng(0)
        adj = np.asfortranarray(rng.standard_normal((nsh * nt, prob["ng"])).astype(np.float32) * 1e-3)
        sw = np.zeros(nsh * nt * prob["ng"], dtype=np.float32)
        g = [np.zeros(nz * nx, dtype=np.float32) for _ in range(3)]
        t0 = time.perf_counter()
        failed = L.oracle_batch_f(p(ivf), p(ivb), *[p(f32[k]) for k in O.COEF_ORDER], p(s), p(sx), nsh, p(adj),
                                  p(sw), p(g[0]), p(g[1]), p(g[2]), nthreads)
        wall = time.perf_counter() - t0
        if failed:
            raise RuntimeError("CPU baseline shots failed")
        return wall

    per_shot = nzbc * nxbc * (nt + nt - 1)
    wall = run(threads, threads)
    one = run(1, 1)
    return {"value": round(threads * per_shot / wall / 1e9, 3), "unit": "Gcell-steps/s", "cores": threads,
            "kind": "port",
            "one_core_shot_s": round(one, 2), "one_core_gcell_steps_per_s": round(per_shot / one / 1e9, 3),
            "nproc": os.cpu_count(), "affinity_cpus": avail, "cpu_model": _cpu_model(),
            "sample": f"{threads} shots (one per thread, {threads} threads) of the 132x660 grid at the full nt={nt}, "
                      f"forward + adjoint with imaging, FP32 C restatement of the MEX statement loops (one array pass "
                      f"per Eigen statement), {wall:.1f} s; one shot alone on one core {one:.1f} s. FD only: the "
                      f"Radon / pick / ADWDgrad_1 stages are not part of this CPU sample"}


def copy_peak(nbytes=1 << 30, reps=10):
    """HBM rates (read + write bytes per second) of two device copies of
    1 GiB, measured now: the best streaming copy (swd_copy_peak, the practical
    ceiling beside the 8 TB/s figure) and the plain float4 copy whose counter
    readings calibrate the traffic figures (swd_calib_copy)."""
    import ctypes
    from swd_amd._lib import check, load
    out = []
    for fn in (load().swd_copy_peak, load().swd_calib_copy):
        ms = ctypes.c_double(0)
        check(fn(nbytes, 2, ctypes.byref(ms)))
        check(fn(nbytes, reps, ctypes.byref(ms)))
        out.append(2.0 * nbytes / (ms.value * 1e-3) / 1e9)
    return out


N_SIMD = 1024           # MI355X: 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9        # MI355X peak engine clock


def measured_profile(kname, shape, args):
    """Per-launch HBM traffic (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a
    copy kernel of known size) and VALU instruction count (SQ_INSTS_VALU) of
    the kernel over the same workload, from swd_amd/fd_traffic.json (written
    by ``python -m swd_amd.profile_fd summarize``), when the profiled launch
    shape is the one that ran; else {}."""
    path = os.path.join(ROOT, "swd_amd", "fd_traffic.json")
    if not os.path.exists(path) or args.shots != 60 or args.nt != 6000:
        return {}
    with open(path) as fh:
        t = json.load(fh)
    k = kname.replace("_march", "_tb") if shape[0] > 1 else kname
    rec = t.get("by_shape", {}).get(f"{k}:{shape[0]}x{shape[1]}")
    if rec is None:
        return {}
    return {"bytes_per_launch": int(rec["bytes_per_launch"]), "valu_insts": rec.get("valu_insts")}


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def uneven_leg(steps=2, graph=True):
    """The uneven multi-model batch: 8 faulted layered models of different grid
    size and nt (16 shots each) in shared FD launches, steps per launch chosen
    by the autotuner on the batch (band heights from occupancy per launch
    table).  Also times the same 8 gradients one model after another, each at
    the launch shape the autotuner picks for it alone, and checks each
    model's gradient alone == in the mix."""
    from swd_amd import problems as P
    from swd_amd.engine import Engine, batch_autotune, batch_gradient
    models = P.uneven_models()
    engs = []
    for prob, mod in models:
        e = Engine(prob)
        e.set_wavelet(mod["s"])
        e.set_model(mod["vp_true"], mod["vs_true"])
        e.observe()
        e.set_model(mod["vp"], mod["vs"])
        engs.append(e)
    progress("uneven batch: engines ready")
    t0 = time.perf_counter()
    tune = batch_autotune(engs)
    tune_s = time.perf_counter() - t0
    for e in engs:                         # the batch and the one-model-at-a-time runs replay graphs
        e.set_graph(graph)
    batch_gradient(engs)
    t0 = time.perf_counter()
    for _ in range(steps):
        outs, _ = batch_gradient(engs)
    t_batch = (time.perf_counter() - t0) / steps
    cell_steps = 0
    for e in engs:
        i = e.info()
        cell_steps += e.nshots * i["nzbc"] * i["nxbc"] * (2 * e.problem["nt"] - 1)
    # one model at a time, each with the launch shape the autotuner picks for
    # it alone (the batch's shape can be far from it: 89-row adjoint bands
    # made the same gradients 2.2x slower), graphs captured before timing
    seq_shapes = []
    for e in engs:
        t = e.autotune()
        seq_shapes.append([list(t["fwd"]), list(t["adj"])])
        e.gradient()
    t0 = time.perf_counter()
    alone = [e.gradient()["dk_vs"] for e in engs]
    t_seq = time.perf_counter() - t0
    same = all(np.array_equal(a, b, equal_nan=True) for a, b in zip(alone, outs))
    for e in engs:
        e.close()
    return {"value": round(cell_steps / t_batch / 1e9, 3), "unit": "Gcell-steps/s",
            "ms_per_batch_gradient": round(1000 * t_batch, 2),
            "sequential_ms": round(1000 * t_seq, 2), "batch_speedup_vs_sequential": round(t_seq / t_batch, 3),
            "sequential_fd_launch": seq_shapes,
            "models": [[p["nz"], p["nx"], p["nt"], len(p["sx"])] for p, _ in models],
            "fd_launch": {"fwd": list(tune["fwd"]), "adj": list(tune["adj"]), "candidates": tune["candidates"],
                          "tune_s": round(tune_s, 2)},
            "alone_equals_in_mix": bool(same),
            "what": "8 models (nz x nx x nt x shots as listed), gradient of all 128 shots per step; FP32 error "
                    "against the FP64 oracle is a test (tests/test_gpu_configs.py::test_uneven_batch_full_size), "
                    "not part of the timed run"}


def iteration_leg(eng, mod, prob):
    """One full SWD iteration of the production-shots workload on the device
    (swd_engine_iterate: gradient, smooth2a, line search of forward-only
    batches, model update), after the timed gradients; the engine's model
    changes, so it runs last on this engine."""
    vsmin, vsmax = float(np.min(mod["vs_true"])), float(np.max(mod["vs_true"]))
    t0 = time.perf_counter()
    r = eng.iterate(vsmin, vsmax, pur=-0.05, smooth_z=1, smooth_x=2, state={"f1": 0.5})
    ms = 1000.0 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    eng.forward_picks()
    fwd_ms = 1000.0 * (time.perf_counter() - t0)
    return {"ms": round(ms, 1), "forward_batches": r["forwards"], "forward_batch_ms": round(fwd_ms, 1),
            "residual": r["residual"], "res0": r["res0"], "gama": r["gama"], "alpha": r["alpha"],
            "what": "SWD.m:171-335 for one k on 60 shots: gradient + smooth2a + line search (forward FD without "
                    "snapshots, Radon, picks per trial) + model update, device-resident"}


def field_leg(graph=True):
    """The outcrop survey at full size (fieldexamples/outcrop/outcrop.m: 40 x
    240 cells at 0.25 m, 40-cell borders, 60 shots, 60 receivers, nt 7450,
    npair 651 at df 0.1 Hz, 50-trace windows) with the outcrop variant
    (norm_trace, strongest-column pick start, windowed misfit, peak-normalised
    adjoint sources, vp = 1.732 vs): one gradient and one device iteration.
    The observed picks are synthetic (a forward of a model with a +8% Vs
    anomaly), standing in for the survey's manual picks."""
    from swd_amd import field as F
    from swd_amd.engine import Engine
    prob, mod = F.outcrop(offset=50)
    vs_true = mod["vs"] * (1 + 0.08 * np.exp(-((np.arange(240)[None, :] - 120) ** 2 +
                                                (np.arange(40)[:, None] - 15.0) ** 2) / (2 * 12.0 ** 2)))
    e = Engine(prob)
    try:
        e.set_variant("outcrop")
        e.set_wavelet(mod["s"])
        e.set_model(1.732 * vs_true, vs_true)
        e.observe()
        e.set_model(mod["vp"], mod["vs"])
        tune = e.autotune()
        e.set_graph(graph)
        e.gradient()
        t0 = time.perf_counter()
        e.gradient()
        g_s = time.perf_counter() - t0
        t = e.timings()
        i = e.info()
        cell_steps = e.nshots * i["nzbc"] * i["nxbc"] * (2 * prob["nt"] - 1)
        vsmin, vsmax = F.OUTCROP_VS_BOUNDS
        t0 = time.perf_counter()
        r = e.iterate(vsmin, vsmax, pur=-0.05, smooth_z=1, smooth_x=2, state={"f1": 0.5})
        it_ms = 1000.0 * (time.perf_counter() - t0)
    finally:
        e.close()
    return {"value": round(cell_steps / g_s / 1e9, 3), "unit": "Gcell-steps/s", "ms_per_gradient": round(1000 * g_s, 2),
            "stage_ms": {k: round(t[k], 2) for k in ("fwd_ms", "disp_ms", "adj_ms", "asm_ms")},
            "fd_launch": {"fwd": list(tune["fwd"]), "adj": list(tune["adj"])},
            "npair": i["npair"], "grid": [i["nzbc"], i["nxbc"]], "nt": prob["nt"], "shots": e.nshots,
            "iteration_ms": round(it_ms, 1), "iteration_forward_batches": r["forwards"],
            "residual": r["residual"], "res0": r["res0"],
            "what": "outcrop.m survey, one SWD gradient and one device iteration with the field variant"}


PEAK_MFMA_FP32_TFLOPS = 157.3  # MI355X dense FP32 MFMA


def dispersion_leg(betas=(5.0, 10.0, 20.0), reps=3):
    """The dispersion path alone: 60 chirp gathers of 6000 x 600 (dt 8e-5,
    0.5 m traces, phase velocity 250 + 300 exp(-f/20) m/s, 10% band-limited
    noise), Radon over 151 frequencies x 701 velocities for every (shot,
    side) window, picks, then adjoint sources for each beta.  FLOPs: the DFT
    at the bins and the inverse DFT are real GEMMs of 2 x (2 npair) x nt x
    ng per shot; each slant stack is 8 x np x npair x (total window length)."""
    from swd_amd import problems as P
    from swd_amd.dispersion import DispersionBatch
    prob, _ = P.production_shots(nshots=60)
    g = P.chirp_gathers(nshots=60, nt=prob["nt"], ng=prob["ng"], dt=prob["dt"], dx=prob["ds"], sx=prob["sx"])
    D = DispersionBatch(prob)
    D.set_gathers(g)
    D.set_observed_picks(np.ones((D.npair, 60)), np.ones((D.npair, 60)))
    D.pick()
    pick_ms = min(D.pick()["ms"] for _ in range(reps))
    r = D.pick()
    adj_ms = {}
    for b in betas:
        D.adjoint(b, want=False)
        adj_ms[f"{b:g}"] = round(min(D.adjoint(b, want=False)[1] for _ in range(reps)), 3)
    # the BF16 split-precision variant of the DFT / inverse-DFT GEMMs
    from swd_amd.dispersion import set_radon_precision
    old = set_radon_precision(1)
    try:
        D.pick()
        pick16 = min(D.pick()["ms"] for _ in range(reps))
        r16 = D.pick()
        D.adjoint(10.0, want=False)
        adj16 = min(D.adjoint(10.0, want=False)[1] for _ in range(reps))
    finally:
        set_radon_precision(old)
    # the field-script variant with the QadimahFault FK == 1 branch (fk_filter per window) and norm_trace
    D.set_variant({"fk_filter": 1, "norm_trace": 1})
    D.pick()
    pick_fk = min(D.pick()["ms"] for _ in range(reps))
    D.adjoint(10.0, want=False)
    adj_fk = min(D.adjoint(10.0, want=False)[1] for _ in range(reps))
    D.close()
    same_bin = float(np.mean(np.abs(r16["cr_r"] - r["cr_r"]) <= 1.0))
    nt, ng, npair, S = prob["nt"], prob["ng"], D.npair, 60
    dft = 2.0 * (2 * npair) * min(nt, D.ccn) * ng * S
    slant = 8.0 * D.np * npair * D.sum_l
    f = prob["fmin"] + prob["df"] * np.arange(npair)
    c = P.chirp_phase_velocity(f)
    band = (f >= 15) & (f <= 70)
    # picks are cr_pre = index + vmin, one bin above the axis velocity; judged on windows of >= 200 traces
    # (100 m), where the aperture resolves the chirp (offset 600 >= ng: right window ng - sx + 1, left sx)
    sx = np.asarray(prob["sx"])
    ok = [np.abs((r["cr_r"][:, s] - 1.0) - c)[band] <= 1.0 for s in range(S) if ng - sx[s] + 1 >= 200] + \
         [np.abs((r["cr_l"][:, s] - 1.0) - c)[band] <= 1.0 for s in range(S) if sx[s] >= 200]
    amean = float(np.mean(list(adj_ms.values())))
    return {"pick_ms": round(pick_ms, 3), "adjoint_ms": adj_ms, "shots": S,
            "shots_per_s": round(S / ((pick_ms + amean) * 1e-3), 1),
            "mfma": {"forward_gflop": round((dft + slant) / 1e9, 1), "adjoint_gflop": round((dft + slant) / 1e9, 1),
                     "forward_tflops": round((dft + slant) / (pick_ms * 1e-3) / 1e12, 2),
                     "adjoint_tflops": round((dft + slant) / (amean * 1e-3) / 1e12, 2),
                     "peak_fp32_tflops": PEAK_MFMA_FP32_TFLOPS},
            "bf16x3": {"pick_ms": round(pick16, 3), "adjoint_ms_beta10": round(adj16, 3),
                       "forward_tflops": round((dft + slant) / (pick16 * 1e-3) / 1e12, 2),
                       "adjoint_tflops": round((dft + slant) / (adj16 * 1e-3) / 1e12, 2),
                       "right_picks_within_1_bin_of_fp32": round(same_bin, 4),
                       "what": "DFT and inverse-DFT GEMMs in BF16 split precision (bf16 hi + lo, 3 MFMA products); "
                               "slant stacks stay FP32"},
            "fk_norm": {"pick_ms": round(pick_fk, 3), "adjoint_ms_beta10": round(adj_fk, 3),
                        "what": "norm_trace + fk_filter on every window (one more DFT-sized GEMM and the window "
                                "Hilbert mix)"},
            "picks_within_1_bin_of_c_15_70hz": round(float(np.mean(np.concatenate(ok))), 4),
            "pick_columns_judged": len(ok),
            "what": "times include the small kernels of each stage (gather, |mm|, picks, softargmax, placement)"}


def serving_leg(rates=(2.0, 5.0, 10.0, 20.0), n=32, seed=0):
    """The serving loop: Poisson arrivals of forward + gradient requests, each
    a random model of the uneven set with 1..32 shots.  "batched" is the
    continuous-batching server (HBM-budget admission, up to 16 requests on the
    GPU, least remaining work first); "one_slot" the same server with one
    request on the GPU at a time; "per_request" the plain path, an Engine per
    request one after another.  Same arrival trace for all three."""
    from swd_amd import problems as P
    from swd_amd.serve import Server, poisson_trace, request_templates, run_trace, run_trace_per_request
    templates = request_templates(P.uneven_models(), shots=32)
    legs = {}
    same = True
    for rate in rates:
        trace = poisson_trace(templates, rate, n, seed)
        row = {}
        outs = {}
        for name, slots in (("batched", 16), ("one_slot", 1), ("per_request", 0)):
            if slots:
                srv = Server(max_active=slots, nt_block=(5, 4))
                r = run_trace(srv, templates, trace)
                st = srv.stats()
                srv.close()
            else:
                r = run_trace_per_request(templates, trace)
                st = {"peak_active": 1, "peak_bytes": 0}
            outs[name] = r["results"]
            progress(f"serving {rate:g} req/s {name}: p50 {r['p50_ms']:.0f} ms p99 {r['p99_ms']:.0f} ms")
            row[name] = {"p50_ms": round(r["p50_ms"], 1), "p99_ms": round(r["p99_ms"], 1),
                         "mean_queued_ms": round(r["mean_queued_ms"], 1), "req_per_s": round(r["req_per_s"], 2),
                         "gcell_steps_per_s": round(r["gcell_steps_per_s"], 2), "peak_active": st["peak_active"]}
            if slots:
                row[name]["peak_hbm_gb"] = round(st["peak_bytes"] / 1e9, 1)
        # every request: the server's result is the plain path's result, bitwise
        for a, b in zip(outs["batched"], outs["per_request"]):
            same &= bool(np.array_equal(a["dk_vs"], b["dk_vs"], equal_nan=True))
        legs[f"{rate:g}"] = row
    return {"rates_req_per_s": legs, "requests_per_rate": n, "alone_equals_in_mix": same,
            "what": "Poisson arrivals; request = forward + gradient of one of the 8 uneven models with 1..32 shots; "
                    "latency = arrival to result on the host; alone_equals_in_mix: every batched result equals the "
                    "per_request result bitwise"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=60)
    ap.add_argument("--nt", type=int, default=6000)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--time-block", type=int, default=0,
                    help="FD steps per launch (1: single-step kernels; 0: chosen by the autotuner)")
    ap.add_argument("--band-rows", type=int, default=0, help="band rows of the time-blocked kernels (0: automatic)")
    ap.add_argument("--no-graph", action="store_true", help="launch the FD phases eagerly instead of replaying HIP graphs")
    ap.add_argument("--plain-shots", type=int, default=2, help="shots for the unfused comparator (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven multi-model batch leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-dispersion", action="store_true", help="skip the dispersion-path leg")
    ap.add_argument("--no-iteration", action="store_true", help="skip the device-resident inversion iteration")
    ap.add_argument("--no-field", action="store_true", help="skip the outcrop-survey leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): libswd has no multi-GPU "
              "path (shots would split across ranks with one image all-reduce per gradient)", file=sys.stderr)
        sys.exit(2)

    from swd_amd import problems as P
    from swd_amd.engine import Engine

    prob, mod = P.production_shots(nshots=args.shots, nt=args.nt)
    eng = Engine(prob)
    if args.tile_rows:
        eng.set_tile_rows(args.tile_rows)
    eng.set_wavelet(mod["s"])
    eng.set_model(mod["vp_true"], mod["vs_true"])
    eng.observe()                          # observed picks cr_0 / cr_0l (untimed, SWD.m:150-162)
    eng.set_model(mod["vp"], mod["vs"])
    tuned = None
    if args.time_block > 0:
        eng.set_time_block(args.time_block, args.band_rows)
    else:                                  # untimed: the autotuner picks the FD launch shape
        t0 = time.perf_counter()
        tune = eng.autotune()
        tuned = {"fwd": list(tune["fwd"]), "adj": list(tune["adj"]), "candidates": tune["candidates"],
                 "seconds": round(time.perf_counter() - t0, 2)}
        tt = eng.timings()                 # the snapshot placements the autotuner timed, and the one kept
        tuned["snapshot_placement_ms"] = tt["snapshot_placement_ms"]
        tuned["snapshot_placement_kept"] = tt["snapshot_placement_kept"]
    eng.set_graph(not args.no_graph)       # FD phases replay HIP graphs, captured at the first gradient
    for _ in range(args.warmup):
        eng.gradient()
    per_step, fwd_us, adj_us, stage = [], [], [], []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        g = eng.gradient()                 # synchronous: returns after the last kernel and copy
        per_step.append(time.perf_counter() - t0)
        t = eng.timings()
        fwd_us.append(t["fwd_kernel_us"])
        adj_us.append(t["adj_kernel_us"])
        stage.append(t)
    wall = time.perf_counter() - t_all
    info = eng.info()
    S, nt = prob["sx"].size, prob["nt"]
    cells = info["nzbc"] * info["nxbc"]
    cell_steps = S * cells * (nt + nt - 1)
    ms_step = 1000.0 * wall / args.steps
    value = cell_steps / (wall / args.steps) / 1e9

    # roofline of the dominant kernel (forward or adjoint march), its launch
    # duration event-timed in the run; frac = compulsory bytes per launch (what
    # a time-blocked launch must move, fd_compulsory_bytes_per_launch) over the
    # duration, against the 8 TB/s HBM peak
    f_us, a_us = float(np.mean(fwd_us)), float(np.mean(adj_us))
    kf, ka = int(stage[-1]["fwd_steps_per_launch"]), int(stage[-1]["adj_steps_per_launch"])
    copy_gbs, plain_copy_gbs = copy_peak()

    def roofline(adjoint):
        k, kus = (ka, a_us) if adjoint else (kf, f_us)
        kname = f"fd_{'adj' if adjoint else 'fwd'}_{'tb' if k > 1 else 'march'}"
        kshape = [k, int(stage[-1]["adj_band_rows" if adjoint else "fwd_band_rows"])]
        comp = fd_compulsory_bytes_per_launch(prob, S, adjoint, k)
        nominal = k * fd_bytes_per_launch(prob, S, adjoint)
        prof = measured_profile(kname, kshape, args)
        traffic = prof.get("bytes_per_launch")
        achieved = comp / (kus * 1e-6) / 1e9
        r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
             "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
             "kernel": kname, "launch_shape": kshape, "kernel_us": round(kus, 2), "steps_per_launch": k,
             "bytes_per_launch": int(comp),
             "bytes_model": "compulsory bytes of one launch: state read + written once (40 B per padded cell and "
                            "shot), coefficients once per model (20 B per padded cell), per imaging step 12 B of "
                            "snapshot per interior cell and shot, adjoint also 24 B of FP32 partial-image RMW and "
                            "48 B / 32 of FP64 flush, receivers 4 B per trace sample",
             "traffic_over_compulsory": round(traffic / comp, 3) if traffic else None,
             "hbm_gbs_measured": round(traffic / (kus * 1e-6) / 1e9, 1) if traffic else None,
             "hbm_frac_measured": round(traffic / (kus * 1e-6) / 1e9 / PEAK_HBM_GBS, 4) if traffic else None,
             "copy_peak_gbs": round(copy_gbs, 1),
             "traffic_over_copy_peak": round(traffic / (kus * 1e-6) / 1e9 / copy_gbs, 3) if traffic else None,
             "plain_copy_gbs": round(plain_copy_gbs, 1),
             "copy_model": "copy_peak_gbs: 1 GiB copied by the fastest streaming copy (non-temporal 16-byte loads "
                           "and stores, four in flight per lane); plain_copy_gbs: the grid-stride float4 copy that "
                           "calibrates the traffic counters",
             "nominal_bytes_per_launch": int(nominal),
             "nominal_gbs": round(nominal / (kus * 1e-6) / 1e9, 1),
             "nominal_model": "k x the unblocked per-step model (60 B per padded cell-step + snapshot and image "
                              "streams); exceeds the HBM peak because time blocking keeps the intermediate steps "
                              "on chip"}
        if prof.get("valu_insts"):
            r["valu_busy_frac"] = round(prof["valu_insts"] * 4.0 / (N_SIMD * CLOCK_HZ * kus * 1e-6), 3)
            r["valu_model"] = ("SQ_INSTS_VALU per launch x 4 cycles per wave64 instruction over 1024 SIMDs at the "
                               "2.4 GHz peak engine clock, from the committed SQ counter pass")
        return r

    adj_dominant = a_us * (nt - 1) / ka > f_us * nt / kf
    roof = roofline(adj_dominant)
    roof_other = roofline(not adj_dominant)

    out = {"metric": "Gcell-steps/s FD fwd+adj (% HBM roofline), shots-grad/s; p50/p99 req latency",
           "value": round(value, 3), "unit": "Gcell-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (checkerboard true model, homogeneous current model, Ricker 25 Hz)",
           "config": {"workload": "single-model production shots: 60 shots, 100x600 cells (132x660 padded), "
                                  "nt=6000, dt=8e-5, 600 receivers; Radon 151 freq x 701 vel; beta=10",
                      "shots": S, "nt": nt, "grid": [info["nzbc"], info["nxbc"]], "parallelism": "1 GPU, shot batch"},
           "interior_gcell_steps_per_s": round(S * prob["nz"] * prob["nx"] * (nt + nt - 1) / (wall / args.steps) / 1e9, 3),
           "shots_grad_per_s": round(S / (wall / args.steps), 3),
           "gradient_latency_ms": {"p50": round(1000 * float(np.percentile(per_step, 50)), 2),
                          "p99": round(1000 * float(np.percentile(per_step, 99)), 2)},
           "stage_ms": {k: round(float(np.mean([s[k] for s in stage])), 2)
                        for k in ("fwd_ms", "disp_ms", "adj_ms", "asm_ms")},
           "roofline": roof,
           "roofline_other_phase": roof_other,
           "fd_launch": {"fwd": [kf, int(stage[-1]["fwd_band_rows"])], "adj": [ka, int(stage[-1]["adj_band_rows"])],
                         "autotune": tuned, "hip_graph": not args.no_graph},
           "device_bytes": info["device_bytes"],
           "residual_r_mean": float(np.mean(g["res_r"]))}

    if args.plain_shots > 0:
        q2, m2 = P.production_shots(nshots=args.plain_shots, nt=args.nt)
        e2 = Engine(q2)
        e2.set_wavelet(m2["s"])
        e2.set_observed_picks(np.ones((e2.npair, args.plain_shots)), np.ones((e2.npair, args.plain_shots)))
        e2.set_model(m2["vp"], m2["vs"])
        e2.set_fd_path(1)
        t0 = time.perf_counter()
        e2.gradient()
        tp = time.perf_counter() - t0
        plain = args.plain_shots * cells * (nt + nt - 1) / tp / 1e9
        out["plain_path"] = {"value": round(plain, 3), "unit": "Gcell-steps/s",
                             "what": f"unfused per-statement kernels, one shot per launch, {args.plain_shots} shots",
                             "speedup": round(value / plain, 2)}
        e2.close()
    if not args.no_iteration:
        out["iteration"] = iteration_leg(eng, mod, prob)
        progress(f"iteration: {out['iteration']['ms']} ms")
    eng.close()
    progress(f"headline: {value:.1f} Gcell-steps/s")
    if not args.no_dispersion:
        out["dispersion"] = dispersion_leg()
        progress(f"dispersion: picks {out['dispersion']['pick_ms']} ms for 60 shots")
    if not args.no_field:
        out["field_outcrop"] = field_leg(graph=not args.no_graph)
        progress(f"outcrop survey: {out['field_outcrop']['ms_per_gradient']} ms per gradient")
    if not args.no_uneven:
        out["uneven_batch"] = uneven_leg(graph=not args.no_graph)
    if not args.no_serving:
        out["serving"] = serving_leg()
        mid = out["serving"]["rates_req_per_s"].get("10", {}).get("batched")
        if mid:
            out["req_latency_ms"] = {"rate_req_per_s": 10, "p50": mid["p50_ms"], "p99": mid["p99_ms"]}
    if not args.no_cpu_baseline:
        progress("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(prob, mod)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
