"""Python handle of the continuous-batching gradient server (swd_server_* in
include/swd_abi.h) and a Poisson load generator for it.

A request is one forward + gradient solve of SWD.m:171-201 (all its shots)
for one model.  The native server admits requests in arrival order while a
slot is free and their buffers fit in the device budget, advances the shots
of every admitted request in shared FD launches (requests join and leave the
launch sequence independently, earliest deadline first when they do not
all fit the GPU at once) and runs each request's Radon / picking /
adjoint-source and assembly stages on side streams.  ``max_active=1`` keeps
one request on the GPU at a time; :func:`run_trace_per_request` is the plain
per-request path (an Engine per request, one after another).
"""
from __future__ import annotations

import ctypes
import threading
import time

import numpy as np

from ._lib import check, fortran, load, ptr
from .engine import Engine, make_problem, problem_npair

ERR_BUSY = -8


class Server:
    def __init__(self, hbm_bytes: int = 0, max_active: int = 16, nt_block=(4, 4)):
        h = ctypes.c_void_p()
        check(load().swd_server_create(int(hbm_bytes), int(max_active), int(nt_block[0]), int(nt_block[1]),
                                       ctypes.byref(h)))
        self.h = h
        self._meta = {}

    def submit(self, prob: dict, vp, vs, wavelet, cr0_r, cr0_l) -> int:
        """Queue one forward + gradient solve; returns its ticket."""
        p, keep = make_problem(prob)
        t = ctypes.c_longlong(0)
        check(load().swd_server_submit(self.h, ctypes.byref(p), ptr(fortran(vp, np.float64)),
                                       ptr(fortran(vs, np.float64)),
                                       ptr(np.ascontiguousarray(wavelet, dtype=np.float64)),
                                       ptr(fortran(cr0_r, np.float64)), ptr(fortran(cr0_l, np.float64)),
                                       ctypes.byref(t)))
        self._meta[t.value] = (prob["nz"], prob["nx"], int(np.size(prob["sx"])))
        return t.value

    def wait(self, ticket: int, timeout: float | None = None) -> dict:
        """Result of a request: dk_vs (nz x nx), res_r / res_l, and its
        queued / on-GPU / total times (ms) as measured by the server."""
        nz, nx, S = self._meta[ticket]
        dk = np.zeros((nz, nx), order="F")
        rr, rl, tm = np.zeros(S), np.zeros(S), np.zeros(4)
        rc = load().swd_server_wait(self.h, ticket, -1.0 if timeout is None else 1000.0 * timeout, ptr(dk), ptr(rr),
                                    ptr(rl), ptr(tm))
        if rc == ERR_BUSY and b"still running" in (load().swd_last_error() or b""):
            raise TimeoutError(f"request {ticket} still running")
        del self._meta[ticket]
        check(rc)   # a cancelled request raises SwdError (SWD_ERR_BUSY)
        return {"dk_vs": dk, "res_r": rr, "res_l": rl, "queued_ms": tm[0], "gpu_ms": tm[1], "total_ms": tm[2],
                "submit_ms": tm[3]}

    def cancel(self, ticket: int) -> None:
        """Cancel a queued or running request (swd_server_cancel); its wait()
        then raises SwdError with SWD_ERR_BUSY."""
        check(load().swd_server_cancel(self.h, int(ticket)))

    def set_deadline_slack(self, slack: float) -> None:
        """Selection deadlines = arrival + slack x work at the nominal rate."""
        check(load().swd_server_set_deadline_slack(self.h, float(slack)))

    def set_wave_slots(self, waves: int) -> None:
        """Resident waves the launch selection fills (0: the device's count)."""
        check(load().swd_server_set_wave_slots(self.h, int(waves)))

    def stats(self) -> dict:
        v = np.zeros(10, dtype=np.int64)
        check(load().swd_server_stats(self.h, ptr(v), 10))
        keys = ("queued", "active", "completed", "failed", "bytes_in_use", "budget", "fd_launches", "peak_active",
                "peak_bytes", "rebuilds")
        return dict(zip(keys, (int(x) for x in v)))

    def close(self):
        if getattr(self, "h", None):
            load().swd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def subset(prob: dict, k: int) -> dict:
    """The problem restricted to its first k shots."""
    q = dict(prob)
    for key in ("sx", "sz", "is_index"):
        if q.get(key) is not None:
            q[key] = np.asarray(q[key])[:k]
    return q


def request_templates(models, shots: int = 32):
    """Observed picks of each model with `shots` shots (one engine at a time);
    a request of k shots of a model uses the first k shots and the first k
    pick columns (a shot's picks depend on that shot's gather only)."""
    out = []
    for prob, mod in models:
        q = dict(prob)
        sx = np.linspace(1, q["nx"], shots).round().astype(int)
        q.update(sx=sx, sz=np.ones(shots, dtype=int), is_index=sx)
        e = Engine(q)
        e.set_wavelet(mod["s"])
        e.set_model(mod["vp_true"], mod["vs_true"])
        cr_r, cr_l = e.observe()
        e.close()
        out.append((q, mod, cr_r, cr_l))
    return out


def cell_steps(prob: dict) -> float:
    nzbc = prob["nz"] + 2 + prob["nbc"]
    nxbc = prob["nx"] + 2 * prob["nbc"]
    return float(np.size(prob["sx"])) * nzbc * nxbc * (2 * prob["nt"] - 1)


def poisson_trace(templates, rate: float, n: int, seed: int = 0, max_shots: int = 32):
    """n requests: Poisson arrival times at `rate` per second, a random model
    and 1..max_shots shots each."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    t -= t[0]
    return [(float(t[i]), int(rng.integers(len(templates))), int(rng.integers(1, max_shots + 1)))
            for i in range(n)]


def run_trace(server: Server, templates, trace) -> dict:
    """Submit the trace's requests at their arrival times from a separate
    thread, wait for all, and summarise latency and throughput."""
    tickets = [None] * len(trace)
    errors = []

    def submitter():
        t0 = time.perf_counter()
        try:
            for i, (ta, ti, k) in enumerate(trace):
                d = ta - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                q, mod, cr_r, cr_l = templates[ti]
                tickets[i] = server.submit(subset(q, k), mod["vp"], mod["vs"], mod["s"], cr_r[:, :k], cr_l[:, :k])
        except Exception as ex:   # reported by the caller
            errors.append(ex)

    th = threading.Thread(target=submitter)
    t_start = time.perf_counter()
    th.start()
    results = []
    for i in range(len(trace)):
        while tickets[i] is None:
            if errors:
                th.join()
                raise errors[0]
            time.sleep(0.0005)
        results.append(server.wait(tickets[i]))
    th.join()
    wall = time.perf_counter() - t_start
    span_ms = max(r["submit_ms"] + r["total_ms"] for r in results) - min(r["submit_ms"] for r in results)
    return _summary(trace, templates, [r["total_ms"] for r in results], span_ms, [r["queued_ms"] for r in results],
                    wall, results)


def _summary(trace, templates, lat_ms, span_ms, queued_ms, wall, results):
    work = sum(cell_steps(subset(templates[ti][0], k)) for _, ti, k in trace)
    lat = np.asarray(lat_ms)
    return {"requests": len(trace), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_queued_ms": float(np.mean(queued_ms)), "req_per_s": len(trace) / (span_ms / 1000.0),
            "gcell_steps_per_s": work / (span_ms / 1000.0) / 1e9, "wall_s": wall, "results": results}


def run_trace_per_request(templates, trace) -> dict:
    """The plain per-request path on the same trace: each request, in arrival
    order, through its own Engine (allocate, upload, gradient, free), one at
    a time; latency from its arrival time to its result."""
    t0 = time.perf_counter()
    lat, queued, results = [], [], []
    for ta, ti, k in trace:
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        start = time.perf_counter() - t0
        q, mod, cr_r, cr_l = templates[ti]
        e = Engine(subset(q, k))
        e.set_wavelet(mod["s"])
        e.set_observed_picks(cr_r[:, :k], cr_l[:, :k])
        e.set_model(mod["vp"], mod["vs"])
        g = e.gradient()
        e.close()
        done = time.perf_counter() - t0
        lat.append(1000.0 * (done - ta))
        queued.append(1000.0 * (start - ta))
        results.append(g)
    wall = time.perf_counter() - t0
    return _summary(trace, templates, lat, 1000.0 * wall, queued, wall, results)
