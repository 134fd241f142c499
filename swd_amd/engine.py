"""Python handle of the batched device-resident SWD engine (swd_engine_* in
include/swd_abi.h).  One Engine = one model with all the shots of an SWD
iteration resident in HBM; ``gradient()`` is the parfor body of
SWD.m:171-194 for every shot plus the assembly of SWD.m:199-201."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, fortran, load, ptr

_I, _D, _P = ctypes.c_int, ctypes.c_double, ctypes.c_void_p


class Problem(ctypes.Structure):
    _fields_ = [("nz", _I), ("nx", _I), ("nbc", _I), ("fd_order", _I), ("nt", _I), ("in_wf", _I),
                ("dx", _D), ("dt", _D), ("nshots", _I), ("sx", _P), ("sz", _P),
                ("ng", _I), ("gx0", _I), ("dg", _I), ("gz", _I),
                ("df", _D), ("fmin", _D), ("fmax", _D), ("vmin", _D), ("vmax", _D), ("ds", _D),
                ("offset", _I), ("m", _I), ("M", _I), ("ini", _I), ("w", _D), ("beta", _D),
                ("is_index", _P), ("ns", _I)]


class Iteration(ctypes.Structure):
    """swd_iteration of include/swd_abi.h: parameters, carried state and results of one SWD iteration."""
    _fields_ = [("pur", _D), ("vsmin", _D), ("vsmax", _D), ("smooth_z", _I), ("smooth_x", _I), ("f1", _D),
                ("f2", _D), ("res2", _D), ("residual", _D), ("res0", _D), ("alpha", _D), ("gama", _D),
                ("res1", _D), ("res3", _D), ("forwards", _I)]


PROBLEM_KEYS = ("nz", "nx", "nbc", "fd_order", "nt", "in_wf", "dx", "dt", "ng", "gx0", "dg", "gz", "df", "fmin",
                "fmax", "vmin", "vmax", "ds", "offset", "m", "M", "ini", "w", "beta", "ns")

def lib():
    return load()


def make_problem(q: dict):
    """The swd_problem struct of a problem dict, with the int32 arrays it
    points to (keep them alive as long as the struct is used)."""
    sx = np.ascontiguousarray(q["sx"], dtype=np.int32)
    sz = np.ascontiguousarray(q["sz"], dtype=np.int32)
    isx = None if q.get("is_index") is None else np.ascontiguousarray(q["is_index"], dtype=np.int32)
    p = Problem()
    for k in PROBLEM_KEYS:
        setattr(p, k, q.get(k, 0) if k != "fd_order" else q.get(k, 22))
    p.nshots = int(sx.size)
    p.sx = sx.ctypes.data
    p.sz = sz.ctypes.data
    p.is_index = None if isx is None else isx.ctypes.data
    return p, (sx, sz, isx)


def problem_npair(q: dict) -> int:
    """Radon frequency bins of a problem (rows of its pick arrays)."""
    p, keep = make_problem(q)
    n = ctypes.c_int(0)
    check(lib().swd_problem_npair(ctypes.byref(p), ctypes.byref(n)))
    return int(n.value)


class Engine:
    """Batched SWD engine for one model (FP32 FD + Radon, FP64 reductions)."""

    def __init__(self, problem: dict):
        self.problem = dict(problem)
        p, (self._sx, self._sz, self._is) = make_problem(self.problem)
        self._p = p
        h = ctypes.c_void_p()
        check(lib().swd_engine_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        info = self.info()
        self.npair = int(info["npair"])
        self.nshots = p.nshots

    def close(self):
        if getattr(self, "h", None):
            lib().swd_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        v = np.zeros(10, dtype=np.int64)
        check(lib().swd_engine_info(self.h, ptr(v), 10))
        keys = ("nzbc", "nxbc", "pitch", "nrec", "npair", "np", "device_bytes", "ccn", "lf", "tile_rows")
        return dict(zip(keys, (int(x) for x in v)))

    def set_model(self, vp, vs):
        self._vs = fortran(vs, np.float64)
        check(lib().swd_engine_set_model(self.h, ptr(fortran(vp, np.float64)), ptr(self._vs)))

    def set_wavelet(self, s):
        check(lib().swd_engine_set_wavelet(self.h, ptr(np.ascontiguousarray(s, dtype=np.float64))))

    def set_observed_picks(self, cr0_r, cr0_l):
        a = fortran(cr0_r, np.float64)
        b = fortran(cr0_l, np.float64)
        check(lib().swd_engine_set_observed_picks(self.h, ptr(a), ptr(b)))

    def observe(self):
        a = np.zeros((self.npair, self.nshots), order="F")
        b = np.zeros((self.npair, self.nshots), order="F")
        check(lib().swd_engine_observe(self.h, ptr(a), ptr(b)))
        return a, b

    def gradient(self) -> dict:
        q = self.problem
        dk = np.zeros((q["nz"], q["nx"]), order="F")
        rr, rl = np.zeros(self.nshots), np.zeros(self.nshots)
        cr = np.zeros((self.npair, self.nshots), order="F")
        cl = np.zeros((self.npair, self.nshots), order="F")
        check(lib().swd_engine_gradient(self.h, ptr(dk), ptr(rr), ptr(rl), ptr(cr), ptr(cl)))
        return {"dk_vs": dk, "res_r": rr, "res_l": rl, "cr_r": cr, "cr_l": cl}

    def forward_picks(self) -> dict:
        rr, rl = np.zeros(self.nshots), np.zeros(self.nshots)
        cr = np.zeros((self.npair, self.nshots), order="F")
        cl = np.zeros((self.npair, self.nshots), order="F")
        check(lib().swd_engine_forward_picks(self.h, ptr(rr), ptr(rl), ptr(cr), ptr(cl)))
        return {"res_r": rr, "res_l": rl, "cr_r": cr, "cr_l": cl}

    STAGES = {"forward": 1, "dispersion": 2, "adjoint_source": 3, "adjoint": 4, "assemble": 5}

    def iterate(self, vsmin: float, vsmax: float, pur: float = -0.05, smooth_z: int = 1, smooth_x: int = 2,
                state: dict | None = None) -> dict:
        """One SWD iteration on the device (swd_engine_iterate): gradient,
        smooth2a, line search and model update.  ``state`` carries f1 / f2 /
        res2 between iterations (f1 = 0.5 at the first); the returned dict
        holds the results and the new state."""
        st = state or {}
        it = Iteration(pur=pur, vsmin=vsmin, vsmax=vsmax, smooth_z=smooth_z, smooth_x=smooth_x,
                       f1=st.get("f1", 0.5), f2=st.get("f2", float("nan")), res2=st.get("res2", float("nan")))
        check(lib().swd_engine_iterate(self.h, ctypes.byref(it)))
        return {name: getattr(it, name) for name, _ in Iteration._fields_}

    def model(self) -> np.ndarray:
        """The engine's current vs (nz x nx)."""
        q = self.problem
        vs = np.zeros((q["nz"], q["nx"]), order="F")
        check(lib().swd_engine_read_model(self.h, ptr(vs)))
        return vs

    def run(self, stage: str):
        """One stage of gradient() on the resident buffers."""
        check(lib().swd_engine_run(self.h, self.STAGES[stage]))

    def set_picks(self, cr_r=None, cr_l=None):
        """Overwrite the current picks (npair x nshots) between 'dispersion' and 'adjoint_source'."""
        a = None if cr_r is None else fortran(cr_r, np.float64)
        b = None if cr_l is None else fortran(cr_l, np.float64)
        check(lib().swd_engine_set_picks(self.h, ptr(a), ptr(b)))

    def set_variant(self, variant="swd"):
        """Field-script variant of the dispersion passes and the model update
        (swd_engine_set_variant): a name of :data:`swd_amd.field.VARIANTS`
        or a dict of its fields."""
        from .field import variant_struct
        v = variant_struct(variant)
        check(lib().swd_engine_set_variant(self.h, ctypes.byref(v)))
        self.variant = v

    def set_graph(self, on: bool = True) -> bool:
        """Replay each FD phase as a captured HIP graph (swd_engine_set_graph); returns the previous setting."""
        return bool(lib().swd_engine_set_graph(self.h, int(bool(on))))

    def set_fd_path(self, mode: int) -> int:
        return lib().swd_engine_set_fd_path(self.h, int(mode))

    def set_tile_rows(self, tz: int):
        check(lib().swd_engine_set_tile_rows(self.h, int(tz)))

    def set_time_block(self, nt_block: int, rows: int = 0, phase: int = -1):
        """Steps per fused FD launch (1 = single-step kernels) and band rows
        (0 = chosen from occupancy) of the forward (phase 0), adjoint (1) or
        both (-1) FD phases."""
        check(lib().swd_engine_set_phase_time_block(self.h, int(phase), int(nt_block), int(rows)))

    def autotune(self, steps: int = 60) -> dict:
        """Time the FD march settings on this engine's shots and install the
        fastest (see :func:`batch_autotune`)."""
        return batch_autotune([self], steps)

    def timings(self) -> dict:
        v = np.zeros(15)
        check(lib().swd_engine_timings(self.h, ptr(v), 15))
        out = dict(zip(("fwd_ms", "disp_ms", "adj_ms", "asm_ms", "fwd_kernel_us", "adj_kernel_us",
                        "fwd_steps_per_launch", "adj_steps_per_launch", "fwd_band_rows", "adj_band_rows"), v))
        # snapshot placements timed by autotune(): full forward + adjoint march ms each, and the one kept
        n = int(v[10])
        out["snapshot_placement_ms"] = [round(float(x), 2) for x in v[12:12 + n]]
        out["snapshot_placement_kept"] = int(v[11])
        return out

    def read(self, what: str) -> np.ndarray:
        q = self.problem
        S, nt, ng, nz, nx = self.nshots, q["nt"], q["ng"], q["nz"], q["nx"]
        info = self.info()
        shapes = {"gathers": (0, np.float32, (S, nt, ng)), "adjoint_sources": (1, np.float32, (S, nt, ng)),
                  "gradient": (2, np.float64, (4, nx, nz)), "spectra": (3, np.float32, (2 * S, self.npair, info["np"])),
                  "images": (4, np.float64, (S, 3, nz, nx)), "smoothed_gradient": (5, np.float64, (nx, nz))}
        code, dt, shape = shapes[what]
        out = np.zeros(shape, dtype=dt)
        check(lib().swd_engine_read(self.h, code, ptr(out), out.nbytes))
        if what == "gradient":
            return np.transpose(out, (0, 2, 1))     # -> [4][nz][nx]
        if what == "smoothed_gradient":
            return out.T                             # column-major nz x nx
        return out


def batch_autotune(engines, steps: int = 60) -> dict:
    """Autotune the FD march of a batch of engines: every time block 1..6 and
    band heights splitting the tallest grid evenly are timed over `steps`
    forward and adjoint steps on the engines' own buffers (several engines:
    the time blocks only, band heights from occupancy per launch table,
    reported as rows 0); the fastest forward and adjoint settings are
    installed on every engine.  Overwrites FD state, gathers, snapshots,
    adjoint sources and images."""
    n = len(engines)
    arr = (ctypes.c_void_p * n)(*[e.h.value for e in engines])
    choice = np.zeros(4, dtype=np.int32)
    table = np.zeros((256, 4))
    nrows = np.zeros(1, dtype=np.int32)
    check(lib().swd_batch_autotune(arr, n, int(steps), ptr(choice), ptr(table), table.shape[0], ptr(nrows)))
    t = table[: min(int(nrows[0]), table.shape[0])]
    return {"fwd": (int(choice[0]), int(choice[1])), "adj": (int(choice[2]), int(choice[3])),
            "table": t, "candidates": int(nrows[0])}


def batch_gradient(engines) -> tuple:
    """Gradients of several engines with shared FD launches (uneven models)."""
    n = len(engines)
    arr = (ctypes.c_void_p * n)(*[e.h.value for e in engines])
    outs = [np.zeros((e.problem["nz"], e.problem["nx"]), order="F") for e in engines]
    parr = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
    us = np.zeros(2)
    check(lib().swd_batch_gradient(arr, n, parr, ptr(us)))
    return outs, us
