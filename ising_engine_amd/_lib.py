"""ctypes binding of the C ABI in include/ising_gpu.h.

The shared library is built in-tree (``ising_engine_amd/libising_gpu.so``).
There is no CPU fallback: if the library is missing or cannot be loaded this
module raises, so a run can never silently measure or test something else.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libising_gpu.so")

IGE_OK = 0
IGE_ERR_INVALID = -1
IGE_ERR_HIP = -2
IGE_ERR_NOMEM = -3
IGE_ERR_UNSUPPORTED = -4
IGE_ERR_RANGE = -5

KERNEL_AUTO = 0
KERNEL_GENERIC = 1
KERNEL_MULTISPIN = 2
KERNEL_DENSE = 3
RUN_EXCHANGE = 1
RUN_TRACK_BEST = 2
RUN_HEATBATH = 4
RUN_ICM = 8
RUN_EAGER = 16
STEPS_UNLIMITED = 0xFFFFFFFF

EXPORTED = (
    "ige_context_create", "ige_context_destroy", "ige_last_error", "ige_version",
    "ige_model_from_pairs", "ige_model_destroy", "ige_model_get_info",
    "ige_model_coloring", "ige_model_order", "ige_model_csr", "ige_model_linear", "ige_state_create", "ige_state_destroy",
    "ige_set_ladder", "ige_set_replica_temperatures", "ige_exchange_replicas", "ige_run",
    "ige_exchange", "ige_synchronize", "ige_read_spins", "ige_write_spins",
    "ige_read_energies", "ige_read_temp_of", "ige_read_fields",
    "ige_recompute_energies", "ige_read_stats", "ige_profile_enable",
    "ige_profile_read", "ige_anneal_schedule", "ige_ultimate_ladder", "ige_batch_create", "ige_batch_destroy",
    "ige_batch_reset", "ige_batch_run", "ige_batch_autotune", "ige_batch_synchronize",
    "ige_batch_info", "ige_batch_read", "ige_batch_energies", "ige_batch_timing", "ige_server_create", "ige_server_destroy",
    "ige_server_submit", "ige_server_poll", "ige_server_wait", "ige_server_stats",
    "ige_batch_configure", "ige_server_configure", "ige_steepest_descent", "ige_steepest_descent_noalloc", "ige_update_best",
    "ige_read_best", "ige_model_polish", "ige_server_solution", "ige_problem_parse",
    "ige_problem_destroy", "ige_problem_get_info", "ige_problem_pairs", "ige_problem_energy",
    "ige_icm", "ige_set_populations", "ige_resample", "ige_pa_normalize",
    "ige_pa_free_energy_increment", "ige_pa_systematic_resample", "ige_pa_parents",
    "ige_pa_family_entropy",
    "ige_pa_adapt_spacings",
    "ige_pa_tune_ladder",
)

FORMAT_RUDY = 0
FORMAT_GSET = 1
FORMAT_ORLIB = 2
FORMAT_BIQMAC = 3
FORMAT_QPLIB = 4

REQ_QUEUED = 0
REQ_RUNNING = 1
REQ_DONE = 2
REQ_FAILED = -1

BATCH_PERSISTENT = 0
BATCH_PLAIN = 1


class ModelInfo(ctypes.Structure):
    _fields_ = [
        ("n", c_uint32), ("num_colors", c_uint32), ("nnz", c_uint64),
        ("max_degree", c_uint32), ("multispin_eligible", c_int32),
        ("coupling_magnitude", c_int32), ("dense_scale", c_int32),
        ("max_local_field", c_int64),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("sweeps", c_uint64), ("exchanges", c_uint64), ("accepted", c_uint64),
        ("proposed", c_uint64),
        ("phase", c_uint32), ("kernel", c_uint32),
        ("graph_sweeps", c_uint64), ("graph_fallbacks", c_uint64), ("resident_sweeps", c_uint64),
        ("chain_sweeps", c_uint64),
    ]


class InstanceDesc(ctypes.Structure):
    _fields_ = [
        ("n", c_uint32), ("offset", c_int64), ("linear", POINTER(c_int32)),
        ("npairs", c_uint64), ("pi", POINTER(c_uint32)), ("pj", POINTER(c_uint32)),
        ("q", POINTER(c_int32)), ("seed", c_uint64), ("sweeps", c_uint32),
        ("beta_lo", c_double), ("beta_hi", c_double), ("track_every", c_uint32),
        ("polish", c_uint32), ("num_temps", c_uint32), ("temp_max", c_double),
        ("temp_min", c_double),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("id", c_uint64), ("status", c_int32), ("n", c_uint32), ("best_energy", c_int64),
        ("best_replica", c_uint32), ("cells", c_uint32), ("accepted", c_uint64),
        ("latency_ms", c_double), ("queue_ms", c_double), ("solution_energy", c_int64),
        ("solution_sweep", c_uint32), ("polish_flips", c_uint32),
    ]


class ServerCounters(ctypes.Structure):
    _fields_ = [
        ("queued", c_uint64), ("live", c_uint64), ("completed", c_uint64),
        ("rounds", c_uint64), ("attempts", c_uint64), ("arena_used", c_uint64),
        ("arena_capacity", c_uint64), ("max_live", c_uint64), ("busy_ms", c_double),
    ]


class DescentReport(ctypes.Structure):
    _fields_ = [("steps", c_uint32), ("converged", c_uint32), ("flips", c_uint64),
                ("proposed", c_uint64)]


class ProblemInfo(ctypes.Structure):
    _fields_ = [("n", c_uint32), ("problems_in_file", c_uint32), ("npairs", c_uint64),
                ("maximize", c_int32), ("integral", c_int32), ("offset", c_double)]


class IgeError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"ige error {code}: {message}")
        self.code = code


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(POINTER(ctype))


_lib = None


def load() -> ctypes.CDLL:
    """Load and type the library once; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    vp = c_void_p
    sig = {
        "ige_context_create": (c_int32, [c_int32, POINTER(vp)]),
        "ige_context_destroy": (c_int32, [vp]),
        "ige_last_error": (c_char_p, []),
        "ige_version": (c_char_p, []),
        "ige_model_from_pairs": (c_int32, [vp, c_uint32, c_int64, POINTER(c_int32), c_uint64,
                                           POINTER(c_uint32), POINTER(c_uint32),
                                           POINTER(c_int32), POINTER(vp)]),
        "ige_model_destroy": (c_int32, [vp]),
        "ige_model_get_info": (c_int32, [vp, POINTER(ModelInfo)]),
        "ige_model_coloring": (c_int32, [vp, POINTER(c_uint32)]),
        "ige_model_order": (c_int32, [vp, POINTER(c_uint32), POINTER(c_uint32)]),
        "ige_model_linear": (c_int32, [vp, POINTER(c_int32), POINTER(c_int64)]),
        "ige_model_csr": (c_int32, [vp, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_int32)]),
        "ige_state_create": (c_int32, [vp, vp, c_uint32, c_uint64, POINTER(c_uint64), c_uint32,
                                       POINTER(vp)]),
        "ige_state_destroy": (c_int32, [vp]),
        "ige_set_ladder": (c_int32, [vp, POINTER(c_double)]),
        "ige_set_replica_temperatures": (c_int32, [vp, POINTER(c_double), c_uint32]),
        "ige_exchange_replicas": (c_int32, [vp, c_uint32]),
        "ige_run": (c_int32, [vp, c_uint32, c_uint32, POINTER(c_double)]),
        "ige_exchange": (c_int32, [vp, c_uint32]),
        "ige_synchronize": (c_int32, [vp]),
        "ige_read_spins": (c_int32, [vp, POINTER(c_uint64)]),
        "ige_write_spins": (c_int32, [vp, POINTER(c_uint64)]),
        "ige_read_energies": (c_int32, [vp, POINTER(c_int64)]),
        "ige_read_temp_of": (c_int32, [vp, POINTER(c_uint32)]),
        "ige_read_fields": (c_int32, [vp, POINTER(c_int32)]),
        "ige_recompute_energies": (c_int32, [vp, POINTER(c_int64)]),
        "ige_read_stats": (c_int32, [vp, POINTER(Stats), POINTER(c_uint64), POINTER(c_uint64)]),
        "ige_profile_enable": (c_int32, [vp, c_int32]),
        "ige_anneal_schedule": (c_int32, [c_double, c_double, c_uint32, POINTER(c_double)]),
        "ige_ultimate_ladder": (c_int32, [c_double, c_double, c_uint32, POINTER(c_double)]),
        "ige_batch_create": (c_int32, [vp, c_uint32, POINTER(InstanceDesc), POINTER(vp)]),
        "ige_batch_destroy": (c_int32, [vp]),
        "ige_batch_reset": (c_int32, [vp]),
        "ige_batch_run": (c_int32, [vp, c_uint32, c_uint32, c_uint32, c_uint32]),
        "ige_batch_autotune": (c_int32, [vp, c_uint32, POINTER(c_uint32), POINTER(c_double)]),
        "ige_batch_synchronize": (c_int32, [vp]),
        "ige_batch_info": (c_int32, [vp, POINTER(c_uint64), POINTER(c_uint64)]),
        "ige_batch_read": (c_int32, [vp, c_uint32, POINTER(c_uint64), POINTER(c_int64),
                                     POINTER(c_uint64)]),
        "ige_batch_energies": (c_int32, [vp, POINTER(c_int64), POINTER(c_int64)]),
        "ige_batch_timing": (c_int32, [vp, POINTER(c_double), POINTER(c_double)]),
        "ige_server_create": (c_int32, [vp, c_uint64, c_double, c_uint32, POINTER(vp)]),
        "ige_server_destroy": (c_int32, [vp]),
        "ige_server_submit": (c_int32, [vp, POINTER(InstanceDesc), POINTER(c_uint64)]),
        "ige_server_poll": (c_int32, [vp, c_uint64, c_int32, POINTER(RequestInfo),
                                      POINTER(c_uint64), POINTER(c_int64)]),
        "ige_server_wait": (c_int32, [vp, c_uint64, c_double, POINTER(RequestInfo)]),
        "ige_server_stats": (c_int32, [vp, POINTER(ServerCounters)]),
        "ige_batch_configure": (c_int32, [vp, c_uint32, c_uint32, c_uint32]),
        "ige_server_configure": (c_int32, [vp, c_uint32, c_uint32, c_uint32]),
        "ige_profile_read": (c_int32, [vp, POINTER(c_double), POINTER(c_uint64),
                                       POINTER(c_double), POINTER(c_uint64)]),
        "ige_steepest_descent": (c_int32, [vp, c_uint32, POINTER(ctypes.c_uint8),
                                           POINTER(DescentReport)]),
        "ige_steepest_descent_noalloc": (c_int32, [vp, c_uint32, POINTER(ctypes.c_uint8),
                                                   POINTER(DescentReport)]),
        "ige_update_best": (c_int32, [vp]),
        "ige_icm": (c_int32, [vp, c_uint32, POINTER(c_uint64)]),
        "ige_set_populations": (c_int32, [vp, c_uint32]),
        "ige_resample": (c_int32, [vp, POINTER(c_uint32)]),
        "ige_pa_normalize": (c_int32, [POINTER(c_double), c_uint32, POINTER(c_double),
                                       POINTER(c_double)]),
        "ige_pa_free_energy_increment": (c_int32, [POINTER(c_double), c_uint32,
                                                   POINTER(c_double)]),
        "ige_pa_systematic_resample": (c_int32, [POINTER(c_double), c_uint32, c_double,
                                                 POINTER(c_uint32)]),
        "ige_pa_parents": (c_int32, [POINTER(c_uint32), c_uint32, c_uint64, c_uint32,
                                     POINTER(c_uint32)]),
        "ige_pa_family_entropy": (c_int32, [POINTER(c_uint32), c_uint32, POINTER(c_double),
                                            POINTER(c_uint32)]),
        "ige_pa_adapt_spacings": (c_int32, [POINTER(c_double), POINTER(c_double), c_uint32,
                                            POINTER(c_double)]),
        "ige_pa_tune_ladder": (c_int32, [c_void_p, c_double, c_double, c_uint32, c_uint32,
                                         POINTER(c_double)]),
        "ige_problem_parse": (c_int32, [c_char_p, c_uint64, c_uint32, c_uint32, POINTER(vp)]),
        "ige_problem_destroy": (c_int32, [vp]),
        "ige_problem_get_info": (c_int32, [vp, POINTER(ProblemInfo)]),
        "ige_problem_pairs": (c_int32, [vp, POINTER(c_int64), POINTER(c_int32),
                                        POINTER(c_uint32), POINTER(c_uint32), POINTER(c_int32)]),
        "ige_problem_energy": (c_int32, [vp, POINTER(ctypes.c_uint8), POINTER(c_double)]),
        "ige_server_solution": (c_int32, [vp, c_uint64, POINTER(ctypes.c_uint8),
                                          POINTER(c_int64)]),
        "ige_read_best": (c_int32, [vp, POINTER(ctypes.c_uint8), POINTER(c_int64),
                                    POINTER(c_uint32)]),
        "ige_model_polish": (c_int32, [vp, c_uint32, POINTER(ctypes.c_uint8),
                                       POINTER(ctypes.c_uint8), POINTER(c_int64),
                                       POINTER(c_uint32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    """Message of this thread's last failed call."""
    return load().ige_last_error().decode(errors="replace")


def check(code: int) -> None:
    if code != IGE_OK:
        msg = load().ige_last_error().decode(errors="replace")
        raise IgeError(code, msg)


class Context:
    def __init__(self, device: int = 0):
        lib = load()
        self._lib = lib
        self.handle = c_void_p()
        check(lib.ige_context_create(device, ctypes.byref(self.handle)))

    def close(self):
        if self.handle:
            self._lib.ige_context_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model:
    """Integer QUBO model on the device (ProblemIR::from_pairs semantics)."""

    def __init__(self, ctx: Context, n: int, offset: int, linear, pi, pj, q):
        self._lib = ctx._lib
        self.ctx = ctx
        linear = np.ascontiguousarray(linear, dtype=np.int32)
        pi = np.ascontiguousarray(pi, dtype=np.uint32)
        pj = np.ascontiguousarray(pj, dtype=np.uint32)
        q = np.ascontiguousarray(q, dtype=np.int32)
        if linear.shape != (n,):
            raise ValueError("linear must have length n")
        if not (pi.shape == pj.shape == q.shape):
            raise ValueError("pair arrays must have equal length")
        self.handle = c_void_p()
        check(self._lib.ige_model_from_pairs(
            ctx.handle, n, int(offset), _ptr(linear, c_int32), len(q), _ptr(pi, c_uint32),
            _ptr(pj, c_uint32), _ptr(q, c_int32), ctypes.byref(self.handle)))
        self.info = ModelInfo()
        check(self._lib.ige_model_get_info(self.handle, ctypes.byref(self.info)))
        self.n = n
        self.offset = int(offset)

    def coloring(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.uint32)
        check(self._lib.ige_model_coloring(self.handle, _ptr(out, c_uint32)))
        return out

    def order(self):
        """(order, class_start): the sweep's colour-major position order."""
        order = np.empty(self.n, dtype=np.uint32)
        cs = np.empty(self.info.num_colors + 1, dtype=np.uint32)
        check(self._lib.ige_model_order(self.handle, _ptr(order, c_uint32), _ptr(cs, c_uint32)))
        return order, cs

    def csr(self):
        rp = np.empty(self.n + 1, dtype=np.uint32)
        ci = np.empty(self.info.nnz, dtype=np.uint32)
        w = np.empty(self.info.nnz, dtype=np.int32)
        check(self._lib.ige_model_csr(self.handle, _ptr(rp, c_uint32), _ptr(ci, c_uint32),
                                      _ptr(w, c_int32)))
        return rp, ci, w

    def polish(self, x, clamped=None):
        """steepest_descent_1opt on every row of x ([count][n] bytes, site order), in place
        on a copy.  Returns (x, energies, flips)."""
        x = np.array(x, dtype=np.uint8, order="C", copy=True)
        if x.ndim == 1:
            x = x[None, :]
        if x.shape[1] != self.n:
            raise ValueError("x must be [count][n]")
        cl = None
        if clamped is not None:
            cl = np.ascontiguousarray(clamped, dtype=np.uint8)
            if cl.shape != (self.n,):
                raise ValueError("clamped must have length n")
        e = np.empty(x.shape[0], dtype=np.int64)
        f = np.empty(x.shape[0], dtype=np.uint32)
        check(self._lib.ige_model_polish(self.handle, x.shape[0], _ptr(x, ctypes.c_uint8),
                                         _ptr(cl, ctypes.c_uint8), _ptr(e, c_int64),
                                         _ptr(f, c_uint32)))
        return x, e, f

    def close(self):
        if self.handle:
            self._lib.ige_model_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Problem:
    """A parsed instance file in minimisation form (host only; no device needed)."""

    def __init__(self, text, fmt: int, index: int = 0):
        self._lib = load()
        data = text.encode() if isinstance(text, str) else bytes(text)
        self.handle = c_void_p()
        check(self._lib.ige_problem_parse(data, len(data), fmt, index, ctypes.byref(self.handle)))
        self.info = ProblemInfo()
        check(self._lib.ige_problem_get_info(self.handle, ctypes.byref(self.info)))
        self.n = self.info.n

    def pairs(self):
        """(offset, linear int32[n], pi, pj, q int32[npairs]); raises IGE_ERR_RANGE when the
        instance is not integral."""
        off = c_int64()
        lin = np.empty(self.n, dtype=np.int32)
        m = self.info.npairs
        pi = np.empty(max(m, 1), dtype=np.uint32)
        pj = np.empty(max(m, 1), dtype=np.uint32)
        q = np.empty(max(m, 1), dtype=np.int32)
        check(self._lib.ige_problem_pairs(self.handle, ctypes.byref(off), _ptr(lin, c_int32),
                                          _ptr(pi, c_uint32), _ptr(pj, c_uint32), _ptr(q, c_int32)))
        return off.value, lin, pi[:m], pj[:m], q[:m]

    def energy(self, x) -> float:
        x = np.ascontiguousarray(x, dtype=np.uint8)
        if x.shape != (self.n,):
            raise ValueError("x must have length n")
        e = c_double()
        check(self._lib.ige_problem_energy(self.handle, _ptr(x, ctypes.c_uint8), ctypes.byref(e)))
        return e.value

    def model(self, ctx: "Context") -> "Model":
        off, lin, pi, pj, q = self.pairs()
        return Model(ctx, self.n, off, lin, pi, pj, q)

    def close(self):
        if self.handle:
            self._lib.ige_problem_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class State:
    """`cells` slots x 64 replicas of a model, resident on the device."""

    def __init__(self, model: Model, cells: int, seed: int, init_words=None,
                 kernel: int = KERNEL_AUTO):
        self._lib = model._lib
        self.model = model
        self.cells = int(cells)
        self.n = model.n
        self.handle = c_void_p()
        iw = None
        if init_words is not None:
            iw = np.ascontiguousarray(init_words, dtype=np.uint64).reshape(self.cells * self.n)
        check(self._lib.ige_state_create(model.ctx.handle, model.handle, self.cells,
                                         int(seed) & 0xFFFFFFFFFFFFFFFF,
                                         _ptr(iw, c_uint64), kernel, ctypes.byref(self.handle)))

    def set_ladder(self, beta) -> None:
        b = np.ascontiguousarray(beta, dtype=np.float64)
        if b.shape != (self.cells,):
            raise ValueError("ladder needs one beta per cell")
        check(self._lib.ige_set_ladder(self.handle, _ptr(b, c_double)))
        self.replica_view = False

    def set_replica_temperatures(self, temps) -> None:
        """Replica rho = slot * 64 + lane runs at T = temps[rho % len(temps)]."""
        t = np.ascontiguousarray(temps, dtype=np.float64).reshape(-1)
        check(self._lib.ige_set_replica_temperatures(self.handle, _ptr(t, c_double), t.size))
        self.replica_view = True

    def exchange_replicas(self, rounds: int) -> None:
        """ReplicaExchange::apply with a budget of `rounds`."""
        check(self._lib.ige_exchange_replicas(self.handle, int(rounds)))

    def run(self, sweeps: int, exchange: bool = False, beta_schedule=None,
            track_best: bool = False, heatbath: bool = False, icm: bool = False,
            eager: bool = False) -> None:
        """`sweeps` sweep steps (ige_run).  Runs of 32 or more replay the state's
        captured sweep-loop graph unless `eager`; both give the same bits."""
        sched = None
        if beta_schedule is not None:
            sched = np.ascontiguousarray(beta_schedule, dtype=np.float64)
            if sched.shape != (sweeps, self.cells):
                raise ValueError("beta_schedule must be [sweeps][cells]")
        self._sched_keepalive = sched
        flags = ((RUN_EXCHANGE if exchange else 0) | (RUN_TRACK_BEST if track_best else 0) |
                 (RUN_HEATBATH if heatbath else 0) | (RUN_ICM if icm else 0) |
                 (RUN_EAGER if eager else 0))
        check(self._lib.ige_run(self.handle, int(sweeps), flags, _ptr(sched, c_double)))

    def steepest_descent(self, max_steps: int = STEPS_UNLIMITED, clamped=None) -> DescentReport:
        cl = None if clamped is None else np.ascontiguousarray(clamped, dtype=np.uint8)
        if cl is not None and cl.shape != (self.n,):
            raise ValueError("clamped must have length n")
        rep = DescentReport()
        check(self._lib.ige_steepest_descent(self.handle, int(max_steps),
                                             _ptr(cl, ctypes.c_uint8), ctypes.byref(rep)))
        return rep

    def icm(self, rounds: int = 1) -> int:
        """ICM sweeps; returns the total size of the clusters flipped."""
        f = c_uint64()
        check(self._lib.ige_icm(self.handle, int(rounds), ctypes.byref(f)))
        return f.value

    def set_populations(self, pops: int) -> None:
        check(self._lib.ige_set_populations(self.handle, int(pops)))
        self.pops = int(pops)

    def tune_ladder(self, temp_max: float, temp_min: float, iters: int, steps: int,
                    alphas=None) -> np.ndarray:
        """Acceptance-uniform ladder tuning on this (throwaway) state; the tuned alphas."""
        T = self.cells // getattr(self, "pops", 1)
        a = (np.array([t / max(T - 1, 1) for t in range(T)]) if alphas is None
             else np.array(alphas, dtype=np.float64))
        check(self._lib.ige_pa_tune_ladder(self.handle, float(temp_max), float(temp_min),
                                           int(iters), int(steps), _ptr(a, c_double)))
        return a

    def resample(self, parents) -> None:
        par = np.ascontiguousarray(parents, dtype=np.uint32)
        check(self._lib.ige_resample(self.handle, _ptr(par, c_uint32)))

    def update_best(self) -> None:
        check(self._lib.ige_update_best(self.handle))

    def best(self):
        """(x bytes in site order, energy, sweeps completed when found)."""
        x = np.empty(self.n, dtype=np.uint8)
        e = c_int64()
        sw = c_uint32()
        check(self._lib.ige_read_best(self.handle, _ptr(x, ctypes.c_uint8), ctypes.byref(e),
                                      ctypes.byref(sw)))
        return x, e.value, sw.value

    def exchange(self, rounds: int) -> None:
        check(self._lib.ige_exchange(self.handle, int(rounds)))

    def synchronize(self) -> None:
        check(self._lib.ige_synchronize(self.handle))

    def spins(self) -> np.ndarray:
        out = np.empty((self.cells, self.n), dtype=np.uint64)
        check(self._lib.ige_read_spins(self.handle, _ptr(out, c_uint64)))
        return out

    def write_spins(self, words) -> None:
        """Replace the words ([cells][n], site order); tracked energies are recomputed."""
        w = np.ascontiguousarray(words, dtype=np.uint64).reshape(self.cells, self.n)
        check(self._lib.ige_write_spins(self.handle, _ptr(w, c_uint64)))

    def energies(self) -> np.ndarray:
        out = np.empty((self.cells, 64), dtype=np.int64)
        check(self._lib.ige_read_energies(self.handle, _ptr(out, c_int64)))
        return out

    def temp_of(self) -> np.ndarray:
        out = np.empty((self.cells, 64), dtype=np.uint32)
        check(self._lib.ige_read_temp_of(self.handle, _ptr(out, c_uint32)))
        return out

    def fields(self) -> np.ndarray:
        out = np.empty((self.cells, self.n, 64), dtype=np.int32)
        check(self._lib.ige_read_fields(self.handle, _ptr(out, c_int32)))
        return out

    def recompute_energies(self) -> np.ndarray:
        out = np.empty((self.cells, 64), dtype=np.int64)
        check(self._lib.ige_recompute_energies(self.handle, _ptr(out, c_int64)))
        return out

    def stats(self):
        st = Stats()
        pairs = (self.cells * 64 if getattr(self, "replica_view", False) else self.cells) - 1
        pa = np.zeros(max(pairs, 1), dtype=np.uint64)
        pc = np.zeros(max(pairs, 1), dtype=np.uint64)
        check(self._lib.ige_read_stats(self.handle, ctypes.byref(st), _ptr(pa, c_uint64),
                                       _ptr(pc, c_uint64)))
        return st, pa[:pairs], pc[:pairs]

    def profile(self, on: bool) -> None:
        check(self._lib.ige_profile_enable(self.handle, 1 if on else 0))

    def profile_read(self):
        ms = c_double()
        n = c_uint64()
        ms2 = c_double()
        n2 = c_uint64()
        check(self._lib.ige_profile_read(self.handle, ctypes.byref(ms), ctypes.byref(n),
                                         ctypes.byref(ms2), ctypes.byref(n2)))
        return ms.value, n.value, ms2.value, n2.value

    def close(self):
        if self.handle:
            self._lib.ige_state_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pa_normalize(log_w):
    """(normalized weights, effective sample size) of accumulated log-weights."""
    lw = np.ascontiguousarray(log_w, dtype=np.float64)
    w = np.empty_like(lw)
    ess = c_double()
    check(load().ige_pa_normalize(_ptr(lw, c_double), lw.size, _ptr(w, c_double),
                                  ctypes.byref(ess)))
    return w, ess.value


def pa_free_energy_increment(log_w) -> float:
    lw = np.ascontiguousarray(log_w, dtype=np.float64)
    out = c_double()
    check(load().ige_pa_free_energy_increment(_ptr(lw, c_double), lw.size, ctypes.byref(out)))
    return out.value


def pa_systematic_resample(w_norm, u: float) -> np.ndarray:
    w = np.ascontiguousarray(w_norm, dtype=np.float64)
    counts = np.empty(w.size, dtype=np.uint32)
    check(load().ige_pa_systematic_resample(_ptr(w, c_double), w.size, float(u),
                                            _ptr(counts, c_uint32)))
    return counts


def pa_parents(counts, seed: int, event: int) -> np.ndarray:
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    out = np.empty(c.size, dtype=np.uint32)
    check(load().ige_pa_parents(_ptr(c, c_uint32), c.size, int(seed) & 0xFFFFFFFFFFFFFFFF,
                                int(event), _ptr(out, c_uint32)))
    return out


def pa_family_entropy(family):
    f = np.ascontiguousarray(family, dtype=np.uint32)
    rho = c_double()
    k = c_uint32()
    check(load().ige_pa_family_entropy(_ptr(f, c_uint32), f.size, ctypes.byref(rho),
                                       ctypes.byref(k)))
    return rho.value, k.value


def pa_adapt_spacings(spacings, acceptance) -> np.ndarray:
    d = np.ascontiguousarray(spacings, dtype=np.float64)
    a = np.ascontiguousarray(acceptance, dtype=np.float64)
    out = np.empty_like(d)
    check(load().ige_pa_adapt_spacings(_ptr(d, c_double), _ptr(a, c_double), d.size,
                                       _ptr(out, c_double)))
    return out


def ultimate_ladder(temp_max: float, temp_min: float, num_temps: int) -> np.ndarray:
    """UltimateSolver's betas (hottest first) for a PT solve request."""
    out = np.empty(num_temps, dtype=np.float64)
    check(load().ige_ultimate_ladder(float(temp_max), float(temp_min), int(num_temps),
                                     _ptr(out, c_double)))
    return out


def anneal_schedule(beta_lo: float, beta_hi: float, sweeps: int) -> np.ndarray:
    """The deterministic geometric schedule batch instances follow."""
    out = np.empty(sweeps, dtype=np.float64)
    check(load().ige_anneal_schedule(float(beta_lo), float(beta_hi), int(sweeps),
                                     _ptr(out, c_double)))
    return out


def _fill_desc(d, qb, seed, sweeps, beta_lo, beta_hi, keep):
    lin = np.ascontiguousarray(qb.linear, dtype=np.int32)
    pi = np.ascontiguousarray(qb.pi, dtype=np.uint32)
    pj = np.ascontiguousarray(qb.pj, dtype=np.uint32)
    q = np.ascontiguousarray(qb.q, dtype=np.int32)
    keep += [lin, pi, pj, q]
    d.n = qb.n
    d.offset = int(qb.offset)
    d.linear = _ptr(lin, c_int32)
    d.npairs = len(q)
    d.pi = _ptr(pi, c_uint32)
    d.pj = _ptr(pj, c_uint32)
    d.q = _ptr(q, c_int32)
    d.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    d.sweeps = int(sweeps)
    d.beta_lo = float(beta_lo)
    d.beta_hi = float(beta_hi)


class Server:
    """Continuous-batching request loop (a background thread in the library)."""

    def __init__(self, ctx: Context, hbm_bytes: int = 8 << 30, round_ms: float = 10.0,
                 waves: int = 0):
        self._lib = ctx._lib
        self.ctx = ctx
        self.handle = c_void_p()
        check(self._lib.ige_server_create(ctx.handle, int(hbm_bytes), float(round_ms), int(waves),
                                          ctypes.byref(self.handle)))
        self.sizes = {}

    def configure(self, max_gang=16, gang_sites=2048, spin_limit=1 << 24):
        check(self._lib.ige_server_configure(self.handle, max_gang, gang_sites, spin_limit))

    def submit(self, qb, seed, sweeps, beta_lo=0.0, beta_hi=0.0, track_every: int = 0,
               polish: bool = False, num_temps: int = 0, temp_max: float = 0.0,
               temp_min: float = 0.0) -> int:
        """An anneal (num_temps 0, beta_lo..beta_hi over `sweeps`) or UltimateSolver's
        loop (num_temps cells on the ladder temp_max..temp_min, `sweeps` steps, the
        incumbent every track_every steps)."""
        d = InstanceDesc()
        keep = []
        _fill_desc(d, qb, seed, sweeps, beta_lo, beta_hi, keep)
        d.track_every = int(track_every)
        d.polish = 1 if polish else 0
        d.num_temps = int(num_temps)
        d.temp_max = float(temp_max)
        d.temp_min = float(temp_min)
        rid = c_uint64()
        check(self._lib.ige_server_submit(self.handle, ctypes.byref(d), ctypes.byref(rid)))
        self.sizes[rid.value] = (qb.n, max(1, int(num_temps)))
        return rid.value

    def poll(self, rid: int, release: bool = False, want_words: bool = True):
        """(info, words [cells][n], energies [cells][64]) of a request."""
        info = RequestInfo()
        n, cells = self.sizes[rid]
        words = np.empty((cells, n), dtype=np.uint64) if want_words else None
        e = np.empty((cells, 64), dtype=np.int64) if want_words else None
        check(self._lib.ige_server_poll(self.handle, rid, 1 if release else 0, ctypes.byref(info),
                                        _ptr(words, c_uint64), _ptr(e, c_int64)))
        if release and info.status == REQ_DONE:
            self.sizes.pop(rid, None)
        return info, words, e

    def solution(self, rid: int):
        """(x bytes in site order, energy) of a finished request."""
        x = np.empty(self.sizes[rid][0], dtype=np.uint8)
        e = c_int64()
        check(self._lib.ige_server_solution(self.handle, rid, _ptr(x, ctypes.c_uint8),
                                            ctypes.byref(e)))
        return x, e.value

    def wait(self, rid: int, timeout_ms: float = 60000.0):
        info = RequestInfo()
        check(self._lib.ige_server_wait(self.handle, rid, float(timeout_ms), ctypes.byref(info)))
        return info

    def stats(self) -> ServerCounters:
        st = ServerCounters()
        check(self._lib.ige_server_stats(self.handle, ctypes.byref(st)))
        return st

    def close(self):
        if self.handle:
            self._lib.ige_server_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent instances (one cell of 64 replicas each) in one arena."""

    def __init__(self, ctx: Context, qubos, seeds, sweeps, beta_lo, beta_hi):
        self._lib = ctx._lib
        self.ctx = ctx
        self.count = len(qubos)
        self.ns = [q.n for q in qubos]
        descs = (InstanceDesc * self.count)()
        keep = []
        for i, qb in enumerate(qubos):
            _fill_desc(descs[i], qb, seeds[i], sweeps[i], beta_lo[i], beta_hi[i], keep)
        self.handle = c_void_p()
        check(self._lib.ige_batch_create(ctx.handle, self.count, descs, ctypes.byref(self.handle)))
        ts = c_uint64()
        ta = c_uint64()
        check(self._lib.ige_batch_info(self.handle, ctypes.byref(ts), ctypes.byref(ta)))
        self.total_sites = ts.value
        self.total_attempts = ta.value

    def reset(self):
        check(self._lib.ige_batch_reset(self.handle))

    def configure(self, max_gang=16, gang_sites=2048, spin_limit=1 << 24):
        check(self._lib.ige_batch_configure(self.handle, max_gang, gang_sites, spin_limit))

    def run(self, mode=BATCH_PERSISTENT, waves=0, k0=0, k1=0xFFFFFFFF):
        check(self._lib.ige_batch_run(self.handle, mode, waves, k0, k1))

    def autotune(self, probe_sweeps=20):
        w = c_uint32()
        ms = (c_double * 5)()
        check(self._lib.ige_batch_autotune(self.handle, probe_sweeps, ctypes.byref(w), ms))
        return w.value, list(ms)

    def synchronize(self):
        check(self._lib.ige_batch_synchronize(self.handle))

    def read(self, i):
        words = np.empty(self.ns[i], dtype=np.uint64)
        e = np.empty(64, dtype=np.int64)
        acc = c_uint64()
        check(self._lib.ige_batch_read(self.handle, i, _ptr(words, c_uint64), _ptr(e, c_int64),
                                       ctypes.byref(acc)))
        return words, e, acc.value

    def energies(self):
        """(tracked, recomputed): [count][64] tracked energies and a fresh device
        recompute from the words."""
        t = np.empty((self.count, 64), dtype=np.int64)
        r = np.empty((self.count, 64), dtype=np.int64)
        check(self._lib.ige_batch_energies(self.handle, _ptr(t, c_int64), _ptr(r, c_int64)))
        return t, r

    def timing_ms(self):
        """Per-instance device time (ms) of the last persistent run that advanced it."""
        ms = np.empty(self.count, dtype=np.float64)
        check(self._lib.ige_batch_timing(self.handle, None, _ptr(ms, c_double)))
        return ms

    def launch_ms(self):
        """Duration (ms) of the last persistent run, from HIP events on the batch's stream."""
        v = c_double()
        check(self._lib.ige_batch_timing(self.handle, ctypes.byref(v), None))
        return v.value

    def close(self):
        if self.handle:
            self._lib.ige_batch_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
