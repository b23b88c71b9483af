"""ctypes binding of oracle/build/liboracle.so (the CPU checker).

TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg).
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int64, c_uint32, c_uint64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

VIEW_WORD = 0
VIEW_PLANE = 1
VIEW_HEATBATH = 2  # heat-bath rule, word-view uniforms

_lib = None


def _p(a, t):
    return None if a is None else a.ctypes.data_as(POINTER(t))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    u32p, u64p, i64p, dp = POINTER(c_uint32), POINTER(c_uint64), POINTER(c_int64), POINTER(c_double)
    lib.or_philox4x32_10.argtypes = [u32p, u32p, u32p]
    lib.or_philox4x32_10.restype = None
    lib.or_det_exp.argtypes = [c_double]
    lib.or_det_exp.restype = c_double
    lib.or_threshold.argtypes = [c_double, c_int64]
    lib.or_threshold.restype = c_uint32
    lib.or_heatbath_threshold.argtypes = [c_double, c_int64]
    lib.or_heatbath_threshold.restype = c_uint64
    lib.or_exchange_threshold.argtypes = [c_double]
    lib.or_exchange_threshold.restype = c_uint32
    lib.or_ir_from_pairs.argtypes = [c_uint32, c_uint64, u32p, u32p, i64p, u32p, u32p, i64p]
    lib.or_ir_from_pairs.restype = c_int64
    lib.or_ir_energy.argtypes = [c_uint32, u32p, u32p, i64p, i64p, c_int64,
                                 POINTER(ctypes.c_uint8)]
    lib.or_ir_energy.restype = c_int64
    lib.or_greedy_coloring.argtypes = [c_uint32, u32p, u32p, u32p]
    lib.or_greedy_coloring.restype = None
    lib.or_color_order.argtypes = [c_uint32, u32p, u32p, u32p]
    lib.or_color_order.restype = c_uint32
    lib.or_fields.argtypes = [c_uint32, u32p, u32p, i64p, i64p, c_uint32, u64p, i64p]
    lib.or_fields.restype = None
    lib.or_energies.argtypes = [c_uint32, u32p, u32p, i64p, i64p, c_int64, c_uint32, u64p, i64p]
    lib.or_energies.restype = None
    lib.or_init_words.argtypes = [c_uint32, c_uint32, c_uint64, u64p]
    lib.or_init_words.restype = None
    lib.or_run.argtypes = [c_uint32, u32p, u32p, i64p, i64p, c_uint32, c_uint64, c_uint32,
                           c_uint32, c_uint32, c_uint32, c_uint32, dp, c_uint32, c_uint32, u64p,
                           u32p, i64p, u64p]
    lib.or_run.restype = c_int
    lib.or_run_temps.argtypes = [c_uint32, u32p, u32p, i64p, i64p, c_uint32, c_uint32, c_uint64,
                                 c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, dp, c_uint32,
                                 c_uint32, u64p, u32p, i64p, u64p]
    lib.or_run_temps.restype = c_int
    lib.or_exchange_replicas.argtypes = [c_uint32, c_uint64, c_uint32, c_uint32, dp, i64p, u32p,
                                         u64p]
    lib.or_exchange_replicas.restype = None
    lib.or_exchange.argtypes = [c_uint32, c_uint64, c_uint32, dp, i64p, u32p, u64p]
    lib.or_exchange_pops.argtypes = [c_uint32, c_uint32, c_uint64, c_uint32, dp, i64p, u32p,
                                     u64p]
    lib.or_exchange_pops.restype = None
    lib.or_resample.argtypes = [c_uint32, c_uint32, c_uint32, u32p, u64p, u32p, i64p]
    lib.or_resample.restype = None
    lib.or_exchange.restype = None
    lib.or_geometric_ladder.argtypes = [c_uint32, c_double, c_double, dp]
    lib.or_geometric_ladder.restype = None
    lib.or_suggested_range.argtypes = [c_uint32, u32p, i64p, i64p, dp, dp]
    lib.or_suggested_range.restype = None
    u8p, i32p = POINTER(ctypes.c_uint8), POINTER(ctypes.c_int32)
    lib.or_steepest_descent.argtypes = [c_uint32, u32p, u32p, i64p, i64p, u8p, c_uint32,
                                        c_uint32, u64p, i64p, u32p]
    lib.or_steepest_descent.restype = c_uint32
    lib.or_polish_1opt.argtypes = [c_uint32, u32p, u32p, i64p, i64p, u8p, u8p, u64p]
    lib.or_polish_1opt.restype = c_int64
    lib.or_icm_sweep.argtypes = [c_uint32, u32p, u32p, i64p, i64p, c_uint32, c_uint64,
                                 c_uint32, u64p, u32p, i64p, u32p]
    lib.or_icm_sweep.restype = c_uint32
    lib.or_update_best.argtypes = [c_uint32, c_uint32, u64p, u32p, i64p, i64p, i32p, u8p]
    lib.or_update_best.restype = c_int
    lib.or_update_best_view.argtypes = [c_uint32, c_uint32, c_uint32, u64p, u32p, i64p, i64p,
                                        i32p, u8p]
    lib.or_update_best_view.restype = c_int
    _lib = lib
    return lib


def philox(ctr, key):
    lib = load()
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib.or_philox4x32_10(_p(c, c_uint32), _p(k, c_uint32), _p(o, c_uint32))
    return o


def det_exp(x: float) -> float:
    return load().or_det_exp(float(x))


def threshold(beta: float, de: int) -> int:
    return load().or_threshold(float(beta), int(de))


def heatbath_threshold(beta: float, h: int) -> int:
    return load().or_heatbath_threshold(float(beta), int(h))


def exchange_threshold(delta: float) -> int:
    return load().or_exchange_threshold(float(delta))


class CSR:
    """ProblemIR restated: CSR with both triangles (int64 containers)."""

    def __init__(self, n, offset, linear, pi, pj, q):
        lib = load()
        self.n = int(n)
        self.offset = int(offset)
        self.linear = np.ascontiguousarray(linear, dtype=np.int64)
        pi = np.ascontiguousarray(pi, dtype=np.uint32)
        pj = np.ascontiguousarray(pj, dtype=np.uint32)
        q = np.ascontiguousarray(q, dtype=np.int64)
        m = len(q)
        self.row_ptr = np.zeros(self.n + 1, dtype=np.uint32)
        ci = np.zeros(max(2 * m, 1), dtype=np.uint32)
        w = np.zeros(max(2 * m, 1), dtype=np.int64)
        nnz = lib.or_ir_from_pairs(self.n, m, _p(pi, c_uint32), _p(pj, c_uint32), _p(q, c_int64),
                                   _p(self.row_ptr, c_uint32), _p(ci, c_uint32), _p(w, c_int64))
        if nnz < 0:
            raise ValueError("pairs must satisfy i < j < n")
        self.col_idx = ci[:nnz].copy()
        self.w = w[:nnz].copy()
        if nnz == 0:
            self.col_idx = np.zeros(1, dtype=np.uint32)
            self.w = np.zeros(1, dtype=np.int64)
        self.nnz = int(nnz)

    @classmethod
    def from_qubo(cls, qb):
        return cls(qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)

    def _args(self):
        return (self.n, _p(self.row_ptr, c_uint32), _p(self.col_idx, c_uint32),
                _p(self.w, c_int64))

    def energy(self, x) -> int:
        x = np.ascontiguousarray(x, dtype=np.uint8)
        n, rp, ci, w = self._args()
        return load().or_ir_energy(n, rp, ci, w, _p(self.linear, c_int64), self.offset,
                                   x.ctypes.data_as(POINTER(ctypes.c_uint8)))

    def coloring(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.uint32)
        n, rp, ci, _ = self._args()
        load().or_greedy_coloring(n, rp, ci, _p(out, c_uint32))
        return out

    def color_order(self):
        colors = self.coloring()
        order = np.zeros(self.n, dtype=np.uint32)
        cs = np.zeros(self.n + 2, dtype=np.uint32)
        k = load().or_color_order(self.n, _p(colors, c_uint32), _p(order, c_uint32),
                                  _p(cs, c_uint32))
        return order, cs[: k + 1].copy()

    def fields(self, words) -> np.ndarray:
        words = np.ascontiguousarray(words, dtype=np.uint64)
        cells = words.shape[0]
        out = np.zeros((cells, self.n, 64), dtype=np.int64)
        n, rp, ci, w = self._args()
        load().or_fields(n, rp, ci, w, _p(self.linear, c_int64), cells, _p(words, c_uint64),
                         _p(out, c_int64))
        return out

    def energies(self, words) -> np.ndarray:
        words = np.ascontiguousarray(words, dtype=np.uint64)
        cells = words.shape[0]
        out = np.zeros((cells, 64), dtype=np.int64)
        n, rp, ci, w = self._args()
        load().or_energies(n, rp, ci, w, _p(self.linear, c_int64), self.offset, cells,
                           _p(words, c_uint64), _p(out, c_int64))
        return out


    def steepest_descent(self, words, energies, max_steps, clamped=None, flips=None):
        """SteepestDescent::apply on every replica, in place on words / energies.
        Returns (steps that flipped, per-replica flip counts [cells][64])."""
        cells = words.shape[0]
        if flips is None:
            flips = np.zeros((cells, 64), dtype=np.uint32)
        cl = None if clamped is None else np.ascontiguousarray(clamped, dtype=np.uint8)
        n, rp, ci, w = self._args()
        steps = load().or_steepest_descent(n, rp, ci, w, _p(self.linear, c_int64),
                                           _p(cl, ctypes.c_uint8), cells, max_steps,
                                           _p(words, c_uint64), _p(energies, c_int64),
                                           _p(flips, c_uint32))
        return int(steps), flips

    def polish_1opt(self, x, clamped=None):
        """steepest_descent_1opt on one byte state, in place. Returns (improvement, flips)."""
        assert x.dtype == np.uint8 and x.flags.c_contiguous and x.shape == (self.n,)
        cl = None if clamped is None else np.ascontiguousarray(clamped, dtype=np.uint8)
        flips = c_uint64()
        n, rp, ci, w = self._args()
        imp = load().or_polish_1opt(n, rp, ci, w, _p(self.linear, c_int64),
                                    _p(cl, ctypes.c_uint8), _p(x, ctypes.c_uint8),
                                    ctypes.byref(flips))
        return int(imp), int(flips.value)


class Incumbent:
    """update_incumbent restated: best (energy, bytes) seen over tracked energies."""

    def __init__(self, n: int):
        self.n = n
        self.energy = ctypes.c_int64(0)
        self.valid = ctypes.c_int32(0)
        self.x = np.zeros(n, dtype=np.uint8)

    def update(self, words, temp_of, energies, replica_view=False) -> bool:
        words = np.ascontiguousarray(words, dtype=np.uint64)
        temp_of = np.ascontiguousarray(temp_of, dtype=np.uint32)
        energies = np.ascontiguousarray(energies, dtype=np.int64)
        return bool(load().or_update_best_view(self.n, words.shape[0], 1 if replica_view else 0,
                                               _p(words, c_uint64), _p(temp_of, c_uint32),
                                               _p(energies, c_int64), ctypes.byref(self.energy),
                                               ctypes.byref(self.valid),
                                               _p(self.x, ctypes.c_uint8)))


def init_words(n: int, cells: int, seed: int) -> np.ndarray:
    out = np.zeros((cells, n), dtype=np.uint64)
    load().or_init_words(n, cells, seed & 0xFFFFFFFFFFFFFFFF, _p(out, c_uint64))
    return out


class OracleState:
    """Mirror of a device state, advanced by the oracle's or_run."""

    def __init__(self, csr: CSR, cells: int, seed: int, words=None, view: int = VIEW_WORD):
        self.csr = csr
        self.cells = int(cells)
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.view = view
        self.words = (init_words(csr.n, cells, seed) if words is None
                      else np.array(words, dtype=np.uint64).reshape(cells, csr.n))
        self.temp_of = np.repeat(np.arange(self.cells, dtype=np.uint32)[:, None], 64, axis=1).copy()
        self.energy = csr.energies(self.words)
        self.sweep = 0
        self.xround = 0
        self.icm_round = 0
        self.pops = 1
        self.stats = np.zeros(2 + 2 * max(self.cells - 1, 0), dtype=np.uint64)
        self.replica_betas = None  # the replica view's beta per replica index, once set

    def set_replica_temperatures(self, temps):
        """The replica view (metropolis_sweep.rs:115, replica_exchange.rs:78-85): replica
        rho = slot * 64 + lane runs at T = temps[rho % len(temps)].  The first call makes
        temp_of the identity over replica indices and resizes the pair statistics."""
        temps = np.asarray(temps, dtype=np.float64)
        R = self.cells * 64
        if self.replica_betas is None:
            # the configuration at (temperature t, lane) becomes replica t * 64 + lane
            self.temp_of = (self.temp_of.astype(np.uint32) * 64
                            + np.arange(64, dtype=np.uint32)[None, :]).astype(np.uint32)
            self.stats = np.concatenate([self.stats[:2],
                                         np.zeros(2 * (R - 1), dtype=np.uint64)])
        self.replica_betas = np.ascontiguousarray(1.0 / temps[np.arange(R) % temps.size])

    def exchange_replicas(self, rounds: int):
        """ReplicaExchange::apply with budget `rounds`: round k of this call pairs from
        start k & 1; the uniforms are keyed on the state's exchange counter."""
        for k in range(rounds):
            load().or_exchange_replicas(self.cells, self.seed, self.xround, k & 1,
                                        _p(self.replica_betas, c_double),
                                        _p(self.energy, c_int64), _p(self.temp_of, c_uint32),
                                        _p(self.stats, c_uint64))
            self.xround += 1

    def run(self, sweeps: int, exchange: bool, ladder=None, schedule=None):
        if self.replica_betas is not None:
            assert schedule is None and not exchange
            n, rp, ci, w = self.csr._args()
            R = self.cells * 64
            rc = load().or_run_temps(n, rp, ci, w, _p(self.csr.linear, c_int64), self.cells, R,
                                     self.seed, self.view, 0, self.sweep, self.xround, sweeps,
                                     _p(self.replica_betas, c_double), 0, self.cells,
                                     _p(self.words, c_uint64), _p(self.temp_of, c_uint32),
                                     _p(self.energy, c_int64), _p(self.stats, c_uint64))
            if rc != 0:
                raise ValueError("or_run_temps rejected its arguments")
            self.sweep += sweeps
            return
        if schedule is not None:
            betas = np.ascontiguousarray(schedule, dtype=np.float64).reshape(sweeps, self.cells)
            per = 1
        else:
            betas = np.ascontiguousarray(ladder, dtype=np.float64).reshape(self.cells)
            per = 0
        n, rp, ci, w = self.csr._args()
        rc = load().or_run(n, rp, ci, w, _p(self.csr.linear, c_int64), self.cells, self.seed,
                           self.view, 1 if exchange else 0, self.sweep, self.xround, sweeps,
                           _p(betas, c_double), per, self.cells // self.pops,
                           _p(self.words, c_uint64),
                           _p(self.temp_of, c_uint32), _p(self.energy, c_int64),
                           _p(self.stats, c_uint64))
        if rc != 0:
            raise ValueError("or_run rejected its arguments")
        self.sweep += sweeps
        if exchange and self.cells >= 2:
            self.xround += sweeps

    def icm(self, rounds: int = 1):
        """ICM sweeps (icm.rs:163-185), keyed on the state's ICM round counter.
        Returns the per-pair cluster sizes of the last round ([cells][32])."""
        sizes = np.zeros((self.cells, 32), dtype=np.uint32)
        n, rp, ci, w = self.csr._args()
        for _ in range(rounds):
            load().or_icm_sweep(n, rp, ci, w, _p(self.csr.linear, c_int64), self.cells,
                                self.seed, self.icm_round, _p(self.words, c_uint64),
                                _p(self.temp_of, c_uint32), _p(self.energy, c_int64),
                                _p(sizes, c_uint32))
            self.icm_round += 1
        return sizes

    def exchange(self, rounds: int, ladder):
        """Exchange-only rounds (ReplicaExchange::apply) on the tracked energies."""
        if self.cells < 2:
            return
        beta = np.ascontiguousarray(ladder, dtype=np.float64)
        for _ in range(rounds):
            load().or_exchange_pops(self.cells, self.cells // self.pops,
                                    self.seed & 0xFFFFFFFFFFFFFFFF, self.xround,
                                    _p(beta, c_double), _p(self.energy, c_int64),
                                    _p(self.temp_of, c_uint32), _p(self.stats, c_uint64))
            self.xround += 1

    def resample(self, parents):
        """apply_resample: member m = p*64 + r takes member parents[m]'s configurations."""
        par = np.ascontiguousarray(parents, dtype=np.uint32)
        assert par.shape == (self.pops * 64,)
        load().or_resample(self.csr.n, self.cells, self.cells // self.pops, _p(par, c_uint32),
                           _p(self.words, c_uint64), _p(self.temp_of, c_uint32),
                           _p(self.energy, c_int64))


def exchange(cells, seed, round_, beta, energies, temp_of, stats):
    """Exchange number round_ (phase round_ & 1) in place on temp_of / stats."""
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    energies = np.ascontiguousarray(energies, dtype=np.int64)
    load().or_exchange(cells, seed & 0xFFFFFFFFFFFFFFFF, round_, _p(beta, c_double),
                       _p(energies, c_int64), _p(temp_of, c_uint32), _p(stats, c_uint64))


def geometric_ladder(k, t_hi, t_lo):
    out = np.zeros(max(k, 1), dtype=np.float64)
    load().or_geometric_ladder(k, t_hi, t_lo, _p(out, c_double))
    return out


def suggested_range(csr: CSR):
    tmax = c_double()
    tmin = c_double()
    load().or_suggested_range(csr.n, _p(csr.row_ptr, c_uint32), _p(csr.w, c_int64),
                              _p(csr.linear, c_int64), ctypes.byref(tmax), ctypes.byref(tmin))
    return tmax.value, tmin.value
