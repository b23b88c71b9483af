"""Pins the CPU baseline (the port of src/solver/engine.rs::step timed by bench.py).

The baseline is reported next to the GPU figure, never used to check it; these
tests make sure the thing timed is a correct Metropolis + DEO sweep:
fast_exp within the error bound its construction gives, tracked energies equal
a recompute, and the sampled distribution equals exact Boltzmann weights.
"""
import ctypes
import itertools
import math
import os
from ctypes import POINTER, c_double, c_int, c_int8, c_uint32, c_uint64

import numpy as np
import pytest

from oracle import binding as ob
from ising_engine_amd import instances

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cb():
    path = os.path.join(ROOT, "oracle", "build", "libcpu_baseline.so")
    if not os.path.exists(path):
        pytest.skip("oracle not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    lib.cb_run.argtypes = [c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double),
                           POINTER(c_double), c_uint32, POINTER(c_double), c_uint32, c_uint64,
                           c_int, POINTER(c_int8), POINTER(c_double), POINTER(c_int)]
    lib.cb_run.restype = c_int
    lib.cb_energies.argtypes = [c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double),
                                POINTER(c_double), c_uint32, POINTER(c_int8), POINTER(c_double)]
    lib.cb_energies.restype = None
    lib.cb_fast_exp.argtypes = [c_double]
    lib.cb_fast_exp.restype = c_double
    return lib


class Baseline:
    def __init__(self, lib, qb, cells, temps, seed=5):
        self.lib = lib
        csr = ob.CSR.from_qubo(qb)
        self.n = qb.n
        self.rp, self.ci = csr.row_ptr, csr.col_idx
        self.w = csr.w.astype(np.float64)
        self.lin = csr.linear.astype(np.float64)
        self.cells = cells
        self.temps = np.ascontiguousarray(temps, dtype=np.float64)
        rng = np.random.default_rng(seed)
        self.spins = rng.integers(0, 2, size=(cells, qb.n, 64), dtype=np.int8)
        self.energies = self.recompute()
        self.phase = c_int(0)
        self.seed = seed

    def _p(self, a, t):
        return a.ctypes.data_as(POINTER(t))

    def recompute(self):
        e = np.zeros((self.cells, 64), dtype=np.float64)
        self.lib.cb_energies(self.n, self._p(self.rp, c_uint32), self._p(self.ci, c_uint32),
                             self._p(self.w, c_double), self._p(self.lin, c_double), self.cells,
                             self._p(self.spins, c_int8), self._p(e, c_double))
        return e

    def run(self, steps, threads=2):
        self.seed += 1
        rc = self.lib.cb_run(self.n, self._p(self.rp, c_uint32), self._p(self.ci, c_uint32),
                             self._p(self.w, c_double), self._p(self.lin, c_double), self.cells,
                             self._p(self.temps, c_double), steps, self.seed, threads,
                             self._p(self.spins, c_int8), self._p(self.energies, c_double),
                             ctypes.byref(self.phase))
        assert rc == 0


def test_fast_exp_error_bound(cb):
    # fast_exp(x) = p(x/16)^16 with p the degree-6 Taylor polynomial of exp.
    # For r = x/16 in [-1.25, 0] the truncation error is at most |r|^7/7!, so
    # |fast_exp(x)/exp(x) - 1| <= (1 + |r|^7/(7! e^r))^16 - 1 (plus rounding).
    xs = np.linspace(-20.0, 0.0, 20001)
    for x in xs:
        r = x / 16.0
        bound = (1.0 + abs(r) ** 7 / (math.factorial(7) * math.exp(r))) ** 16 - 1.0
        got = cb.cb_fast_exp(float(x))
        assert abs(got / math.exp(x) - 1.0) <= bound * 1.001 + 1e-14, x
    # the clamp: arguments below -20 give exp(-20)'s approximation, above 0 give 1
    assert cb.cb_fast_exp(-100.0) == cb.cb_fast_exp(-20.0)
    assert cb.cb_fast_exp(3.0) == 1.0


def test_tracked_energies_equal_recompute_with_exchange(cb):
    qb = instances.hypercubic_pm_j(8, 2, seed=3)
    temps = 1.0 / instances.geometric_betas(3.0, 0.5, 6)
    b = Baseline(cb, qb, 6, temps)
    for _ in range(5):
        b.run(4)
        np.testing.assert_array_equal(b.energies, b.recompute())
    assert b.phase.value == 0  # 20 steps: the DEO phase toggled an even number of times


def test_samples_exact_boltzmann(cb):
    n, lin, pi, pj, q = 5, [1, -2, 1, 0, 2], [0, 1, 2, 0, 1], [1, 2, 3, 4, 3], [-2, 3, -1, 2, -2]
    qb = instances.Qubo(n=n, offset=0, linear=np.array(lin, dtype=np.int64),
                        pi=np.array(pi, dtype=np.uint32), pj=np.array(pj, dtype=np.uint32),
                        q=np.array(q, dtype=np.int64))
    beta = 0.6
    states = list(itertools.product([0, 1], repeat=n))
    es = np.array([qb.energy(np.array(s)) for s in states], dtype=np.float64)
    p = np.exp(-beta * (es - es.min()))
    p /= p.sum()
    b = Baseline(cb, qb, 1, [1.0 / beta])
    b.run(20, threads=1)
    counts = np.zeros(len(states))
    for _ in range(300):
        b.run(1, threads=1)
        s = b.spins[0]  # [n][64]
        idx = sum(s[i].astype(np.int64) << (n - 1 - i) for i in range(n))
        counts += np.bincount(idx, minlength=len(states))
    from scipy.stats import chisquare
    assert chisquare(counts, p * counts.sum())[1] > 1e-4
