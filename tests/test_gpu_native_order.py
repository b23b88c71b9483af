"""The device's colour-ordered, counter-keyed parallel tempering against the reference's
native sequential order, statistically: per-temperature mean energy and per-pair swap
acceptance within 99% intervals.

Native order: the port of src/solver/engine.rs::step (oracle/cpu_baseline.c): every cell
swept in ascending site order with per-cell Xoshiro streams and fast_exp acceptance, then a
DEO exchange that swaps spins.  Device: the multispin kernel in colour order with Philox
plane uniforms and temperature-index exchange.  Both run K = 12 independent seeds of
- a 2D +-J L=16 lattice on 8 temperatures geometric in T 0.5 -> 2.0 (the 3D workload's
  range), 2000 burn-in and 4000 measured steps;
- a 3D +-J L=6 lattice on 16 temperatures geometric in T 0.5 -> 2.0, 1000 burn-in and
  2000 measured steps;
each run contributes one mean per quantity.  A quantity passes when |mean_gpu - mean_cpu|
<= t * sqrt(se_gpu^2 + se_cpu^2) with t the two-sided 99% Student quantile (df = K - 1),
Bonferroni-adjusted over the case's 2T - 1 quantities so that each case's check holds at
99% when the two laws agree.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int8, c_uint32, c_uint64

import numpy as np
import pytest
from scipy.stats import t as student_t

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
K, EVERY = 12, 40
CASES = {
    "2d_L16_8T": dict(L=16, dim=2, seed=11, temps=np.geomspace(0.5, 2.0, 8), burn=2000,
                      measure=4000),
    "3d_L6_16T": dict(L=6, dim=3, seed=12, temps=np.geomspace(0.5, 2.0, 16), burn=1000,
                      measure=2000),
}


def cpu_run(lib, csr, seed, TEMPS, BURN, MEASURE):
    cells, n = len(TEMPS), csr.n
    rp, ci = csr.row_ptr, csr.col_idx
    w, lin = csr.w.astype(np.float64), csr.linear.astype(np.float64)
    temps = np.ascontiguousarray(TEMPS)
    rng = np.random.default_rng(seed)
    spins = rng.integers(0, 2, size=(cells, n, 64), dtype=np.int8)
    e = np.zeros((cells, 64))
    P = lambda a, t: a.ctypes.data_as(POINTER(t))
    lib.cb_energies(n, P(rp, c_uint32), P(ci, c_uint32), P(w, c_double), P(lin, c_double), cells,
                    P(spins, c_int8), P(e, c_double))
    phase = c_int(0)
    stats = np.zeros(2 * (cells - 1), dtype=np.uint64)

    def run(steps, s, st):
        assert lib.cb_run_stats(n, P(rp, c_uint32), P(ci, c_uint32), P(w, c_double),
                                P(lin, c_double), cells, P(temps, c_double), steps, s, 4,
                                P(spins, c_int8), P(e, c_double), ctypes.byref(phase),
                                P(st, c_uint64)) == 0

    run(BURN, seed * 7919 + 1, np.zeros_like(stats))
    means = []
    for k in range(MEASURE // EVERY):
        run(EVERY, seed * 7919 + 2 + k, stats)
        # cell = temperature (spins are swapped); the port's energies leave out the offset
        means.append(e.mean(axis=1) + csr.offset)
    acc = stats[1::2] / stats[0::2]
    return np.mean(means, axis=0), acc


def gpu_run(ctx, model, seed, TEMPS, BURN, MEASURE):
    cells = len(TEMPS)
    st = _lib.State(model, cells, seed)
    assert st.stats()[0].kernel == _lib.KERNEL_MULTISPIN
    st.set_ladder(1.0 / TEMPS)
    st.run(BURN, exchange=True)
    _, a0, c0 = st.stats()
    means = []
    for _ in range(MEASURE // EVERY):
        st.run(EVERY, exchange=True)
        e, tof = st.energies(), st.temp_of()
        per_t = np.zeros(cells)
        for s in range(cells):
            np.add.at(per_t, tof[s], e[s])
        means.append(per_t / 64)
    _, a1, c1 = st.stats()
    np.testing.assert_array_equal(st.recompute_energies(), st.energies())
    return np.mean(means, axis=0), (c1 - c0) / (a1 - a0)


@pytest.mark.parametrize("case", list(CASES), ids=list(CASES))
def test_pt_matches_native_sequential_order_statistically(ctx, case):
    cfg = CASES[case]
    TEMPS = cfg["temps"]
    path = os.path.join(ROOT, "oracle", "build", "libcpu_baseline.so")
    lib = ctypes.CDLL(path)
    lib.cb_run_stats.restype = c_int
    qb = instances.hypercubic_pm_j(cfg["L"], cfg["dim"], cfg["seed"])
    csr = ob.CSR.from_qubo(qb)
    model = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    print(f"[native-order] {case}: device runs", flush=True)
    g = [gpu_run(ctx, model, 1000 + k, TEMPS, cfg["burn"], cfg["measure"]) for k in range(K)]
    print(f"[native-order] {case}: native-order port runs", flush=True)
    c = [cpu_run(lib, csr, 2000 + k, TEMPS, cfg["burn"], cfg["measure"]) for k in range(K)]
    ge, ga = np.array([x[0] for x in g]), np.array([x[1] for x in g])
    ce, ca = np.array([x[0] for x in c]), np.array([x[1] for x in c])
    q = 2 * len(TEMPS) - 1
    tq = student_t.ppf(1 - 0.01 / (2 * q), df=K - 1)
    report = []
    for name, a, b in (("energy", ge, ce), ("swap acceptance", ga, ca)):
        ma, mb = a.mean(axis=0), b.mean(axis=0)
        se = np.sqrt(a.var(axis=0, ddof=1) / K + b.var(axis=0, ddof=1) / K)
        for i in range(ma.size):
            report.append((name, i, ma[i], mb[i], tq * se[i]))
    print("\n".join(f"{n:16s} {i}: gpu {x:.4f} cpu {y:.4f} +- {h:.4f}"
                    for n, i, x, y, h in report), flush=True)
    for name, i, x, y, h in report:
        assert abs(x - y) <= h, (name, i, x, y, h)
    # and the quantities are not trivially equal: energies fall with temperature
    assert np.all(np.diff(ge.mean(axis=0)) > 0)
