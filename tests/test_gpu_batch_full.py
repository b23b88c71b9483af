"""The 4096-instance sparse QUBO batch at its full size (N log-uniform 64..8192, mean degree
8, Q_ij in [-100, 100], Q_ii in [-200, 200], 2000-sweep anneals), run the way the benchmark
runs it: the autotuned workgroup size, one persistent launch over the packed arena.

- Over a 50-sweep prefix of the schedule the smallest instance, the largest and eight
  random ones equal the oracle run of that instance alone, bit for bit (words, int64
  energies, accepted counts).
- After the whole anneal every instance's tracked energies equal a fresh recompute from
  its words (on the device for all 4096; on the CPU oracle for a sample), and the anneal
  split at sweep 50 equals the same anneal in one launch, for every instance.
"""
import math

import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu

PREFIX = 50


def oracle_schedule(beta_lo, beta_hi, sweeps):
    c = (math.log(beta_hi) - math.log(beta_lo)) / (sweeps - 1)
    return np.array([beta_hi * ob.det_exp(-(sweeps - 1 - k) * c) for k in range(sweeps)])


@pytest.fixture(scope="module")
def full(ctx):
    print("[batch-full] building the 4096-instance workload", flush=True)
    wl = instances.workload_qubo_batch()
    b = _lib.Batch(ctx, wl["qubos"], wl["seeds"], wl["sweeps"], wl["beta_lo"], wl["beta_hi"])
    waves, _ = b.autotune(probe_sweeps=5)
    print(f"[batch-full] {b.count} instances, {b.total_sites} sites, tuned waves {waves}",
          flush=True)
    yield wl, b
    b.close()


def sample_ids(wl, k=8, seed=17):
    ns = np.array([q.n for q in wl["qubos"]])
    rng = np.random.default_rng(seed)
    ids = [int(ns.argmin()), int(ns.argmax())]
    ids += [int(i) for i in rng.choice(len(ns), size=k, replace=False) if i not in ids]
    return ids


def test_full_batch_prefix_matches_oracle(full):
    wl, b = full
    b.reset()
    b.run(_lib.BATCH_PERSISTENT, waves=0, k0=0, k1=PREFIX)
    b.synchronize()
    tracked, recomputed = b.energies()
    np.testing.assert_array_equal(tracked, recomputed)
    for i in sample_ids(wl):
        qb = wl["qubos"][i]
        print(f"[batch-full] prefix instance {i} n={qb.n}", flush=True)
        csr = ob.CSR.from_qubo(qb)
        sched = oracle_schedule(wl["beta_lo"][i], wl["beta_hi"][i], wl["sweeps"][i])
        ost = ob.OracleState(csr, 1, int(wl["seeds"][i]), view=ob.VIEW_WORD)
        ost.run(PREFIX, exchange=False, schedule=sched[:PREFIX, None])
        words, e, acc = b.read(i)
        np.testing.assert_array_equal(words, ost.words[0], err_msg=f"instance {i}")
        np.testing.assert_array_equal(e, ost.energy[0], err_msg=f"instance {i}")
        assert acc == int(ost.stats[0])


def test_full_batch_anneal_energies_exact_and_split_equals_one_launch(full):
    wl, b = full
    # split: sweeps [0, 50) then [50, 2000)
    b.reset()
    b.run(_lib.BATCH_PERSISTENT, waves=0, k0=0, k1=PREFIX)
    b.run(_lib.BATCH_PERSISTENT, waves=0, k0=PREFIX, k1=wl["sweeps"][0])
    b.synchronize()
    tracked, recomputed = b.energies()
    np.testing.assert_array_equal(tracked, recomputed)
    split = [b.read(i) for i in range(b.count)]
    for i in sample_ids(wl, k=16, seed=23):
        csr = ob.CSR.from_qubo(wl["qubos"][i])
        np.testing.assert_array_equal(csr.energies(split[i][0][None, :])[0], split[i][1],
                                      err_msg=f"instance {i}")
    print("[batch-full] split anneal checked; one-launch anneal next", flush=True)
    # one launch: what the benchmark times
    b.reset()
    b.run(_lib.BATCH_PERSISTENT, waves=0)
    b.synchronize()
    tracked1, recomputed1 = b.energies()
    np.testing.assert_array_equal(tracked1, recomputed1)
    np.testing.assert_array_equal(tracked1, tracked)
    for i in range(b.count):
        words, e, acc = b.read(i)
        np.testing.assert_array_equal(words, split[i][0], err_msg=f"instance {i}")
        assert acc == split[i][2], i
    ms = b.timing_ms()
    assert np.all(ms > 0)
    ns = np.array([q.n for q in wl["qubos"]])
    print(f"[batch-full] per-instance device time: largest instance {ms[ns.argmax()]:.1f} ms, "
          f"max {ms.max():.1f} ms, median {np.median(ms):.2f} ms", flush=True)
