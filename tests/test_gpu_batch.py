"""Batched instances: one persistent launch over a packed arena.

Every instance must equal (bit-exact) the oracle run of that instance alone
with the same seed and schedule, whichever launch mode or workgroup size ran it.
"""
import math

import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def oracle_schedule(beta_lo, beta_hi, sweeps):
    # beta_k = beta_hi * exp(-(S-1-k) c), c = (ln beta_hi - ln beta_lo)/(S-1), exp = det_exp
    if sweeps == 1:
        return np.array([beta_hi])
    c = (math.log(beta_hi) - math.log(beta_lo)) / (sweeps - 1)
    return np.array([beta_hi * ob.det_exp(-(sweeps - 1 - k) * c) for k in range(sweeps)])


def small_batch(count=14, sweeps=24, seed=3):
    rng = np.random.default_rng(seed)
    qubos = []
    for i in range(count):
        n = int(rng.choice([1, 5, 37, 64, 100, 257, 400, 613]))
        if i % 5 == 4:
            qubos.append(instances.hypercubic_pm_j(4, 2, 100 + i))  # a +-J model in the mix
        elif i % 7 == 3:
            # |dE| range beyond the LDS threshold table: thresholds from the band test
            qubos.append(instances.random_qubo(n, 8, 1000 + i, offdiag=2000, diag=3000))
        else:
            qubos.append(instances.random_qubo(n, 6, 1000 + i))
    seeds = rng.integers(1, 1 << 62, size=count)
    b_lo, b_hi = [], []
    for qb in qubos:
        tmax, tmin = instances.suggested_temp_range(qb)
        b_lo.append(1.0 / tmax)
        b_hi.append(1.0 / tmin)
    sw = [sweeps - (i % 3) for i in range(count)]  # uneven lengths too
    return qubos, seeds, sw, b_lo, b_hi


def test_schedule_matches_oracle_restatement():
    s = _lib.anneal_schedule(0.01, 3.0, 37)
    np.testing.assert_array_equal(s, oracle_schedule(0.01, 3.0, 37))
    assert s[-1] == 3.0 and s[0] == pytest.approx(0.01, rel=1e-12)


def test_batch_matches_oracle_per_instance(ctx):
    qubos, seeds, sw, b_lo, b_hi = small_batch()
    batch = _lib.Batch(ctx, qubos, seeds, sw, b_lo, b_hi)
    batch.run(_lib.BATCH_PERSISTENT, waves=4)
    for i, qb in enumerate(qubos):
        print(f"[batch] instance {i} n={qb.n}", flush=True)
        csr = ob.CSR.from_qubo(qb)
        ost = ob.OracleState(csr, 1, int(seeds[i]), view=ob.VIEW_WORD)
        sched = oracle_schedule(b_lo[i], b_hi[i], sw[i])
        ost.run(sw[i], exchange=False, schedule=sched[:, None])
        words, e, acc = batch.read(i)
        np.testing.assert_array_equal(words, ost.words[0])
        np.testing.assert_array_equal(e, ost.energy[0])
        assert acc == int(ost.stats[0])
        np.testing.assert_array_equal(csr.energies(words[None, :])[0], e)


def test_modes_and_workgroup_sizes_agree(ctx):
    qubos, seeds, sw, b_lo, b_hi = small_batch(count=20, sweeps=15, seed=5)
    batch = _lib.Batch(ctx, qubos, seeds, sw, b_lo, b_hi)
    results = {}
    for label, mode, waves in [("plain", _lib.BATCH_PLAIN, 0), ("w1", 0, 1), ("w2", 0, 2),
                               ("w4", 0, 4), ("w8", 0, 8)]:
        batch.reset()
        batch.run(mode, waves=waves)
        results[label] = [batch.read(i) for i in range(len(qubos))]
    ref = results["plain"]
    for label, res in results.items():
        for i in range(len(qubos)):
            np.testing.assert_array_equal(res[i][0], ref[i][0], err_msg=f"{label} inst {i}")
            np.testing.assert_array_equal(res[i][1], ref[i][1], err_msg=f"{label} inst {i}")
            assert res[i][2] == ref[i][2]


def test_split_runs_resume_exactly_and_autotune_resets(ctx):
    qubos, seeds, sw, b_lo, b_hi = small_batch(count=9, sweeps=20, seed=7)
    batch = _lib.Batch(ctx, qubos, seeds, sw, b_lo, b_hi)
    batch.run(0, waves=2)
    full = [batch.read(i) for i in range(9)]
    w, ms = batch.autotune(probe_sweeps=3)
    assert w in (1, 2, 4, 8, 16) and len(ms) == 5 and all(m > 0 for m in ms)
    batch.run(0, waves=0, k0=0, k1=7)   # tuned size, in two pieces
    batch.run(0, waves=0, k0=7, k1=100)
    for i in range(9):
        got = batch.read(i)
        np.testing.assert_array_equal(got[0], full[i][0])
        np.testing.assert_array_equal(got[1], full[i][1])


def test_replayed_anneal_graph_reproduces_the_run(ctx):
    """The persistent anneal is captured once per (waves, k0, k1) and replayed: replays after
    a reset, with host reads in between, give the first run's results."""
    qubos, seeds, sw, b_lo, b_hi = small_batch(count=7, sweeps=14, seed=8)
    batch = _lib.Batch(ctx, qubos, seeds, sw, b_lo, b_hi)
    batch.run(0, waves=2)
    first = [batch.read(i) for i in range(7)]
    for _ in range(3):
        batch.reset()
        batch.run(0, waves=2)
        batch.synchronize()
        for i in range(7):
            got = batch.read(i)
            np.testing.assert_array_equal(got[0], first[i][0])
            np.testing.assert_array_equal(got[1], first[i][1])
            assert got[2] == first[i][2]


def test_single_state_with_the_schedule_reproduces_an_instance(ctx):
    qubos, seeds, sw, b_lo, b_hi = small_batch(count=4, sweeps=18, seed=9)
    batch = _lib.Batch(ctx, qubos, seeds, sw, b_lo, b_hi)
    batch.run(0, waves=1)
    for i in range(4):
        qb = qubos[i]
        m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
        st = _lib.State(m, 1, int(seeds[i]), kernel=_lib.KERNEL_GENERIC)
        sched = _lib.anneal_schedule(b_lo[i], b_hi[i], sw[i])
        st.run(sw[i], beta_schedule=sched[:, None])
        words, e, _ = batch.read(i)
        np.testing.assert_array_equal(st.spins()[0], words)
        np.testing.assert_array_equal(st.energies()[0], e)


def test_gangs_give_identical_results(ctx):
    """Large instances split over a gang of workgroups meeting at device-wide barriers
    produce the same words and energies as one workgroup per instance."""
    qubos, seeds, sw, b_lo, b_hi = small_batch(count=14, sweeps=16, seed=11)
    batch = _lib.Batch(ctx, qubos, seeds, sw, b_lo, b_hi)
    batch.configure(max_gang=1)
    batch.run(0, waves=4)
    ref = [batch.read(i) for i in range(14)]
    for max_gang, sites, waves in [(4, 64, 4), (8, 32, 2), (3, 100, 8)]:
        batch.configure(max_gang=max_gang, gang_sites=sites)
        batch.reset()
        batch.run(0, waves=waves)
        batch.synchronize()
        for i in range(14):
            got = batch.read(i)
            np.testing.assert_array_equal(got[0], ref[i][0], err_msg=f"gang {max_gang} inst {i}")
            np.testing.assert_array_equal(got[1], ref[i][1], err_msg=f"gang {max_gang} inst {i}")
            assert got[2] == ref[i][2]


def test_expired_gang_wait_is_reported_not_hung(ctx):
    """A zero spin limit makes a waiting workgroup give up at once: the launch still ends
    and the failure surfaces as an error; the batch works again once reconfigured."""
    qubos, seeds, sw, b_lo, b_hi = small_batch(count=6, sweeps=12, seed=13)
    big = [instances.random_qubo(600, 6, 77 + i) for i in range(2)]
    qubos = qubos + big
    seeds = list(seeds) + [5, 6]
    sw = sw + [12, 12]
    for qb in big:
        tmax, tmin = instances.suggested_temp_range(qb)
        b_lo.append(1.0 / tmax)
        b_hi.append(1.0 / tmin)
    batch = _lib.Batch(ctx, qubos, seeds, sw, b_lo, b_hi)
    batch.configure(max_gang=8, gang_sites=50, spin_limit=0)
    batch.run(0, waves=1)
    with pytest.raises(_lib.IgeError) as ei:
        batch.synchronize()
    assert ei.value.code == _lib.IGE_ERR_HIP
    batch.configure(max_gang=1)
    batch.reset()
    batch.run(0, waves=1)
    batch.synchronize()
    words, e, _ = batch.read(7)
    csr = ob.CSR.from_qubo(big[1])
    np.testing.assert_array_equal(csr.energies(words[None, :])[0], e)
