"""Pins the oracle's replica view, the engine_v2 operators' own semantics:
MetropolisSweep with T = temps[rep % temps.len()] (metropolis_sweep.rs:115) and
ReplicaExchange pairing replica k with k + 1 from start = round & 1, round counted from 0
in every apply (replica_exchange.rs:61-109).

- A replica view whose temperatures are constant over each 64-replica slot is the
  slot-temperature chain already pinned (test_oracle_kat.py): the trajectories agree.
- Exchange is checked against a restatement of the reference's own procedure on
  physically swapped configurations (energies.swap and swap_replicas), its energy-multiset
  invariance (replica_exchange.rs:139-196), exact acceptance rates, and the Boltzmann law
  of the joint chain, which rejects a planted sign mistake in the oracle's exchange.
"""
import itertools
import math

import numpy as np
import pytest
from scipy.stats import chisquare

from oracle import binding as ob
from ising_engine_amd import instances
from test_oracle_kat import BOLTZ_ARGS, frustrated7


def logical(st):
    """(energies, words) of the logical replicas 0..R-1 (replica rho at temp_of == rho)."""
    R = st.cells * 64
    loc = np.empty(R, dtype=np.int64)
    loc[st.temp_of.reshape(-1)] = np.arange(R)
    e = st.energy.reshape(-1)[loc]
    w = st.words  # [cells][n]
    bits = np.array([[(int(w[q // 64, i]) >> int(q % 64)) & 1 for i in range(st.csr.n)]
                     for q in loc], dtype=np.uint8)
    return e, bits


@pytest.mark.parametrize("view", [ob.VIEW_WORD, ob.VIEW_PLANE])
def test_slot_constant_temperatures_reduce_to_the_slot_ladder(view):
    qb = instances.hypercubic_pm_j(6, 2, seed=3)
    csr = ob.CSR.from_qubo(qb)
    ladder = np.array([0.3, 0.9, 2.0])
    a = ob.OracleState(csr, 3, 41, view=view)
    a.run(9, exchange=False, ladder=ladder)
    b = ob.OracleState(csr, 3, 41, view=view)
    b.set_replica_temperatures(np.repeat(1.0 / ladder, 64))   # replica rho -> slot rho // 64
    b.run(9, exchange=False)
    assert np.array_equal(a.words, b.words)
    assert np.array_equal(a.energy, b.energy)
    assert np.array_equal(b.energy, csr.energies(b.words))
    # a transposed replica index (rho = lane * cells + slot) is a different chain
    c = ob.OracleState(csr, 3, 41, view=view)
    c.set_replica_temperatures(np.tile(1.0 / ladder, 64))
    c.run(9, exchange=False)
    assert not np.array_equal(a.words, c.words)


def test_per_replica_temperatures_follow_rep_mod_len():
    # temps shorter than R repeat: T(rho) = temps[rho % len].  Lanes of one slot at
    # different temperatures evolve as lanes of slots at those temperatures would.
    csr = frustrated7()
    temps = np.array([0.25, 4.0])
    st = ob.OracleState(csr, 1, 7)
    st.set_replica_temperatures(temps)
    assert np.allclose(1.0 / st.replica_betas, temps[np.arange(64) % 2])
    st.run(20, exchange=False)
    assert np.array_equal(st.energy, csr.energies(st.words))
    # cold lanes end lower on average than hot lanes
    e = st.energy[0]
    assert e[0::2].mean() < e[1::2].mean()


def reference_exchange(energies, bits, betas, seed, xround0, rounds):
    """replica_exchange.rs:61-109 as the reference performs it: configurations and
    energies swapped physically, round k of the call from start k & 1, uniforms drawn for
    every pair (keyed as the device keys them: Philox(i >> 2, 0, xround, 8 << 24)[i & 3])."""
    e = list(energies)
    x = [row.copy() for row in bits]
    R = len(e)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for k in range(rounds):
        xr = xround0 + k
        i = k & 1
        while i + 1 < R:
            delta = (betas[i] - betas[i + 1]) * float(e[i] - e[i + 1])
            u = ob.philox([i >> 2, 0, xr, 8 << 24], key)[i & 3]
            if delta >= 0.0 or u < ob.exchange_threshold(delta):
                e[i], e[i + 1] = e[i + 1], e[i]
                x[i], x[i + 1] = x[i + 1], x[i]
            i += 2
    return np.array(e), np.array(x)


def test_exchange_matches_the_references_physical_swaps():
    # replica_exchange.rs:139-196 in shape: the frustrated 7-site instance, 128 replicas,
    # temperatures 0.1 + 0.3 k (16 of them, repeating), replicas diverged by sweeps first.
    csr = frustrated7()
    st = ob.OracleState(csr, 2, 19)
    st.set_replica_temperatures(0.1 + 0.3 * np.arange(16))
    st.run(5, exchange=False)
    e0, x0 = logical(st)
    before = np.sort(e0)
    # one apply of 6 rounds, then two applies of 3 (each restarting at start 0)
    e_ref, x_ref = reference_exchange(e0, x0, st.replica_betas, st.seed, st.xround, 6)
    st.exchange_replicas(6)
    e1, x1 = logical(st)
    assert np.array_equal(e1, e_ref) and np.array_equal(x1, x_ref)
    e_ref, x_ref = reference_exchange(e1, x1, st.replica_betas, st.seed, st.xround, 3)
    e_ref, x_ref = reference_exchange(e_ref, x_ref, st.replica_betas, st.seed, st.xround + 3, 3)
    st.exchange_replicas(3)
    st.exchange_replicas(3)
    e2, x2 = logical(st)
    assert np.array_equal(e2, e_ref) and np.array_equal(x2, x_ref)
    assert np.array_equal(np.sort(e2), before)            # energies only permuted
    assert np.array_equal(st.energy, csr.energies(st.words))
    accepted = int(st.stats[3::2].sum())
    assert 0 < accepted < int(st.stats[2::2].sum())
    # starts 0,1,0,1,0,1 then 0,1,0 twice: pair 0 is attempted in the 7 start-0 rounds,
    # pair 1 in the 5 start-1 rounds
    att = st.stats[2::2]
    assert att[0] == 7 and att[1] == 5


def test_exchange_exact_rates():
    csr = ob.CSR(1, 0, [0], [], [], [])
    # favourable: replica 0 hot with higher energy, replica 1 cold with lower -> always swap
    st = ob.OracleState(csr, 1, 3, words=np.zeros((1, 1), np.uint64))
    st.set_replica_temperatures(np.array([2.0, 0.5] * 32))
    st.energy[:] = np.tile([0, 10], 32)  # delta = (0.5 - 2)(0 - 10) > 0 for every even pair
    st.exchange_replicas(1)
    assert int(st.stats[3::2][0::2].sum()) == 32
    # delta = (0 - 1)(1 - 0) = -1 on every even pair: rate e^-1 over many keyed rounds
    acc = tries = 0
    for k in range(200):
        st = ob.OracleState(csr, 1, 11 + k, words=np.zeros((1, 1), np.uint64))
        st.set_replica_temperatures(np.array([np.inf, 1.0] * 32))
        st.energy[:] = np.tile([1, 0], 32)
        st.exchange_replicas(1)
        acc += int(st.stats[3::2][0::2].sum())
        tries += 32
    p = math.exp(-1.0)
    assert abs(acc / tries - p) < 5 * math.sqrt(p * (1 - p) / tries)


def joint_boltzmann_pvalues(flip_sign=False, seed=5):
    """64 replicas of the 5-spin model at 4 temperatures (T = temps[rho % 4]), Metropolis
    sweeps and ReplicaExchange rounds; per temperature, the configurations held there
    against exp(-E/T)/Z by enumeration.  flip_sign plants a sign mistake in the oracle's
    exchange (delta negated)."""
    csr = ob.CSR(*BOLTZ_ARGS)
    n = csr.n
    states = list(itertools.product([0, 1], repeat=n))
    es = np.array([csr.energy(list(s)) for s in states], dtype=np.float64)
    temps = np.array([0.6, 1.0, 1.8, 3.0])
    st = ob.OracleState(csr, 1, seed)
    st.set_replica_temperatures(temps)
    good = st.replica_betas.copy()
    bad = -good
    counts = np.zeros((4, len(states)))
    for it in range(420):
        st.run(1, exchange=False)
        if flip_sign:
            st.replica_betas = bad
        st.exchange_replicas(2)
        st.replica_betas = good
        if it < 20:
            continue
        e, bits = logical(st)
        for rho in range(64):
            idx = int(sum(int(b) << (n - 1 - i) for i, b in enumerate(bits[rho])))
            counts[rho % 4, idx] += 1
    out = []
    for t in range(4):
        p = np.exp(-(es - es.min()) / temps[t])
        p /= p.sum()
        out.append(chisquare(counts[t], p * counts[t].sum())[1])
    return out


def test_replica_view_chain_samples_the_joint_boltzmann_law():
    assert min(joint_boltzmann_pvalues()) > 1e-4


def test_joint_law_check_catches_a_planted_exchange_sign_mistake():
    assert min(joint_boltzmann_pvalues(flip_sign=True)) < 1e-6
