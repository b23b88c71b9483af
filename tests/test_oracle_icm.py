"""Pins the oracle's isoenergetic cluster moves (icm.rs:91-185) before they check the GPU.

Checked without reusing the oracle's arithmetic: the pair energy is conserved exactly
(energies recomputed from the words), the flipped set is the seed's connected component of
the disagreement graph (an independent BFS in Python), both replicas flip the same set,
identical replicas are left alone, and Metropolis sweeps mixed with ICM moves still sample
the exact Boltzmann distribution.
"""
import itertools

import numpy as np
from scipy.stats import chisquare

from oracle import binding as ob
from ising_engine_amd import instances
from test_oracle_kat import BOLTZ_ARGS


def bits(words, lane):
    return ((words >> np.uint64(lane)) & np.uint64(1)).astype(np.uint8)


def component(csr, dis, seed):
    seen = {seed}
    stack = [seed]
    while stack:
        i = stack.pop()
        for e in range(csr.row_ptr[i], csr.row_ptr[i + 1]):
            j = int(csr.col_idx[e])
            if dis[j] and j not in seen:
                seen.add(j)
                stack.append(j)
    return seen


def slot_of(temp_of):
    cells = temp_of.shape[0]
    so = np.zeros((cells, 64), dtype=np.int64)
    for s in range(cells):
        for r in range(64):
            so[temp_of[s, r], r] = s
    return so


def check_round(csr, st, ladder=None):
    before_w = st.words.copy()
    before_e = csr.energies(before_w)
    so = slot_of(st.temp_of)
    rnd = st.icm_round
    sizes = st.icm(1)
    after_e = csr.energies(st.words)
    np.testing.assert_array_equal(after_e, st.energy)           # tracked == recompute
    key = [st.seed & 0xFFFFFFFF, st.seed >> 32]
    moved = 0
    for t in range(st.cells):
        for r in range(32):
            sa, sb = so[t, r], so[t, r + 32]
            xa, xb = bits(before_w[sa], r), bits(before_w[sb], r + 32)
            fa = np.flatnonzero(bits(before_w[sa], r) ^ bits(st.words[sa], r))
            fb = np.flatnonzero(bits(before_w[sb], r + 32) ^ bits(st.words[sb], r + 32))
            np.testing.assert_array_equal(fa, fb)                  # same set in both
            assert before_e[sa, r] + before_e[sb, r + 32] == \
                after_e[sa, r] + after_e[sb, r + 32]               # isoenergetic
            dis = xa ^ xb
            assert sizes[t, r] == fa.size
            if not dis.any():
                assert fa.size == 0
                continue
            u = ob.philox([t, r, rnd, 6 << 24], key)[0]
            k = (int(u) * int(dis.sum())) >> 32
            seed = int(np.flatnonzero(dis)[k])
            assert set(fa.tolist()) == component(csr, dis, seed)   # the seed's component
            moved += 1
    return moved


def test_icm_moves_are_isoenergetic_clusters():
    qb = instances.random_qubo(60, 3, 8, offdiag=7, diag=9)
    csr = ob.CSR.from_qubo(qb)
    st = ob.OracleState(csr, 3, 21)
    ladder = instances.geometric_betas(1.5, 0.2, 3)
    total = 0
    for _ in range(4):
        st.run(2, exchange=True, ladder=ladder)
        total += check_round(csr, st)
    assert total > 0


def test_icm_on_a_lattice_with_exchanged_temperatures():
    qb = instances.hypercubic_pm_j(6, 2, 5)
    csr = ob.CSR.from_qubo(qb)
    st = ob.OracleState(csr, 4, 22, view=ob.VIEW_PLANE)
    ladder = instances.geometric_betas(2.0, 0.5, 4)
    st.run(10, exchange=True, ladder=ladder)
    assert not np.all(st.temp_of == np.arange(4)[:, None])      # temperatures have moved
    assert check_round(csr, st) > 0


def test_identical_replicas_are_left_alone():
    qb = instances.random_qubo(30, 3, 9)
    csr = ob.CSR.from_qubo(qb)
    lo = np.random.default_rng(1).integers(0, 2 ** 32, size=(2, 30), dtype=np.uint64)
    words = lo | (lo << np.uint64(32))                          # lane r == lane r + 32
    st = ob.OracleState(csr, 2, 5, words=words)
    sizes = st.icm(1)
    assert (sizes == 0).all()
    np.testing.assert_array_equal(st.words, words)


def test_metropolis_with_icm_samples_exact_boltzmann():
    csr = ob.CSR(*BOLTZ_ARGS)
    n = csr.n
    beta = 0.6
    states = list(itertools.product([0, 1], repeat=n))
    es = np.array([csr.energy(list(s)) for s in states], dtype=np.float64)
    p = np.exp(-beta * (es - es.min()))
    p /= p.sum()
    st = ob.OracleState(csr, cells=1, seed=31)
    st.run(20, exchange=False, ladder=[beta])
    counts = np.zeros(len(states))
    for _ in range(300):
        st.run(1, exchange=False, ladder=[beta])
        st.icm(1)
        w = st.words[0]
        for r in range(64):
            counts[sum(((int(w[i]) >> r) & 1) << (n - 1 - i) for i in range(n))] += 1
    assert chisquare(counts, p * counts.sum())[1] > 1e-4
    np.testing.assert_array_equal(st.energy, csr.energies(st.words))
