"""Pins the CPU oracle before it is trusted as the GPU checker.

Every check here is one a plausible mistake in the oracle would fail:
published known-answer vectors, the reference's own unit-test cases,
closed forms, brute force and exact enumeration.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import binding as ob
from ising_engine_amd import instances


# --------------------------------------------------------------------------
# Philox4x32-10: the known-answer vectors published with Random123
# (Salmon et al., SC'11), entered by hand.
# --------------------------------------------------------------------------
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    assert list(ob.philox(ctr, key)) == expect


def test_philox_uniformity_and_independence():
    # Mirrors the property checks of tests/test_rng_quality.rs:10-70: mean of
    # U[0,1) within 5 sigma, lane (word) independence, determinism.
    out = np.array([ob.philox([i, 7, 3, 1 << 24], [11, 13]) for i in range(4000)], dtype=np.float64)
    u = out / 2.0 ** 32
    sigma = math.sqrt(1 / 12 / u.size)
    assert abs(u.mean() - 0.5) < 5 * sigma
    c = np.corrcoef(u.T)
    assert np.all(np.abs(c - np.eye(4)) < 0.1)
    assert list(ob.philox([5, 6, 7, 8], [1, 2])) == list(ob.philox([5, 6, 7, 8], [1, 2]))


# --------------------------------------------------------------------------
# det_exp and the integer thresholds
# --------------------------------------------------------------------------
def test_det_exp_matches_libm_within_2ulp():
    xs = np.concatenate([np.linspace(-39.99, -1e-9, 20001), -np.logspace(-12, 1.5, 2001)])
    for x in xs:
        e = ob.det_exp(float(x))
        ref = math.exp(float(x))
        assert abs(e - ref) <= 2 * math.ulp(ref), x


def test_det_exp_special_points():
    assert ob.det_exp(0.0) == 1.0
    assert ob.det_exp(-40.0) == 0.0
    assert ob.det_exp(-1e300) == 0.0
    assert ob.det_exp(float("nan")) == 0.0
    assert ob.det_exp(-math.log(2.0)) == pytest.approx(0.5, abs=1e-16)


def test_threshold_closed_form():
    for beta in [0.0, 0.1, 0.5, 1.0, 2.0, 3.0, 7.5]:
        for de in [1, 2, 3, 4, 8, 12, 100, 1000]:
            p = math.exp(-beta * de)
            t = ob.threshold(beta, de)
            expect = min(math.floor(p * 2 ** 32), 2 ** 32 - 1)
            assert abs(t - expect) <= 1, (beta, de)
    assert ob.threshold(0.0, 5) == 0xFFFFFFFF
    assert ob.threshold(10.0, 5) == 0
    # monotone non-increasing in dE
    ts = [ob.threshold(0.3, d) for d in range(1, 200)]
    assert all(a >= b for a, b in zip(ts, ts[1:]))


def test_exchange_threshold_rates():
    # tests/test_population_annealing.rs:354-415: exp(-15) for an unfavourable
    # swap of h=10 between T=2.0 and T=0.5.
    assert ob.exchange_threshold(-15.0) == math.floor(math.exp(-15.0) * 2 ** 32)


# --------------------------------------------------------------------------
# ProblemIR: the reference's own unit tests (src/engine_v2/ir.rs:164-219)
# --------------------------------------------------------------------------
def brute_min(csr):
    best = None
    for bits in range(1 << csr.n):
        x = [(bits >> i) & 1 for i in range(csr.n)]
        e = csr.energy(x)
        best = e if best is None else min(best, e)
    return best


def test_ir_energy_matches_brute_force_definition():
    csr = ob.CSR(3, 1, [2, -3, 1], [0, 1], [1, 2], [-4, 2])
    assert csr.energy([0, 0, 0]) == 1
    assert csr.energy([1, 1, 0]) == 1 + 2 - 3 - 4
    assert csr.energy([1, 1, 1]) == 1 + 2 - 3 + 1 - 4 + 2
    assert brute_min(csr) == -4


def test_ir_duplicate_pairs_accumulate():
    a = ob.CSR(2, 0, [0, 0], [0, 0], [1, 1], [3, 1])
    b = ob.CSR(2, 0, [0, 0], [0], [1], [4])
    assert a.energy([1, 1]) == b.energy([1, 1]) == 4
    assert list(a.col_idx[: a.nnz]) == list(b.col_idx[: b.nnz])
    z = ob.CSR(2, 0, [0, 0], [0, 0], [1, 1], [3, -3])  # sums to zero -> dropped
    assert z.nnz == 0


def test_ir_rows_sorted_and_symmetric():
    qb = instances.random_qubo(60, 5, seed=3)
    csr = ob.CSR.from_qubo(qb)
    for i in range(csr.n):
        row = csr.col_idx[csr.row_ptr[i]:csr.row_ptr[i + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)
        for e in range(csr.row_ptr[i], csr.row_ptr[i + 1]):
            j = csr.col_idx[e]
            back = csr.col_idx[csr.row_ptr[j]:csr.row_ptr[j + 1]]
            k = np.searchsorted(back, i)
            assert back[k] == i and csr.w[csr.row_ptr[j] + k] == csr.w[e]


def test_ir_rejects_bad_pairs():
    with pytest.raises(ValueError):
        ob.CSR(3, 0, [0, 0, 0], [1], [1], [1])
    with pytest.raises(ValueError):
        ob.CSR(3, 0, [0, 0, 0], [0], [3], [1])


def test_greedy_coloring_triangle_path():
    # ir.rs:182-211: path + triangle 0-1-2-0, 2-3, 3-4.
    csr = ob.CSR(5, 0, [0] * 5, [0, 1, 0, 2, 3], [1, 2, 2, 3, 4], [1] * 5)
    col = csr.coloring()
    assert list(col[:3]) == [0, 1, 2]
    for i in range(5):
        for e in range(csr.row_ptr[i], csr.row_ptr[i + 1]):
            assert col[i] != col[csr.col_idx[e]]


@pytest.mark.parametrize("L,dim", [(4, 2), (32, 2), (8, 3), (16, 3)])
def test_lattice_coloring_is_checkerboard(L, dim):
    csr = ob.CSR.from_qubo(instances.hypercubic_pm_j(L, dim, seed=1))
    col = csr.coloring()
    idx = np.arange(csr.n)
    parity = sum((idx // L ** a) % L for a in range(dim)) % 2
    assert np.array_equal(col, parity)


def test_energy_equals_ising_hamiltonian():
    # Ising -> QUBO lowering (src/bin/sk_parisi_benchmark.rs:56-110) is exact.
    qb = instances.hypercubic_pm_j(4, 2, seed=5)
    csr = ob.CSR.from_qubo(qb)
    rng = np.random.default_rng(0)
    J = -qb.q // 4
    for _ in range(20):
        x = rng.integers(0, 2, size=qb.n)
        s = 1 - 2 * x
        h = -int((J * s[qb.pi.astype(int)] * s[qb.pj.astype(int)]).sum())
        assert csr.energy(x) == h == qb.energy(x)


# --------------------------------------------------------------------------
# Sweep invariants
# --------------------------------------------------------------------------
def frustrated7():
    # metropolis_sweep.rs:156-172 / replica_exchange.rs:120-136 instance.
    return ob.CSR(7, -2, [1, -2, 3, 0, -1, 2, 1],
                  [0, 1, 2, 0, 3, 4, 5, 2], [1, 2, 3, 3, 4, 5, 6, 6],
                  [-1, 2, -3, 1, -1, 2, -2, 1])


@pytest.mark.parametrize("view", [ob.VIEW_WORD, ob.VIEW_PLANE])
def test_tracked_energy_equals_recompute(view):
    # Mirrors tests/test_trotter_correctness.rs:111-140 and the audit() == 0
    # contract of sparse_bitslice.rs:317-346.
    csr = frustrated7()
    st = ob.OracleState(csr, cells=3, seed=99, view=view)
    st.run(50, exchange=True, ladder=[0.2, 0.7, 1.5])
    assert np.array_equal(st.energy, csr.energies(st.words))
    assert st.stats[0] <= st.stats[1] == 50 * 7 * 64 * 3


def test_fields_and_delta_e_match_brute_force():
    csr = frustrated7()
    words = ob.init_words(7, 1, 5)
    f = csr.fields(words)
    for r in range(64):
        x = [(int(words[0, i]) >> r) & 1 for i in range(7)]
        e0 = csr.energy(x)
        for i in range(7):
            y = list(x)
            y[i] ^= 1
            de = csr.energy(y) - e0
            assert de == (1 - 2 * x[i]) * f[0, i, r]


def test_zero_temperature_descent_is_monotone():
    # gibbs_color_sweep.rs:258-292: at T -> 0 energies never increase.
    qb = instances.random_qubo(40, 4, seed=2)
    csr = ob.CSR.from_qubo(qb)
    st = ob.OracleState(csr, cells=1, seed=3)
    prev = st.energy.copy()
    for _ in range(10):
        st.run(1, exchange=False, ladder=[50.0])
        assert np.all(st.energy <= prev)
        prev = st.energy.copy()


def test_infinite_temperature_accepts_everything():
    csr = frustrated7()
    st = ob.OracleState(csr, cells=1, seed=4)
    st.run(3, exchange=False, ladder=[0.0])
    # thr = 2^32-1 at beta = 0: rejected only when u == 2^32-1.
    assert st.stats[0] == st.stats[1]


BOLTZ_ARGS = (5, 0, [1, -2, 1, 0, 2], [0, 1, 2, 0, 1], [1, 2, 3, 4, 3], [-2, 3, -1, 2, -2])


def boltzmann_pvalue(csr_target, beta_target, csr_run, beta_run, view, seed=12):
    """Chi-square p-value of the 64 oracle chains run on (csr_run, beta_run), sampled
    300 times after burn-in, against the exact distribution exp(-beta E)/Z of
    (csr_target, beta_target) by enumeration."""
    n = csr_target.n
    states = list(itertools.product([0, 1], repeat=n))
    es = np.array([csr_target.energy(list(s)) for s in states], dtype=np.float64)
    p = np.exp(-beta_target * (es - es.min()))
    p /= p.sum()
    st = ob.OracleState(csr_run, cells=1, seed=seed, view=view)
    st.run(20, exchange=False, ladder=[beta_run])
    counts = np.zeros(len(states))
    for _ in range(300):
        st.run(1, exchange=False, ladder=[beta_run])
        w = st.words[0]
        for r in range(64):
            idx = sum(((int(w[i]) >> r) & 1) << (n - 1 - i) for i in range(n))
            counts[idx] += 1
    from scipy.stats import chisquare
    return chisquare(counts, p * counts.sum())[1]


@pytest.mark.parametrize("view", [ob.VIEW_WORD, ob.VIEW_PLANE])
def test_metropolis_samples_exact_boltzmann(view):
    # Exact enumeration on a 5-spin model: the empirical distribution of the
    # 64 chains after burn-in matches exp(-beta E)/Z (chi-square, p > 1e-4).
    csr = ob.CSR(*BOLTZ_ARGS)
    assert boltzmann_pvalue(csr, 0.6, csr, 0.6, view) > 1e-4


@pytest.mark.parametrize("view", [ob.VIEW_WORD, ob.VIEW_PLANE])
@pytest.mark.parametrize("mistake", ["beta_doubled", "linear_sign", "coupling_dropped"])
def test_boltzmann_check_catches_planted_mistakes(view, mistake):
    # The same check fails when the chains run a plausibly wrong sweep: a factor
    # of two in beta, a flipped sign of the linear term, one coupling lost.
    n, off, lin, pi, pj, q = BOLTZ_ARGS
    target = ob.CSR(*BOLTZ_ARGS)
    beta = 0.6
    if mistake == "beta_doubled":
        run, beta_run = target, 2 * beta
    elif mistake == "linear_sign":
        run, beta_run = ob.CSR(n, off, [-v for v in lin], pi, pj, q), beta
    else:
        run, beta_run = ob.CSR(n, off, lin, pi[:-1], pj[:-1], q[:-1]), beta
    assert boltzmann_pvalue(target, beta, run, beta_run, view) < 1e-6


def test_views_differ_but_agree_in_law():
    qb = instances.hypercubic_pm_j(8, 2, seed=9)
    csr = ob.CSR.from_qubo(qb)
    a = ob.OracleState(csr, cells=1, seed=1, view=ob.VIEW_WORD)
    b = ob.OracleState(csr, cells=1, seed=1, view=ob.VIEW_PLANE)
    a.run(200, exchange=False, ladder=[0.4])
    b.run(200, exchange=False, ladder=[0.4])
    assert not np.array_equal(a.words, b.words)
    ea, eb = a.energy.mean() / csr.n, b.energy.mean() / csr.n
    assert abs(ea - eb) < 0.05


# --------------------------------------------------------------------------
# Exchange
# --------------------------------------------------------------------------
def test_exchange_exact_rates():
    # Favourable swap: rate 1.  exp(-1): within 5 sigma of e^-1.
    beta = np.array([0.5, 2.0])
    temp_of = np.array([[0] * 64, [1] * 64], dtype=np.uint32)
    stats = np.zeros(4, dtype=np.uint64)
    e = np.array([[0] * 64, [10] * 64], dtype=np.int64)  # hot E=0, cold E=10
    ob.exchange(2, 1, 0, beta, e, temp_of, stats)
    assert stats[3] == 64 and np.all(temp_of == [[1] * 64, [0] * 64])
    acc = 0
    tries = 0
    e = np.array([[1] * 64, [0] * 64], dtype=np.int64)
    beta = np.array([0.0, 1.0])  # delta = (0 - 1)(1 - 0) = -1
    for sweep in range(400):
        temp_of = np.array([[0] * 64, [1] * 64], dtype=np.uint32)
        stats = np.zeros(4, dtype=np.uint64)
        ob.exchange(2, 7, 2 * sweep, beta, e, temp_of, stats)
        acc += int(stats[3])
        tries += int(stats[2])
    p = math.exp(-1.0)
    assert abs(acc / tries - p) < 5 * math.sqrt(p * (1 - p) / tries)


def test_exchange_is_a_per_lane_permutation_and_deo_alternates():
    csr = frustrated7()
    st = ob.OracleState(csr, cells=5, seed=21)
    for k in range(7):
        st.run(1, exchange=True, ladder=[0.1, 0.3, 0.6, 1.0, 2.0])
        for r in range(64):
            assert sorted(st.temp_of[:, r]) == list(range(5))
    # pairs (0,1),(2,3) on even sweeps, (1,2),(3,4) on odd: 4 even, 3 odd sweeps
    att = st.stats[2::2]
    assert list(att) == [4 * 64, 3 * 64, 4 * 64, 3 * 64]


# --------------------------------------------------------------------------
# Ladders
# --------------------------------------------------------------------------
def test_geometric_ladder_closed_form():
    t = ob.geometric_ladder(5, 2.0, 0.125)
    assert np.allclose(t, [2.0, 1.0, 0.5, 0.25, 0.125], rtol=1e-14)
    assert list(ob.geometric_ladder(1, 0.5, 3.0)) == [3.0]
    assert ob.geometric_ladder(3, 1.0, 0.0)[-1] == pytest.approx(1e-6)


def test_suggested_range():
    # schedule.rs:21-43: T_max = max local / ln 2, T_min = min(|w|min / ln 100, T_max/2)
    csr = ob.CSR(3, 0, [2, -3, 1], [0, 1], [1, 2], [-4, 2])
    tmax, tmin = ob.suggested_range(csr)
    assert tmax == pytest.approx((3 + 4 + 2) / math.log(2))
    assert tmin == pytest.approx(1 / math.log(100))
