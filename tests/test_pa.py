"""Population-annealing arithmetic and the resampling / population-exchange oracle.

The C ABI host functions (pa.cpp) against the numpy restatement (oracle/pa.py) and closed
forms; apply_resample and population-restricted exchange in the oracle by their defining
properties.
"""
import math

import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob
from oracle import pa


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_weights_ess_and_free_energy(seed):
    lw = np.random.default_rng(seed).normal(0, 3, size=256)
    w, ess = _lib.pa_normalize(lw)
    np.testing.assert_allclose(w, pa.normalized_weights(lw), rtol=1e-13, atol=0)
    assert abs(ess - pa.effective_sample_size(w)) < 1e-9 * ess
    assert abs(_lib.pa_free_energy_increment(lw) - pa.free_energy_increment(lw)) < 1e-12
    w2, _ = _lib.pa_normalize(lw + 123.0)                      # shift invariance
    np.testing.assert_allclose(w2, w, rtol=1e-12)


def test_closed_forms():
    w, ess = _lib.pa_normalize(np.full(128, -7.5))
    assert np.all(w == 1.0 / 128) and abs(ess - 128) < 1e-9    # equal weights: ESS = R
    lw = np.full(64, -1e4)
    lw[17] = 0.0
    w, ess = _lib.pa_normalize(lw)
    assert w[17] == 1.0 and ess == 1.0                          # one survivor: ESS = 1
    assert abs(_lib.pa_free_energy_increment(np.full(64, 2.5)) - 2.5) < 1e-12
    # two levels: log(mean(e^a, e^b))
    assert abs(_lib.pa_free_energy_increment([0.0, math.log(3.0)]) - math.log(2.0)) < 1e-12
    assert _lib.pa_family_entropy(np.arange(64)) == (1.0, 64)
    assert _lib.pa_family_entropy(np.zeros(64)) == (64.0, 1)
    assert _lib.pa_family_entropy([0, 0, 1, 1]) == (2.0, 2)


@pytest.mark.parametrize("u", [0.0, 0.3, 0.999999])
def test_systematic_resample(u):
    rng = np.random.default_rng(int(u * 1e6) + 5)
    w = pa.normalized_weights(rng.normal(0, 2, size=192))
    c = _lib.pa_systematic_resample(w, u)
    np.testing.assert_array_equal(c, pa.systematic_resample(w, u))
    assert c.sum() == 192
    assert np.all(np.abs(c - 192 * w) < 1.0 + 1e-9)           # |count - R w| < 1
    assert np.all(_lib.pa_systematic_resample(np.full(64, 1 / 64), u) == 1)


def test_parents_shuffle():
    c = _lib.pa_systematic_resample(pa.normalized_weights(np.linspace(0, 3, 128)), 0.4)
    p = _lib.pa_parents(c, seed=99, event=3)
    np.testing.assert_array_equal(p, pa.parents(c, 99, 3))
    np.testing.assert_array_equal(np.bincount(p, minlength=128), c)   # the counts' multiset
    assert not np.array_equal(p, np.sort(p))                           # shuffled
    assert not np.array_equal(p, _lib.pa_parents(c, seed=99, event=4))
    with pytest.raises(_lib.IgeError):
        _lib.pa_parents(np.ones(10, np.uint32) * 2, 1, 0)               # sum != R


def test_population_exchange_stays_inside_populations():
    qb = instances.hypercubic_pm_j(6, 2, 3)
    csr = ob.CSR.from_qubo(qb)
    st = ob.OracleState(csr, 6, 11)
    st.pops = 2                                                # two ladders of 3
    ladder = np.tile(instances.geometric_betas(2.0, 0.4, 3), 2)
    st.run(20, exchange=True, ladder=ladder)
    np.testing.assert_array_equal(st.energy, csr.energies(st.words))
    assert st.stats[2 + 2 * 2] == 0                            # pair (2, 3) crosses: never tried
    assert st.stats[2 + 2 * 0] > 0 and st.stats[2 + 2 * 3] > 0
    for s in range(6):                                         # a lane's temperature stays in
        assert np.all(st.temp_of[s] // 3 == s // 3)            # its population's levels


def test_resample_copies_whole_ladders():
    qb = instances.random_qubo(40, 3, 12)
    csr = ob.CSR.from_qubo(qb)
    st = ob.OracleState(csr, 6, 13)
    st.pops = 2
    st.run(8, exchange=True, ladder=np.tile([1.0, 0.6, 0.3], 2))
    before_w, before_e, tof = st.words.copy(), st.energy.copy(), st.temp_of.copy()
    so = np.zeros((6, 64), dtype=np.int64)
    for s in range(6):
        so[tof[s], np.arange(64)] = s
    par = np.random.default_rng(4).integers(0, 128, size=128).astype(np.uint32)
    st.resample(par)
    for m in range(128):
        p, r = divmod(m, 64)
        ps, rs = divmod(int(par[m]), 64)
        for t in range(3):
            d, s = so[p * 3 + t, r], so[ps * 3 + t, rs]
            np.testing.assert_array_equal((st.words[d] >> np.uint64(r)) & np.uint64(1),
                                          (before_w[s] >> np.uint64(rs)) & np.uint64(1))
            assert st.energy[d, r] == before_e[s, rs]
    np.testing.assert_array_equal(st.energy, csr.energies(st.words))
    np.testing.assert_array_equal(st.temp_of, tof)


# ---- acceptance-uniform ladder tuning (population_annealing.rs:137-148,
# ultimate.rs:599-660) ----

def test_adapt_spacings_reference_cases():
    # tests/test_population_annealing.rs:324-347, on the oracle and on the C ABI
    for f in (pa.adapt_spacings, _lib.pa_adapt_spacings):
        sp = [0.1, 0.3, 0.2, 0.4]
        np.testing.assert_allclose(f(sp, [0.3] * 4), sp, rtol=0, atol=1e-12)
        out = f([0.5, 0.5], [0.9, 0.1])
        assert out[0] > 0.5 and out[1] < 0.5
        assert abs(sum(out) - 1.0) < 1e-12


@pytest.mark.parametrize("seed", [4, 5, 6])
def test_adapt_spacings_abi_equals_the_oracle(seed):
    rng = np.random.default_rng(seed)
    k = int(rng.integers(2, 40))
    sp = rng.uniform(0.001, 0.3, size=k)
    acc = rng.uniform(0.0, 1.0, size=k)
    acc[rng.integers(0, k)] = 0.0           # below the 0.02 floor
    got = _lib.pa_adapt_spacings(sp, acc)
    np.testing.assert_array_equal(got, pa.adapt_spacings(sp, acc))
    # closed form: d sqrt(max(a, 0.02)) times the span ratio
    raw = sp * np.sqrt(np.maximum(acc, 0.02))
    np.testing.assert_allclose(got, raw * sp.sum() / raw.sum(), rtol=1e-12)


def two_level_csr(h):
    return ob.CSR(1, 0, [h], [], [], [])


def test_two_level_swap_rates_match_the_closed_form():
    # one spin, E = h x, 5 temperatures, 64 lanes: the oracle's per-pair DEO acceptance over
    # 4000 steps against the equilibrium closed form
    h = 2
    betas = 1.0 / np.array([4.0, 2.5, 1.6, 1.0, 0.6])
    st = ob.OracleState(two_level_csr(h), 5, 17)
    st.run(200, exchange=True, ladder=betas)           # equilibrate
    before = st.stats.copy()
    st.run(4000, exchange=True, ladder=betas)
    rates = pa.pair_rates(before, st.stats, 1, 5)
    exact = pa.two_level_exchange_rates(h, betas)
    assert np.all(np.abs(rates - exact) < 0.012), (rates, exact)
    # and a planted sign mistake in the closed form would not pass
    wrong = 1.0 - (1.0 - np.exp(-betas * h) / (1 + np.exp(-betas * h)))[:-1] * \
        (np.exp(-betas * h) / (1 + np.exp(-betas * h)))[1:] * \
        (1.0 - np.exp(-(betas[1:] - betas[:-1]) * h))
    assert np.max(np.abs(rates - wrong)) > 0.02


def exact_tuning(h, t_max, t_min, T, iters, transpose=False):
    """adapt_spacings iterated on the two-level closed-form rates; transpose plants an
    index mistake (the rates of pair t applied to spacing T - 2 - t)."""
    alphas = np.array([t / (T - 1) for t in range(T)])
    for _ in range(iters):
        betas = 1.0 / pa.ladder_temperatures(t_max, t_min, alphas)
        rates = pa.two_level_exchange_rates(h, betas)
        if transpose:
            rates = rates[::-1]
        new = pa.adapt_spacings(np.diff(alphas), rates)
        alphas[1:] = np.cumsum(new)
        alphas[-1] = 1.0
    return alphas, pa.two_level_exchange_rates(h, 1.0 / pa.ladder_temperatures(t_max, t_min,
                                                                                  alphas))


def test_tuning_fixed_point_is_uniform_acceptance():
    # the geometric ladder's acceptances on the two-level system are far from uniform; the
    # iteration drives them to one value, and an index mistake does not
    h, t_max, t_min, T = 2.0, 20.0, 0.1, 5
    _, r0 = exact_tuning(h, t_max, t_min, T, 0)
    alphas, r = exact_tuning(h, t_max, t_min, T, 60)
    assert np.ptp(r0) > 0.15
    assert np.ptp(r) < 3e-3
    assert np.all(np.diff(alphas) > 0) and alphas[0] == 0.0 and alphas[-1] == 1.0
    _, rbad = exact_tuning(h, t_max, t_min, T, 60, transpose=True)
    assert np.ptp(rbad) > 0.05


def test_oracle_tuning_on_the_two_level_system_follows_the_closed_form():
    # the oracle's stochastic tuning (4 populations, 200 steps per round) tracks the same
    # 12 iterations run on the closed-form rates, and shrinks the acceptance spread as the
    # reference's tests/test_population_annealing.rs:430-490 requires
    h, t_max, t_min, T, P = 2, 20.0, 0.1, 5, 4
    st = ob.OracleState(two_level_csr(h), P * T, 23)
    st.pops = P
    alphas = pa.tune_ladder(st, t_max, t_min, iters=12, steps=200)
    closed, r = exact_tuning(float(h), t_max, t_min, T, 12)
    assert np.max(np.abs(alphas - closed)) < 0.006, (alphas, closed)
    _, r0 = exact_tuning(float(h), t_max, t_min, T, 0)
    assert np.ptp(r) < 0.5 * np.ptp(r0)
