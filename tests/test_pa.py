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
