"""Pins the oracle's heat-bath (Gibbs) colour sweep, gibbs_color_sweep.rs:119-174.

The new spin of (site, replica) is 1 iff u < floor(2^32 / (1 + exp(beta h))), with h the
local field and u the word-view uniform.  Checks: the threshold against libm and its
exact special values, the rule itself at beta = 0 (where it reads the uniforms' top bit
and nothing else), the exact Boltzmann distribution by enumeration (and its failure on
planted mistakes), the reference's own cold-limit test, and energy bookkeeping.
"""
import math

import numpy as np
import pytest

from oracle import binding as ob
from ising_engine_amd import instances
from test_oracle_descent import ir7
from test_oracle_kat import BOLTZ_ARGS, boltzmann_pvalue


@pytest.mark.parametrize("beta", [0.0, 0.013, 0.37, 1.0, 2.9])
def test_threshold_against_libm(beta):
    for h in list(range(-60, 61)) + [-1000, 999, 12345]:
        p = 1.0 / (1.0 + math.exp(min(beta * h, 700.0)))
        ref = math.floor(p * 2.0 ** 32)
        got = ob.heatbath_threshold(beta, h)
        assert abs(got - ref) <= 1, (beta, h, got, ref)
        assert 0 <= got <= 2 ** 32


def test_threshold_special_values():
    assert ob.heatbath_threshold(0.0, 7) == 2 ** 31          # beta = 0: a fair coin
    assert ob.heatbath_threshold(3.0, 0) == 2 ** 31          # h = 0: a fair coin
    assert ob.heatbath_threshold(1e300, 5) == 0              # T = 0 limit, h > 0: never 1
    assert ob.heatbath_threshold(1e300, -5) == 2 ** 32       # T = 0 limit, h < 0: always 1
    assert ob.heatbath_threshold(1e300, 0) == 2 ** 31
    for beta in (0.1, 0.8, 5.0):
        for h in range(1, 80):
            s = ob.heatbath_threshold(beta, h) + ob.heatbath_threshold(beta, -h)
            assert s in (2 ** 32 - 1, 2 ** 32)                # P(h) + P(-h) = 1
            assert ob.heatbath_threshold(beta, h) <= ob.heatbath_threshold(beta, h - 1)


def test_beta_zero_reads_the_top_bit_of_the_uniform():
    """At beta = 0 the new spin is 1 iff u < 2^31, whatever the field: after one sweep,
    x[site] of replica r equals NOT bit 31 of the word-view uniform of (pos, slot, sweep, r)."""
    qb = instances.random_qubo(40, 4, 3)
    csr = ob.CSR.from_qubo(qb)
    order, _ = csr.color_order()
    seed, cells = 0xB0B, 2
    st = ob.OracleState(csr, cells, seed, view=ob.VIEW_HEATBATH)
    st.run(3, exchange=False, ladder=[0.7, 0.7])
    sweep = st.sweep
    st.run(1, exchange=False, ladder=[0.0, 0.0])
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for s in range(cells):
        for pos, site in enumerate(order):
            for r in range(64):
                u = ob.philox([pos >> 2, s, sweep, (1 << 24) | r], key)[pos & 3]
                assert ((int(st.words[s, site]) >> r) & 1) == (1 if u < 2 ** 31 else 0)


def test_heatbath_samples_exact_boltzmann():
    csr = ob.CSR(*BOLTZ_ARGS)
    assert boltzmann_pvalue(csr, 0.6, csr, 0.6, ob.VIEW_HEATBATH) > 1e-4


@pytest.mark.parametrize("mistake", ["beta_doubled", "field_sign", "coupling_dropped"])
def test_heatbath_boltzmann_check_catches_planted_mistakes(mistake):
    n, off, lin, pi, pj, q = BOLTZ_ARGS
    target = ob.CSR(*BOLTZ_ARGS)
    beta = 0.6
    if mistake == "beta_doubled":
        run, beta_run = target, 2 * beta
    elif mistake == "field_sign":  # P(x=1) = 1/(1+exp(-beta h)): the whole field negated
        run, beta_run = ob.CSR(n, -off, [-v for v in lin], pi, pj, [-v for v in q]), beta
    else:
        run, beta_run = ob.CSR(n, off, lin, pi[:-1], pj[:-1], q[:-1]), beta
    assert boltzmann_pvalue(target, beta, run, beta_run, ob.VIEW_HEATBATH) < 1e-6


def test_cold_heatbath_is_monotone_descent():
    """The reference's cold_temperature_is_monotone_descent (gibbs_color_sweep.rs:256-290):
    at T = 1e-9 every replica's energy is non-increasing, sweep after sweep."""
    csr = ob.CSR.from_qubo(ir7())
    st = ob.OracleState(csr, 1, 5, words=np.zeros((1, 7), dtype=np.uint64),
                        view=ob.VIEW_HEATBATH)
    prev = st.energy.copy()
    for _ in range(5):
        st.run(1, exchange=False, ladder=[1e9])
        assert (st.energy <= prev).all()
        prev = st.energy.copy()
    np.testing.assert_array_equal(st.energy, csr.energies(st.words))


def test_tracked_energies_and_exchange():
    qb = instances.hypercubic_pm_j(6, 2, 4)
    csr = ob.CSR.from_qubo(qb)
    st = ob.OracleState(csr, 4, 77, view=ob.VIEW_HEATBATH)
    st.run(30, exchange=True, ladder=instances.geometric_betas(2.5, 0.4, 4))
    np.testing.assert_array_equal(st.energy, csr.energies(st.words))
    assert st.stats[0] > 0 and st.stats[3::2].sum() > 0
