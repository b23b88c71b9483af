"""Heat-bath (Gibbs) colour sweeps on the device against the oracle, bit-exact: words, int64
energies, temperature permutation, flip counts and exchange statistics."""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def make(ctx, qb):
    return _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)


def compare(st, ost):
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    np.testing.assert_array_equal(st.temp_of(), ost.temp_of)
    s, pa, pc = st.stats()
    assert s.accepted == int(ost.stats[0]) and s.proposed == int(ost.stats[1])
    np.testing.assert_array_equal(pa, ost.stats[2::2])
    np.testing.assert_array_equal(pc, ost.stats[3::2])
    np.testing.assert_array_equal(st.recompute_energies(), ost.energy)


CASES = [
    # name, factory, state kernel, cells, exchange, schedule
    ("qubo_n300_generic_pt", lambda: instances.random_qubo(300, 6, 61, offdiag=30, diag=50),
     _lib.KERNEL_GENERIC, 4, True, False),
    ("qubo_n200_wide_range_global_table", lambda: instances.random_qubo(200, 8, 62, offdiag=3000,
                                                                         diag=5000),
     _lib.KERNEL_GENERIC, 3, True, False),
    ("pmJ_3d_L6_multispin_state_pt", lambda: instances.hypercubic_pm_j(6, 3, 63),
     _lib.KERNEL_MULTISPIN, 4, True, False),
    ("er_pmJ_n500_anneal", lambda: instances.erdos_renyi_pm_j(500, 6, 64),
     _lib.KERNEL_AUTO, 1, False, True),
    ("sk_n150_dense_state_anneal", lambda: instances.sk_int8(150, 65), _lib.KERNEL_DENSE, 2,
     False, True),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_heatbath_parity_with_oracle(ctx, case):
    name, factory, kernel, cells, exchange, use_sched = case
    print(f"[heatbath] {name}", flush=True)
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    seed = 0x6188_5000 + len(name)
    st = _lib.State(model, cells, seed, kernel=kernel)
    ost = ob.OracleState(csr, cells, seed, view=ob.VIEW_HEATBATH)
    t_max, t_min = instances.suggested_temp_range(qb)
    sweeps = 12
    if use_sched:
        sched = np.repeat(instances.anneal_schedule(1.0 / t_max, 1.0 / t_min, sweeps), cells,
                          axis=1)
        st.run(5, beta_schedule=sched[:5], heatbath=True)
        st.run(sweeps - 5, beta_schedule=sched[5:], heatbath=True)
        ost.run(sweeps, exchange=False, schedule=sched)
    else:
        ladder = instances.geometric_betas(t_max, t_min, cells)
        st.set_ladder(ladder)
        st.run(5, exchange=exchange, heatbath=True)
        st.run(sweeps - 5, exchange=exchange, heatbath=True)
        ost.run(sweeps, exchange=exchange, ladder=ladder)
    compare(st, ost)
    s = st.stats()[0]
    assert 0 < s.accepted < s.proposed


def test_heatbath_and_metropolis_interleave(ctx):
    """A run may switch rules between calls: the sweep counter keys both, the state's own
    (multispin) kernel takes the Metropolis sweeps with its plane-view uniforms."""
    qb = instances.hypercubic_pm_j(8, 2, 66)
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    st = _lib.State(model, 2, 99)
    assert st.stats()[0].kernel == _lib.KERNEL_MULTISPIN
    ladder = [0.9, 0.4]
    st.set_ladder(ladder)
    hb = ob.OracleState(csr, 2, 99, view=ob.VIEW_HEATBATH)
    st.run(4, exchange=True, heatbath=True)
    hb.run(4, exchange=True, ladder=ladder)
    np.testing.assert_array_equal(st.spins(), hb.words)
    met = ob.OracleState(csr, 2, 99, words=hb.words, view=ob.VIEW_PLANE)
    met.temp_of[:] = hb.temp_of
    met.sweep, met.xround = hb.sweep, hb.xround
    st.run(3, exchange=True)
    met.run(3, exchange=True, ladder=ladder)
    np.testing.assert_array_equal(st.spins(), met.words)
    np.testing.assert_array_equal(st.energies(), met.energy)
    np.testing.assert_array_equal(st.temp_of(), met.temp_of)


def test_cold_heatbath_is_monotone_on_device(ctx):
    qb = instances.random_qubo(400, 5, 67)
    model = make(ctx, qb)
    st = _lib.State(model, 2, 3, kernel=_lib.KERNEL_GENERIC)
    st.set_ladder([1e9, 1e9])
    prev = st.energies().copy()
    for _ in range(6):
        st.run(1, heatbath=True)
        e = st.energies()
        assert (e <= prev).all()
        prev = e.copy()
    np.testing.assert_array_equal(st.recompute_energies(), prev)
