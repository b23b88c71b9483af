"""GPU parity: the HIP path against the CPU oracle, bit-exact.

Spins, tracked int64 energies, temperature permutation, acceptance and swap
counters, and local fields must equal the oracle's after the same sweeps on
the same seeded inputs.  Larger sizes are checked through size-independent
properties (tracked energy == recompute, per-lane permutation).
"""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def make(ctx, qb):
    return _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)


def frustrated7():
    # metropolis_sweep.rs:156-172 instance (integer couplings).
    return instances.Qubo(n=7, offset=-2, linear=np.array([1, -2, 3, 0, -1, 2, 1]),
                          pi=np.array([0, 1, 2, 0, 3, 4, 5, 2], dtype=np.uint32),
                          pj=np.array([1, 2, 3, 3, 4, 5, 6, 6], dtype=np.uint32),
                          q=np.array([-1, 2, -3, 1, -1, 2, -2, 1]), name="frustrated7")


def compare(st, ost, csr, fields=True):
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    np.testing.assert_array_equal(st.temp_of(), ost.temp_of)
    s, pa, pc = st.stats()
    assert s.accepted == int(ost.stats[0])
    assert s.proposed == int(ost.stats[1])
    np.testing.assert_array_equal(pa, ost.stats[2::2])
    np.testing.assert_array_equal(pc, ost.stats[3::2])
    np.testing.assert_array_equal(st.recompute_energies(), ost.energy)
    if fields:
        np.testing.assert_array_equal(st.fields().astype(np.int64), csr.fields(ost.words))


CASES = [
    # name, qubo factory, cells, kernel, view, exchange, sweeps, schedule?
    ("frustrated7_generic_pt", frustrated7, 3, _lib.KERNEL_GENERIC, ob.VIEW_WORD, True, 20, False),
    ("2d_L8_multispin_anneal", lambda: instances.hypercubic_pm_j(8, 2, 3), 1,
     _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE, False, 30, True),
    ("2d_L8_generic_anneal", lambda: instances.hypercubic_pm_j(8, 2, 3), 1,
     _lib.KERNEL_GENERIC, ob.VIEW_WORD, False, 30, True),
    ("3d_L4_multispin_pt", lambda: instances.hypercubic_pm_j(4, 3, 4), 4,
     _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE, True, 25, False),
    ("3d_L4_generic_pt", lambda: instances.hypercubic_pm_j(4, 3, 4), 4,
     _lib.KERNEL_GENERIC, ob.VIEW_WORD, True, 25, False),
    ("3d_L6_multispin_pt_odd_cells", lambda: instances.hypercubic_pm_j(6, 3, 5), 5,
     _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE, True, 12, False),
    ("qubo_n300_generic_pt", lambda: instances.random_qubo(300, 8, 6), 2,
     _lib.KERNEL_GENERIC, ob.VIEW_WORD, True, 10, False),
    ("er_n400_d6_multispin_pt", lambda: instances.erdos_renyi_pm_j(400, 6, 7), 2,
     _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE, True, 10, False),
    ("er_n300_d12_multispin", lambda: instances.erdos_renyi_pm_j(300, 12, 8), 1,
     _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE, False, 8, False),
    ("er_n200_d18_multispin", lambda: instances.erdos_renyi_pm_j(200, 18, 10), 1,
     _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE, False, 6, False),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_parity_with_oracle(ctx, case):
    name, factory, cells, kernel, view, exchange, sweeps, use_sched = case
    print(f"[parity] {name}", flush=True)
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    # model build: CSR and colouring equal the oracle's ProblemIR restatement
    rp, ci, w = model.csr()
    np.testing.assert_array_equal(rp, csr.row_ptr)
    np.testing.assert_array_equal(ci, csr.col_idx[: csr.nnz])
    np.testing.assert_array_equal(w.astype(np.int64), csr.w[: csr.nnz])
    np.testing.assert_array_equal(model.coloring(), csr.coloring())
    # the sweep's position order equals the oracle's (pinned by test_oracle_order.py)
    order, cs = model.order()
    o_order, o_cs = csr.color_order()
    np.testing.assert_array_equal(order, o_order)
    np.testing.assert_array_equal(cs, o_cs)
    seed = 0x1234_5678_9ABC_DEF0 + len(name)
    st = _lib.State(model, cells, seed, kernel=kernel)
    ost = ob.OracleState(csr, cells, seed, view=view)
    compare(st, ost, csr, fields=False)
    ladder = instances.geometric_betas(3.0, 0.4, cells)
    if use_sched:
        sched = instances.anneal_schedule(0.1, 3.0, sweeps)
        st.run(sweeps, exchange=exchange, beta_schedule=sched)
        ost.run(sweeps, exchange=exchange, schedule=sched)
    else:
        st.set_ladder(ladder)
        half = sweeps // 2
        st.run(half, exchange=exchange)
        st.run(sweeps - half, exchange=exchange)  # resume across calls
        ost.run(sweeps, exchange=exchange, ladder=ladder)
    compare(st, ost, csr)


def test_isolated_sites_and_single_site(ctx):
    qb = instances.Qubo(n=5, offset=3, linear=np.array([4, -1, 0, 2, -7]),
                        pi=np.array([0], dtype=np.uint32), pj=np.array([3], dtype=np.uint32),
                        q=np.array([5]))
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    assert model.info.multispin_eligible == 0
    st = _lib.State(model, 2, 77)
    ost = ob.OracleState(csr, 2, 77)
    st.set_ladder([0.5, 1.5])
    st.run(9, exchange=True)
    ost.run(9, exchange=True, ladder=[0.5, 1.5])
    compare(st, ost, csr)
    one = instances.Qubo(n=1, offset=0, linear=np.array([3]), pi=np.zeros(0, np.uint32),
                         pj=np.zeros(0, np.uint32), q=np.zeros(0, np.int64))
    csr1 = ob.CSR.from_qubo(one)
    st1 = _lib.State(make(ctx, one), 1, 5)
    ost1 = ob.OracleState(csr1, 1, 5)
    st1.set_ladder([0.7])
    st1.run(5)
    ost1.run(5, exchange=False, ladder=[0.7])
    compare(st1, ost1, csr1)


def test_user_initial_words_and_auto_kernel(ctx):
    qb = instances.hypercubic_pm_j(4, 2, 11)
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    assert model.info.multispin_eligible == 1 and model.info.coupling_magnitude == 4
    rng = np.random.default_rng(1)
    words = rng.integers(0, 2 ** 63, size=(2, qb.n), dtype=np.uint64)
    st = _lib.State(model, 2, 9, init_words=words)  # AUTO -> multispin
    assert st.stats()[0].kernel == _lib.KERNEL_MULTISPIN
    ost = ob.OracleState(csr, 2, 9, words=words, view=ob.VIEW_PLANE)
    st.set_ladder([0.3, 0.9])
    st.run(15, exchange=True)
    ost.run(15, exchange=True, ladder=[0.3, 0.9])
    compare(st, ost, csr)


def test_full_size_3d_pt_properties(ctx):
    """3D +-J L=64, 32 temperatures: the PT workload at its real shape."""
    wl = instances.workload_3d_pt()
    qb = wl["qubo"]
    model = make(ctx, qb)
    assert model.info.num_colors == 2 and model.info.max_degree == 6
    st = _lib.State(model, wl["cells"], 2024)
    st.set_ladder(wl["ladder"])
    st.run(6, exchange=True)
    e = st.energies()
    np.testing.assert_array_equal(st.recompute_energies(), e)
    tof = st.temp_of()
    assert np.all(np.sort(tof, axis=0) == np.arange(32)[:, None])
    s, pa, pc = st.stats()
    assert s.proposed == 6 * qb.n * 64 * 32 and 0 < s.accepted < s.proposed
    # DEO: 3 even sweeps over pairs t = 0, 2, .., 30 and 3 odd over t = 1, .., 29
    assert np.all(pc <= pa) and pa.sum() == 64 * (3 * 16 + 3 * 15)
    assert np.all(pa[0::2] == 3 * 64) and np.all(pa[1::2] == 3 * 64)
    # Ising energy per spin lies in [-3, 3] on the cubic lattice.
    assert np.all(np.abs(e) <= 3 * qb.n)


def test_errors_are_reported(ctx):
    qb = instances.hypercubic_pm_j(4, 2, 1)
    with pytest.raises(_lib.IgeError) as ei:
        _lib.Model(ctx, qb.n, 0, qb.linear, qb.pj, qb.pi, qb.q)  # i > j
    assert ei.value.code == _lib.IGE_ERR_INVALID
    rq = instances.random_qubo(50, 4, 1)
    m = make(ctx, rq)
    with pytest.raises(_lib.IgeError) as ei:
        _lib.State(m, 1, 1, kernel=_lib.KERNEL_MULTISPIN)
    assert ei.value.code == _lib.IGE_ERR_UNSUPPORTED
    st = _lib.State(m, 2, 1)
    with pytest.raises(_lib.IgeError):
        st.run(3)  # no ladder yet
    with pytest.raises(_lib.IgeError):
        st.set_ladder([float("nan"), 1.0])
    big = instances.Qubo(n=2, offset=0, linear=np.array([2 ** 31 - 1, 0]),
                         pi=np.array([0], np.uint32), pj=np.array([1], np.uint32),
                         q=np.array([5]))
    with pytest.raises(_lib.IgeError) as ei:
        make(ctx, big)
    assert ei.value.code == _lib.IGE_ERR_RANGE


@pytest.mark.parametrize("kernel,view", [(_lib.KERNEL_MULTISPIN, ob.VIEW_PLANE),
                                         (_lib.KERNEL_GENERIC, ob.VIEW_WORD)])
def test_exchange_only_rounds_interleaved(ctx, kernel, view):
    """ReplicaExchange::apply as its own operator (replica_exchange.rs:61-109): exchange-only
    rounds between sweeps continue the same DEO phase and keys as fused exchanges."""
    qb = instances.hypercubic_pm_j(4, 3, 21)
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    ladder = instances.geometric_betas(2.5, 0.3, 6)
    st = _lib.State(model, 6, 31, kernel=kernel)
    ost = ob.OracleState(csr, 6, 31, view=view)
    st.set_ladder(ladder)
    for sweeps, rounds, fused in [(3, 2, True), (2, 3, False), (4, 1, True), (1, 5, False)]:
        st.run(sweeps, exchange=fused)
        ost.run(sweeps, exchange=fused, ladder=ladder)
        st.exchange(rounds)
        ost.exchange(rounds, ladder)
        compare(st, ost, csr, fields=False)
    s = st.stats()[0]
    assert s.exchanges == ost.xround and s.phase == ost.xround & 1
