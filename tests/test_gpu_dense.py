"""Dense couplings (SK, dense QUBO): the matrix-core field recompute and the
sequential dense sweep against the CPU oracle, bit-exact.

The dense kernel visits positions one at a time in colour-major order with the
word-view uniforms, so its trajectory equals the oracle's (and the generic
kernel's) exactly.
"""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def make(ctx, qb):
    return _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)


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


def sk_betas(qb, k):
    t_max, t_min = instances.suggested_temp_range(qb)
    return instances.geometric_betas(t_max, t_min, k)


CASES = [
    # name, factory, cells, exchange, sweeps, schedule?
    ("sk_n300_anneal", lambda: instances.sk_int8(300, 1), 1, False, 12, True),
    ("sk_n64_one_block_pt", lambda: instances.sk_int8(64, 2), 3, True, 10, False),
    ("sk_n65_partial_block_pt", lambda: instances.sk_int8(65, 3), 2, True, 10, False),
    ("sk_n1000_pt", lambda: instances.sk_int8(1000, 4), 4, True, 4, False),
    ("dense_qubo_n200_scale3_pt", lambda: instances.dense_qubo(200, 5, scale=3), 3, True, 10,
     False),
    # scale 7: block corrections exceed int16, the int32 update path runs
    ("dense_qubo_n130_scale7_pt", lambda: instances.dense_qubo(130, 8, scale=7), 2, True, 10,
     False),
    ("sparse_qubo_n150_forced_dense", lambda: instances.random_qubo(150, 6, 6), 2, True, 10,
     False),
    ("pmJ_2d_L8_forced_dense", lambda: instances.hypercubic_pm_j(8, 2, 7), 2, True, 10, False),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_dense_parity_with_oracle(ctx, case):
    name, factory, cells, exchange, sweeps, use_sched = case
    print(f"[dense] {name}", flush=True)
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    assert model.info.dense_scale > 0
    seed = 0xD15E_0000_0000_0001 + len(name)
    st = _lib.State(model, cells, seed, kernel=_lib.KERNEL_DENSE)
    assert st.stats()[0].kernel == _lib.KERNEL_DENSE
    ost = ob.OracleState(csr, cells, seed, view=ob.VIEW_WORD)
    compare(st, ost, csr)
    if use_sched:
        t_max, t_min = instances.suggested_temp_range(qb)
        sched = instances.anneal_schedule(1.0 / t_max, 1.0 / t_min, sweeps)
        st.run(sweeps, exchange=exchange, beta_schedule=sched)
        ost.run(sweeps, exchange=exchange, schedule=sched)
    else:
        ladder = sk_betas(qb, cells) if qb.name.startswith("SK") else \
            instances.geometric_betas(3.0, 0.3, cells) / max(1, int(np.abs(qb.q).max()) // 8)
        st.set_ladder(ladder)
        half = sweeps // 2
        st.run(half, exchange=exchange)
        st.run(sweeps - half, exchange=exchange)
        ost.run(sweeps, exchange=exchange, ladder=ladder)
    compare(st, ost, csr)
    s = st.stats()[0]
    assert 0 < s.accepted < s.proposed


@pytest.mark.parametrize("factory", [lambda: instances.sk_int8(200, 11),
                                     lambda: instances.random_qubo(130, 7, 12)],
                         ids=["sk_n200_scale4", "qubo_n130_scale1"])
def test_dense_band_fallback_and_accept_bounds(ctx, factory):
    """A ladder holding a beta below the accept bounds' minimum runs the whole call on the
    in-chain threshold band; the same ladder without it runs on accept bounds.  Both equal
    the oracle, and the two kernels continue each other's state bit for bit."""
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    base = sk_betas(qb, 3) if qb.name.startswith("SK") else \
        instances.geometric_betas(3.0, 0.3, 3) / max(1, int(np.abs(qb.q).max()) // 8)
    cold = np.array(base, dtype=np.float64)
    hot = cold.copy()
    hot[0] = 5e-7  # below the accept bounds' minimum beta (1e-6): the band kernel runs
    seed = 0xBA4D_0000_0000_0003
    st = _lib.State(model, 3, seed, kernel=_lib.KERNEL_DENSE)
    ost = ob.OracleState(csr, 3, seed, view=ob.VIEW_WORD)
    for ladder in (hot, cold, hot, cold):
        print(f"[dense] {qb.name} ladder beta0={ladder[0]:.3g}", flush=True)
        st.set_ladder(ladder)
        st.run(3, exchange=True)
        ost.run(3, exchange=True, ladder=ladder)
        compare(st, ost, csr, fields=False)
    compare(st, ost, csr)


@pytest.mark.parametrize("cells", [1, 3], ids=["one_slot", "three_slots"])
def test_dense_field_helpers(ctx, cells):
    """N=1600 (2048 padded columns, 25 blocks): each slot's gang of the accept-bound sweep
    shares every block's fields with three helper workgroups on its XCD (the smaller models
    above keep them in one workgroup).  Equal to the oracle after exchange rounds and after
    an annealing schedule."""
    qb = instances.sk_int8(1600, 21)
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    seed = 0x6A46_0000_0000_0005 + cells
    st = _lib.State(model, cells, seed, kernel=_lib.KERNEL_DENSE)
    ost = ob.OracleState(csr, cells, seed, view=ob.VIEW_WORD)
    ladder = sk_betas(qb, cells)
    assert ladder.min() >= 1e-6  # the accept-bound kernel runs
    print(f"[dense] helpers n={qb.n} cells={cells}", flush=True)
    st.set_ladder(ladder)
    st.run(2, exchange=True)
    ost.run(2, exchange=True, ladder=ladder)
    compare(st, ost, csr, fields=False)
    t_max, t_min = instances.suggested_temp_range(qb)
    sched = np.repeat(instances.anneal_schedule(1.0 / t_max, 1.0 / t_min, 3), cells, axis=1)
    st.run(3, beta_schedule=sched)
    ost.run(3, exchange=False, schedule=sched)
    compare(st, ost, csr)


def test_dense_equals_generic_kernel(ctx):
    """Same seed, same schedule: the dense and the lane-per-replica kernels walk the same
    trajectory (one position at a time vs a colour class at a time)."""
    qb = instances.sk_int8(500, 9)
    model = make(ctx, qb)
    t_max, t_min = instances.suggested_temp_range(qb)
    sched = instances.anneal_schedule(1.0 / t_max, 1.0 / t_min, 20)
    out = []
    for kernel in (_lib.KERNEL_DENSE, _lib.KERNEL_GENERIC):
        st = _lib.State(model, 1, 4242, kernel=kernel)
        st.run(20, beta_schedule=sched)
        out.append((st.spins(), st.energies(), st.stats()[0].accepted))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_full_size_sk_4096(ctx):
    """SK N=4096 at its real shape: AUTO chooses the dense kernel; one sweep equals the
    oracle; the matrix-core fields equal the CSR fields; after an anneal segment the tracked
    energies equal a recompute."""
    wl = instances.workload_sk_anneal()
    qb = wl["qubo"]
    model = make(ctx, qb)
    assert model.info.dense_scale == 4
    st = _lib.State(model, 1, 77)
    assert st.stats()[0].kernel == _lib.KERNEL_DENSE
    csr = ob.CSR.from_qubo(qb)
    ost = ob.OracleState(csr, 1, 77, view=ob.VIEW_WORD)
    beta = [min(wl["beta_hi"], 1e-3)]
    st.set_ladder(beta)
    st.run(1)
    ost.run(1, exchange=False, ladder=beta)
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    np.testing.assert_array_equal(st.fields().astype(np.int64), csr.fields(ost.words))
    # at the hot end of the schedule nearly every uphill dE has a nonzero threshold, so the
    # band decides far more proposals: two more sweeps there, still equal to the oracle
    hot = [wl["beta_lo"]]
    st.set_ladder(hot)
    st.run(2)
    ost.run(2, exchange=False, ladder=hot)
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    sched = instances.anneal_schedule(wl["beta_lo"], wl["beta_hi"], 40)
    st.run(40, beta_schedule=sched)
    np.testing.assert_array_equal(st.recompute_energies(), st.energies())
    s = st.stats()[0]
    assert s.proposed == 43 * qb.n * 64 and 0 < s.accepted < s.proposed


def test_dense_eligibility(ctx):
    """q that is not scale * int8 cannot take the dense kernel; AUTO then picks generic."""
    qb = instances.random_qubo(120, 100, 3, offdiag=1000)
    model = make(ctx, qb)
    assert model.info.dense_scale == 0
    with pytest.raises(_lib.IgeError) as ei:
        _lib.State(model, 1, 1, kernel=_lib.KERNEL_DENSE)
    assert ei.value.code == _lib.IGE_ERR_UNSUPPORTED
    assert _lib.State(model, 1, 1).stats()[0].kernel == _lib.KERNEL_GENERIC
    sparse = make(ctx, instances.random_qubo(300, 6, 4))
    assert sparse.info.dense_scale == 1  # eligible, but sparse: AUTO stays generic
    assert _lib.State(sparse, 1, 1).stats()[0].kernel == _lib.KERNEL_GENERIC
