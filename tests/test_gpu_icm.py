"""Isoenergetic cluster moves on the device against the oracle, bit-exact: words, tracked
energies, the temperature permutation and the total cluster size; and at the 3D L=64 x 32
bench shape (several launches per ICM sweep) through size-independent checks."""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def make(ctx, qb):
    return _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)


CASES = [
    ("qubo_n200_generic_pt", lambda: instances.random_qubo(200, 3, 71, offdiag=9, diag=12),
     _lib.KERNEL_GENERIC, ob.VIEW_WORD, 3),
    ("pmJ_3d_L6_multispin_pt", lambda: instances.hypercubic_pm_j(6, 3, 72),
     _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE, 4),
    ("sk_n96_dense_pt", lambda: instances.sk_int8(96, 73), _lib.KERNEL_DENSE, ob.VIEW_WORD, 2),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_icm_parity_with_oracle(ctx, case):
    name, factory, kernel, view, cells = case
    print(f"[icm] {name}", flush=True)
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    seed = 0x1C3_0000 + len(name)
    st = _lib.State(model, cells, seed, kernel=kernel)
    ost = ob.OracleState(csr, cells, seed, view=view)
    t_max, t_min = instances.suggested_temp_range(qb)
    ladder = instances.geometric_betas(t_max, t_min, cells)
    st.set_ladder(ladder)
    # per-sweep ICM through the run flag, then explicit rounds
    st.run(6, exchange=True, icm=True)
    for _ in range(6):
        ost.run(1, exchange=True, ladder=ladder)
        ost.icm(1)
    got = st.icm(2)
    sizes = [ost.icm(1).sum() for _ in range(2)]
    assert got == sum(int(v) for v in sizes) > 0
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    np.testing.assert_array_equal(st.temp_of(), ost.temp_of)
    np.testing.assert_array_equal(st.recompute_energies(), ost.energy)


def test_icm_full_size_3d_L64_pt32(ctx):
    """The bench state (n = 262144, 32 temperatures): an ICM sweep needs several launches
    (the move scratch is capped); tracked energies equal a recompute, the temperature
    permutation is intact and every pair's energy sum is conserved."""
    wl = instances.workload_3d_pt()
    qb = wl["qubo"]
    model = make(ctx, qb)
    st = _lib.State(model, wl["cells"], 0x5EED)
    st.set_ladder(wl["ladder"])
    st.run(20, exchange=True)
    e0 = st.energies().copy()
    tof = st.temp_of()
    flipped = st.icm(1)
    assert flipped > 0
    e1 = st.energies()
    np.testing.assert_array_equal(st.recompute_energies(), e1)
    np.testing.assert_array_equal(st.temp_of(), tof)
    so = np.zeros_like(tof)
    for s in range(tof.shape[0]):
        so[tof[s], np.arange(64)] = s
    for t in range(tof.shape[0]):
        for r in range(32):
            a, b = (so[t, r], r), (so[t, r + 32], r + 32)
            assert e0[a] + e0[b] == e1[a] + e1[b]
    st.run(5, exchange=True, icm=True)
    np.testing.assert_array_equal(st.recompute_energies(), st.energies())
