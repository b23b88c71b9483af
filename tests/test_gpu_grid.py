"""Grid sizing of the multispin colour-class launch on large regular lattices, against the
oracle bit for bit.

When the default grid (about 8192 waves of up to 8 sites per lane) would need more than one
round of resident workgroups and its row of workgroups per slot is not a multiple of 8, the
launch gives each wave a contiguous span of up to 640 positions instead, one round of
resident waves.  The sites of a colour class are independent and their uniforms depend only
on (position, slot, sweep), so the split must not change any bit.  The sizes here take that
path on a 256-CU GPU with 7 resident workgroups per CU:
- 3D +-J L=78, 2 slots: classes of 237276 positions, default rows of 927 workgroups, one
  round of 3584 waves per slot with spans of 67 positions (2 per lane);
- 2D +-J L=1018, 1 slot: classes of 518162 positions, default row of 2025 workgroups, one
  round of 7099 waves with spans of 73 positions.
"""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,L,dim,cells,exchange", [
    ("pmJ_3d_L78_two_slots_pt", 78, 3, 2, True),
    ("pmJ_2d_L1018_one_slot", 1018, 2, 1, False),
])
def test_large_class_grid_matches_oracle(ctx, name, L, dim, cells, exchange):
    print(f"[grid] {name}", flush=True)
    qb = instances.hypercubic_pm_j(L, dim, 91)
    csr = ob.CSR.from_qubo(qb)
    model = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    assert model.info.num_colors == 2
    seed = 0x6E1D_0000 + L
    st = _lib.State(model, cells, seed, kernel=_lib.KERNEL_MULTISPIN)
    ost = ob.OracleState(csr, cells, seed, view=ob.VIEW_PLANE)
    ladder = instances.geometric_betas(2.0, 0.5, cells) if cells > 1 else np.array([0.7])
    st.set_ladder(ladder)
    st.run(2, exchange=exchange)
    ost.run(2, exchange=exchange, ladder=ladder)
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    np.testing.assert_array_equal(st.temp_of(), ost.temp_of)
    s = st.stats()[0]
    assert s.accepted == int(ost.stats[0]) and s.proposed == int(ost.stats[1])
    np.testing.assert_array_equal(st.recompute_energies(), st.energies())
