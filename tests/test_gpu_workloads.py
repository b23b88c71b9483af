"""The named workloads on the GPU at their real sizes.

The 2D anneal (+-J L=32, 64 replicas, beta 0.1 -> 3.0 over 1000 sweeps) is
small enough for the oracle, so it is checked bit-exactly end to end.
"""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel,view", [(_lib.KERNEL_MULTISPIN, ob.VIEW_PLANE),
                                         (_lib.KERNEL_GENERIC, ob.VIEW_WORD)])
def test_2d_anneal_full_size_bit_exact(ctx, kernel, view):
    wl = instances.workload_2d_ej()
    qb, sched = wl["qubo"], wl["schedule"]
    print(f"[workload] {wl['name']} kernel={kernel}", flush=True)
    csr = ob.CSR.from_qubo(qb)
    model = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    st = _lib.State(model, 1, 0xC0FFEE, kernel=kernel)
    st.run(wl["sweeps"], beta_schedule=sched)
    ost = ob.OracleState(csr, 1, 0xC0FFEE, view=view)
    ost.run(wl["sweeps"], exchange=False, schedule=sched)
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    np.testing.assert_array_equal(st.recompute_energies(), ost.energy)
    s = st.stats()[0]
    assert s.accepted == int(ost.stats[0]) and s.proposed == 1000 * 1024 * 64
    # annealed to beta = 3: energy per spin near the 2D +-J ground state (~ -1.40)
    assert ost.energy.mean() / qb.n < -1.3
