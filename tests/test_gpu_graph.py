"""The captured sweep loop (ige_run replays a hipGraph of 32 steps whose counters come from a
device step clock) against eager launches of the same steps, bit for bit after every call:
words, int64 energies, temperature permutation, flip counts, exchange statistics and the
incumbent.  Call lengths mix whole replays, remainders and calls too short to replay, and the
flags change between calls so the graph is captured again.  One case also runs against the
oracle directly."""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu

CALLS = [40, 64, 5, 33]


def make(ctx, qb):
    return _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)


def same(a, b, track):
    np.testing.assert_array_equal(a.spins(), b.spins())
    np.testing.assert_array_equal(a.energies(), b.energies())
    np.testing.assert_array_equal(a.temp_of(), b.temp_of())
    sa, pa, ca = a.stats()
    sb, pb, cb = b.stats()
    assert (sa.accepted, sa.proposed) == (sb.accepted, sb.proposed)
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(ca, cb)
    if track:
        xa, ea, wa = a.best()
        xb, eb, wb = b.best()
        assert (ea, wa) == (eb, wb)
        np.testing.assert_array_equal(xa, xb)


CASES = [
    # name, factory, kernel, cells, pops, exchange, schedule, heat-bath, track
    ("pmJ_2d_L16_multispin_pt_track", lambda: instances.hypercubic_pm_j(16, 2, 71),
     _lib.KERNEL_MULTISPIN, 8, 1, True, False, False, True),
    ("pmJ_3d_L8_multispin_two_populations", lambda: instances.hypercubic_pm_j(8, 3, 72),
     _lib.KERNEL_MULTISPIN, 8, 2, True, False, False, False),
    ("er_pmJ_n2000_multispin_anneal", lambda: instances.erdos_renyi_pm_j(2000, 6, 73),
     _lib.KERNEL_AUTO, 1, 1, False, True, False, True),
    ("qubo_n300_generic_pt_schedule", lambda: instances.random_qubo(300, 6, 74, offdiag=30,
                                                                     diag=50),
     _lib.KERNEL_GENERIC, 4, 1, True, True, False, False),
    ("qubo_n300_heatbath_pt_track", lambda: instances.random_qubo(300, 6, 75, offdiag=30,
                                                                   diag=50),
     _lib.KERNEL_GENERIC, 5, 1, True, False, True, True),
    ("pmJ_3d_L6_heatbath_anneal", lambda: instances.hypercubic_pm_j(6, 3, 76),
     _lib.KERNEL_MULTISPIN, 2, 1, False, True, True, False),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_graph_replays_equal_eager_launches(ctx, case):
    name, factory, kernel, cells, pops, exchange, use_sched, heat, track = case
    print(f"[graph] {name}", flush=True)
    qb = factory()
    model = make(ctx, qb)
    seed = 0x6EA9_0000 + len(name)
    g = _lib.State(model, cells, seed, kernel=kernel)
    e = _lib.State(model, cells, seed, kernel=kernel)
    t_max, t_min = instances.suggested_temp_range(qb)
    if pops > 1:
        g.set_populations(pops)
        e.set_populations(pops)
    if not use_sched:
        ladder = np.tile(instances.geometric_betas(t_max, t_min, cells // pops), pops)
        g.set_ladder(ladder)
        e.set_ladder(ladder)
    sched = np.repeat(instances.anneal_schedule(1.0 / t_max, 1.0 / t_min, sum(CALLS)), cells,
                      axis=1)
    done = 0
    for k, sweeps in enumerate(CALLS):
        kw = dict(exchange=exchange and k != 2, track_best=track, heatbath=heat)
        if use_sched:
            kw["beta_schedule"] = sched[done:done + sweeps]
        g.run(sweeps, **kw)
        e.run(sweeps, eager=True, **kw)
        done += sweeps
        same(g, e, track)
    np.testing.assert_array_equal(g.recompute_energies(), g.energies())
    s = g.stats()[0]
    assert 0 < s.accepted < s.proposed
    # the replays really ran: 32 + 64 + 0 + 32 sweeps of the 142 came from the graph
    assert (s.graph_sweeps, s.graph_fallbacks) == (128, 0)
    assert e.stats()[0].graph_sweeps == 0


def test_graph_replays_match_the_oracle(ctx):
    qb = instances.hypercubic_pm_j(12, 2, 77)
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    cells, seed = 6, 0x6EA9_1000
    st = _lib.State(model, cells, seed, kernel=_lib.KERNEL_MULTISPIN)
    ost = ob.OracleState(csr, cells, seed, view=ob.VIEW_PLANE)
    ladder = instances.geometric_betas(2.0, 0.5, cells)
    st.set_ladder(ladder)
    for sweeps in (70, 32):
        st.run(sweeps, exchange=True)
        ost.run(sweeps, exchange=True, ladder=ladder)
        np.testing.assert_array_equal(st.spins(), ost.words)
        np.testing.assert_array_equal(st.energies(), ost.energy)
        np.testing.assert_array_equal(st.temp_of(), ost.temp_of)
        s, pa, pc = st.stats()
        assert s.accepted == int(ost.stats[0]) and s.proposed == int(ost.stats[1])
        np.testing.assert_array_equal(pa, ost.stats[2::2])
        np.testing.assert_array_equal(pc, ost.stats[3::2])
    assert st.stats()[0].graph_sweeps == 64 + 32
