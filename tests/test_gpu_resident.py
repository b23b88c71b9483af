"""Resident runs (a small regular lattice's whole ige_run call in one workgroup per slot, words
in LDS) against the colour-class launches of the same sweeps (IGE_RUN_EAGER), bit for bit after
every call: words, int64 energies, temperatures, flip counts.  Covers both degrees, schedules and
fixed ladders, several slots, the 4096-site limit, calls mixed with exchange runs (which take the
other paths), and one run against the oracle directly."""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def make(ctx, qb):
    return _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)


def same(a, b):
    np.testing.assert_array_equal(a.spins(), b.spins())
    np.testing.assert_array_equal(a.energies(), b.energies())
    np.testing.assert_array_equal(a.temp_of(), b.temp_of())
    sa, pa, ca = a.stats()
    sb, pb, cb = b.stats()
    assert (sa.accepted, sa.proposed, sa.sweeps) == (sb.accepted, sb.proposed, sb.sweeps)
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(ca, cb)


CASES = [
    # name, factory, cells, schedule
    ("pmJ_2d_L16_anneal", lambda: instances.hypercubic_pm_j(16, 2, 81), 1, True),
    ("pmJ_2d_L32_two_slots_ladder", lambda: instances.hypercubic_pm_j(32, 2, 82), 2, False),
    ("pmJ_3d_L8_four_slots_anneal", lambda: instances.hypercubic_pm_j(8, 3, 83), 4, True),
    ("pmJ_3d_L16_at_the_size_limit", lambda: instances.hypercubic_pm_j(16, 3, 84), 2, False),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_resident_runs_equal_colour_launches(ctx, case):
    name, factory, cells, use_sched = case
    print(f"[resident] {name}", flush=True)
    qb = factory()
    model = make(ctx, qb)
    seed = 0x5E51_0000 + len(name)
    r = _lib.State(model, cells, seed, kernel=_lib.KERNEL_MULTISPIN)
    e = _lib.State(model, cells, seed, kernel=_lib.KERNEL_MULTISPIN)
    t_max, t_min = instances.suggested_temp_range(qb)
    ladder = instances.geometric_betas(t_max, t_min, cells)
    r.set_ladder(ladder)
    e.set_ladder(ladder)
    calls = [40, 5, 33, 1]
    sched = np.repeat(instances.anneal_schedule(1.0 / t_max, 1.0 / t_min, sum(calls)), cells,
                      axis=1)
    done = 0
    for sweeps in calls:
        kw = {"beta_schedule": sched[done:done + sweeps]} if use_sched else {}
        r.run(sweeps, **kw)
        e.run(sweeps, eager=True, **kw)
        done += sweeps
        same(r, e)
    # an exchange run in between takes the graph path; resident runs pick up after it
    if cells > 1:
        r.run(33, exchange=True)
        e.run(33, exchange=True, eager=True)
        same(r, e)
        r.run(20)
        e.run(20, eager=True)
        same(r, e)
    np.testing.assert_array_equal(r.recompute_energies(), r.energies())
    s = r.stats()[0]
    assert s.resident_sweeps == 78 + (20 if cells > 1 else 0)
    assert e.stats()[0].resident_sweeps == 0
    assert 0 < s.accepted < s.proposed


def test_resident_run_matches_the_oracle(ctx):
    qb = instances.hypercubic_pm_j(12, 2, 85)
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    cells, seed = 3, 0x5E51_1000
    st = _lib.State(model, cells, seed, kernel=_lib.KERNEL_MULTISPIN)
    ost = ob.OracleState(csr, cells, seed, view=ob.VIEW_PLANE)
    ladder = instances.geometric_betas(2.0, 0.5, cells)
    st.set_ladder(ladder)
    for sweeps in (50, 17):
        st.run(sweeps)
        ost.run(sweeps, exchange=False, ladder=ladder)
        np.testing.assert_array_equal(st.spins(), ost.words)
        np.testing.assert_array_equal(st.energies(), ost.energy)
        s = st.stats()[0]
        assert s.accepted == int(ost.stats[0]) and s.proposed == int(ost.stats[1])
    assert st.stats()[0].resident_sweeps == 67
