"""Resident runs (a small regular lattice's whole ige_run call in one workgroup per slot, words
in LDS), chained runs (a one-slot irregular state's whole call in one launch, a device-wide
barrier per class) and fused small colour classes (trailing classes of a sweep in one launch)
against the
colour-class launches of the same sweeps (IGE_RUN_EAGER), bit for bit after every call: words,
int64 energies, temperatures, flip counts.  Covers both degrees and irregular graphs, schedules
and fixed ladders, several slots, the 4096-site limit, calls mixed with exchange runs, and runs
against the oracle directly."""
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


@pytest.mark.parametrize("name,factory,cells,exchange", [
    ("er_pmJ_n20000_tail_of_small_classes", lambda: instances.erdos_renyi_pm_j(20000, 6, 86), 1,
     False),
    ("er_pmJ_n1500_tail_fused_pt", lambda: instances.erdos_renyi_pm_j(1500, 6, 87), 4, True),
    ("pmJ_2d_L16_both_classes_fused_pt", lambda: instances.hypercubic_pm_j(16, 2, 88), 4, True),
])
def test_fused_small_classes_equal_per_class_launches(ctx, name, factory, cells, exchange):
    """Trailing colour classes of at most 512 sites each run in one launch (one workgroup
    per slot, class after class); IGE_RUN_EAGER launches each class alone."""
    print(f"[fused] {name}", flush=True)
    qb = factory()
    model = make(ctx, qb)
    seed = 0x5E51_2000 + len(name)
    f = _lib.State(model, cells, seed, kernel=_lib.KERNEL_MULTISPIN)
    e = _lib.State(model, cells, seed, kernel=_lib.KERNEL_MULTISPIN)
    t_max, t_min = instances.suggested_temp_range(qb)
    ladder = instances.geometric_betas(t_max, t_min, cells)
    f.set_ladder(ladder)
    e.set_ladder(ladder)
    sched = np.repeat(instances.anneal_schedule(1.0 / t_max, 1.0 / t_min, 45), cells, axis=1)
    for k0, k1 in ((0, 7), (7, 45)):
        f.run(k1 - k0, exchange=exchange, beta_schedule=sched[k0:k1])
        e.run(k1 - k0, exchange=exchange, beta_schedule=sched[k0:k1], eager=True)
        same(f, e)
    np.testing.assert_array_equal(f.recompute_energies(), f.energies())


def test_fused_tail_matches_the_oracle(ctx):
    qb = instances.erdos_renyi_pm_j(20000, 6, 89)
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    st = _lib.State(model, 1, 0x5E51_3000, kernel=_lib.KERNEL_MULTISPIN)
    ost = ob.OracleState(csr, 1, 0x5E51_3000, view=ob.VIEW_PLANE)
    t_max, t_min = instances.suggested_temp_range(qb)
    sched = instances.anneal_schedule(1.0 / t_max, 1.0 / t_min, 6)
    for k in range(6):          # one sweep per call: the colour-class launches and fused tail
        st.run(1, beta_schedule=sched[k:k + 1])
    ost.run(6, exchange=False, schedule=sched)
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    s = st.stats()[0]
    assert s.accepted == int(ost.stats[0])
    assert s.chain_sweeps == 0


@pytest.mark.parametrize("barriers", [False, True], ids=["site_waits", "class_barriers"])
@pytest.mark.parametrize("n,degree", [(20000, 6), (3000, 1), (3000, 5), (3000, 12)],
                         ids=["er20000_d6", "er3000_d1", "er3000_d5", "er3000_d12"])
def test_chained_sweeps_match_the_oracle(ctx, n, degree, barriers, monkeypatch):
    # the whole call in one launch (degree < 8, < 16 and < 32 counter widths), each site
    # waiting for its neighbours' sweep counts, or every class behind a gang barrier:
    # against the oracle on a schedule, then on a fixed ladder, and against the same
    # sweeps run one call each
    if barriers:
        monkeypatch.setenv("IGE_CHAIN_BARRIERS", "1")
    else:
        monkeypatch.delenv("IGE_CHAIN_BARRIERS", raising=False)
    qb = instances.erdos_renyi_pm_j(n, degree, 90 + degree)
    dmax = int(np.bincount(np.concatenate([qb.pi, qb.pj]), minlength=n).max())
    assert dmax < 32
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    print(f"[chain] n={n} degree={degree} dmax={model.info.max_degree} "
          f"colours={model.info.num_colors} barriers={barriers}", flush=True)
    seed = 0xC4A1 + n
    st = _lib.State(model, 1, seed, kernel=_lib.KERNEL_MULTISPIN)
    ost = ob.OracleState(csr, 1, seed, view=ob.VIEW_PLANE)
    t_max, t_min = instances.suggested_temp_range(qb)
    sched = instances.anneal_schedule(1.0 / t_max, 1.0 / t_min, 12)
    st.run(12, beta_schedule=sched)
    ost.run(12, exchange=False, schedule=sched)
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    assert st.stats()[0].chain_sweeps == 12
    assert st.stats()[0].accepted == int(ost.stats[0])
    beta = np.array([1.0 / (0.5 * (t_max + t_min))])
    st.set_ladder(beta)
    st.run(9)
    ost.run(9, exchange=False, ladder=beta)
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    np.testing.assert_array_equal(st.recompute_energies(), st.energies())
    one = _lib.State(model, 1, seed, kernel=_lib.KERNEL_MULTISPIN)
    for k in range(12):
        one.run(1, beta_schedule=sched[k:k + 1])
    one.set_ladder(beta)
    for _ in range(9):
        one.run(1)
    same(st, one)
    assert one.stats()[0].chain_sweeps == 0


@pytest.mark.parametrize("barriers", [False, True], ids=["site_waits", "class_barriers"])
def test_chained_wait_expiry_is_reported(ctx, barriers, monkeypatch):
    # a spin limit of 0 polls makes the first wait that is not already satisfied give up:
    # the launch still ends, the state reports the expired wait when it is next
    # synchronised (a run on a fixed ladder returns before its launch ends), and the next
    # call runs
    if barriers:
        monkeypatch.setenv("IGE_CHAIN_BARRIERS", "1")
    else:
        monkeypatch.delenv("IGE_CHAIN_BARRIERS", raising=False)
    qb = instances.erdos_renyi_pm_j(20000, 6, 96)
    model = make(ctx, qb)
    st = _lib.State(model, 1, 0xE7, kernel=_lib.KERNEL_MULTISPIN)
    st.set_ladder(np.array([0.7]))
    print(f"[chain-expiry] barriers={barriers}", flush=True)
    monkeypatch.setenv("IGE_CHAIN_SPIN_LIMIT", "0")
    with pytest.raises(_lib.IgeError, match="spin limit"):
        st.run(50)
        st.synchronize()
    monkeypatch.delenv("IGE_CHAIN_SPIN_LIMIT")
    before = st.stats()[0].chain_sweeps
    st.run(10)
    st.synchronize()
    assert st.stats()[0].chain_sweeps == before + 10
