"""The replica view on the GPU against the oracle, bit for bit: the engine_v2 operators'
own semantics, MetropolisSweep with T = temps[rep % temps.len()] (metropolis_sweep.rs:115)
and ReplicaExchange pairing replica k with k + 1 from start = round & 1 of each apply
(replica_exchange.rs:61-109), in the shapes of the reference's operator tests
(metropolis_sweep.rs:174-214, replica_exchange.rs:138-196) and with the reference's own
ladder, geometric_ladder(64, hi, lo) (decision.rs:133-142)."""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob
from test_gpu_parity import compare, frustrated7, make

pytestmark = pytest.mark.gpu


def views():
    return [
        ("frustrated7_generic", frustrated7, _lib.KERNEL_GENERIC, ob.VIEW_WORD),
        ("frustrated7_dense", frustrated7, _lib.KERNEL_DENSE, ob.VIEW_WORD),
        ("square_L8_multispin", lambda: instances.hypercubic_pm_j(8, 2, seed=5),
         _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE),
        ("cubic_L6_multispin", lambda: instances.hypercubic_pm_j(6, 3, seed=6),
         _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE),
        ("random_qubo_generic", lambda: instances.random_qubo(90, 5, seed=7),
         _lib.KERNEL_GENERIC, ob.VIEW_WORD),
    ]


@pytest.mark.parametrize("case", views(), ids=[c[0] for c in views()])
def test_metropolis_with_one_temperature_per_replica(ctx, case):
    # metropolis_sweep.rs:174-214 in shape: temperatures 0.2 + 0.05 k, one per replica
    # (here 128 replicas in two slots), 30 sweeps; words, energies, counts equal.
    name, factory, kernel, view = case
    print(f"[replica-view] sweep {name}", flush=True)
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    st = _lib.State(make(ctx, qb), 2, 0xACE + len(name), kernel=kernel)
    ost = ob.OracleState(csr, 2, 0xACE + len(name), view=view)
    temps = 0.2 + 0.05 * np.arange(128)
    st.set_replica_temperatures(temps)
    ost.set_replica_temperatures(temps)
    st.run(30)
    ost.run(30, exchange=False)
    compare(st, ost, csr, fields=False)


@pytest.mark.parametrize("case", views()[:3], ids=[c[0] for c in views()[:3]])
def test_replica_exchange_preserves_the_energy_multiset(ctx, case):
    # replica_exchange.rs:138-196 in shape: 128 replicas at 0.1 + 0.3 k (16 temperatures,
    # repeating), diverged by 5 heat-bath sweeps, then exchange rounds in two applies.
    name, factory, kernel, view = case
    print(f"[replica-view] exchange {name}", flush=True)
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    seed = 0xBEE + len(name)
    st = _lib.State(make(ctx, qb), 2, seed, kernel=kernel)
    ost = ob.OracleState(csr, 2, seed, view=ob.VIEW_HEATBATH)
    temps = 0.1 + 0.3 * np.arange(16)
    st.set_replica_temperatures(temps)
    ost.set_replica_temperatures(temps)
    st.run(5, heatbath=True)
    ost.run(5, exchange=False)
    before = np.sort(st.energies().reshape(-1))
    st.exchange_replicas(6)
    st.exchange_replicas(3)
    ost.exchange_replicas(6)
    ost.exchange_replicas(3)
    compare(st, ost, csr, fields=False)
    assert np.array_equal(np.sort(st.energies().reshape(-1)), before)
    perm = st.temp_of().reshape(-1)
    assert np.array_equal(np.sort(perm), np.arange(128))
    _, pa, pc = st.stats()
    assert pa.size == 127 and 0 < int(pc.sum()) < int(pa.sum())


def test_pt_loop_on_the_reference_geometric_ladder(ctx):
    # A Runtime-style loop: MetropolisSweep(1) then ReplicaExchange(1) per step on
    # geometric_ladder(64, hi, lo), 128 replicas (temperatures repeat per slot), with the
    # incumbent (ties by replica index) checked against the oracle after every step.
    qb = instances.hypercubic_pm_j(8, 2, seed=9)
    csr = ob.CSR.from_qubo(qb)
    temps = ob.geometric_ladder(64, 3.0, 0.3)
    for kernel, view in ((_lib.KERNEL_MULTISPIN, ob.VIEW_PLANE),
                         (_lib.KERNEL_GENERIC, ob.VIEW_WORD)):
        print(f"[replica-view] pt loop kernel {kernel}", flush=True)
        st = _lib.State(make(ctx, qb), 2, 31, kernel=kernel)
        ost = ob.OracleState(csr, 2, 31, view=view)
        inc = ob.Incumbent(csr.n)
        st.set_replica_temperatures(temps)
        ost.set_replica_temperatures(temps)
        for step in range(12):
            st.run(1, track_best=True)
            ost.run(1, exchange=False)
            inc.update(ost.words, ost.temp_of, ost.energy, replica_view=True)
            st.exchange_replicas(1)
            ost.exchange_replicas(1)
        compare(st, ost, csr, fields=False)
        x, e, _ = st.best()
        assert e == inc.energy.value
        np.testing.assert_array_equal(x, inc.x)
        # a long run through the captured sweep loop stays on the oracle's trajectory
        st.run(40, track_best=True)
        for _ in range(40):
            ost.run(1, exchange=False)
            inc.update(ost.words, ost.temp_of, ost.energy, replica_view=True)
        compare(st, ost, csr, fields=False)
        assert st.best()[1] == inc.energy.value


def test_switching_views(ctx):
    qb = frustrated7()
    m = make(ctx, qb)
    st = _lib.State(m, 2, 5, kernel=_lib.KERNEL_GENERIC)
    st.set_ladder([0.5, 1.5])
    st.run(3, exchange=True)
    t_slot = st.temp_of().astype(np.int64)
    e_slot = st.energies().copy()
    st.set_replica_temperatures([1.0, 2.0])
    # the configuration at (temperature t, lane) is now replica t * 64 + lane
    np.testing.assert_array_equal(st.temp_of(), t_slot * 64 + np.arange(64)[None, :])
    np.testing.assert_array_equal(st.energies(), e_slot)
    with pytest.raises(_lib.IgeError):
        st.run(1, exchange=True)          # the replica view exchanges separately
    with pytest.raises(_lib.IgeError):
        st.exchange(1)
    with pytest.raises(_lib.IgeError):
        st.icm(1)
    with pytest.raises(_lib.IgeError):
        st.set_replica_temperatures([1.0, 0.0])
    st.run(2)
    st.exchange_replicas(2)
    st.set_ladder([0.5, 1.5])             # back to slot temperatures, identity permutation
    assert np.array_equal(st.temp_of(), np.repeat(np.arange(2)[:, None], 64, axis=1))
    st.run(3, exchange=True)
    np.testing.assert_array_equal(st.energies(), st.recompute_energies())
    with pytest.raises(_lib.IgeError):
        st.exchange_replicas(1)
