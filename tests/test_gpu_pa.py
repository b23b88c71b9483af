"""Population annealing on the device against the oracle, bit-exact: population-restricted
exchange, ige_resample, and a whole PA-PT solve (stage ladders, reweighting, systematic
resampling, relaxation, ICM, incumbent, polish) against the same stage loop run on the
oracle with the numpy PA arithmetic."""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from ising_engine_amd import population_annealing as pa_drv
from oracle import binding as ob
from oracle import pa

pytestmark = pytest.mark.gpu


def make(ctx, qb):
    return _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)


def test_population_exchange_and_resample_parity(ctx):
    qb = instances.random_qubo(120, 4, 81, offdiag=8, diag=10)
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    st = _lib.State(model, 6, 55, kernel=_lib.KERNEL_GENERIC)
    st.set_populations(3)
    ost = ob.OracleState(csr, 6, 55)
    ost.pops = 3
    ladder = np.tile([1.2, 0.5], 3)
    st.set_ladder(ladder)
    st.run(9, exchange=True)
    ost.run(9, exchange=True, ladder=ladder)
    np.testing.assert_array_equal(st.temp_of(), ost.temp_of)
    s, pa_att, pa_acc = st.stats()
    np.testing.assert_array_equal(pa_att, ost.stats[2::2])
    assert pa_att[1] == 0 and pa_att[3] == 0               # pairs across populations
    par = np.random.default_rng(3).integers(0, 192, size=192).astype(np.uint32)
    st.resample(par)
    ost.resample(par)
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.energies(), ost.energy)
    np.testing.assert_array_equal(st.recompute_energies(), ost.energy)
    with pytest.raises(_lib.IgeError):
        st.resample(np.full(192, 192, np.uint32))
    with pytest.raises(_lib.IgeError):
        st.set_populations(4)


def oracle_pa(csr, seed, P, T, stages, sweeps, tmax, tmin, thr, relax, use_icm,
              tune_iters=0, tune_steps=0):
    alphas = pa_drv.uniform_alphas(T)
    if T >= 3 and tune_iters and tune_steps:
        tune = ob.OracleState(csr, P * T, seed ^ pa_drv.TUNE_SEED_XOR)
        tune.pops = P
        alphas = pa.tune_ladder(tune, tmax, tmin, tune_iters, tune_steps, alphas)
    st = ob.OracleState(csr, P * T, seed)
    st.pops = P
    R = P * 64
    rng = np.random.Generator(np.random.PCG64(seed))
    log_w = np.zeros(R)
    family = np.arange(R, dtype=np.uint32)
    inc = ob.Incumbent(csr.n)
    ess_hist, events, prev = [], 0, None
    for t_cold, temps in pa_drv.stage_temperatures(tmax, tmin, T, stages, alphas):
        ladder = np.tile(1.0 / temps, P)
        beta_cold = 1.0 / t_cold
        if prev is not None:
            ecold = np.empty(R, dtype=np.int64)
            for s in range(P * T):
                for r in range(64):
                    g = int(st.temp_of[s, r])
                    if g % T == T - 1:
                        ecold[(g // T) * 64 + r] = st.energy[s, r]
            log_w -= (beta_cold - prev) * ecold
            w = pa.normalized_weights(log_w)
            ess = pa.effective_sample_size(w)
            ess_hist.append(ess)
            if ess < thr * R:
                counts = pa.systematic_resample(w, rng.random())
                parents = pa.parents(counts, seed, events)
                st.resample(parents)
                family = family[parents]
                events += 1
                log_w[:] = 0.0
                if relax:
                    st.run(relax, exchange=True, ladder=ladder)
        prev = beta_cold
        st.run(sweeps, exchange=True, ladder=ladder)
        if use_icm:
            st.icm(1)
        inc.update(st.words, st.temp_of, st.energy)
    return st, family, ess_hist, events, inc, alphas


@pytest.mark.parametrize("use_icm,tune", [(False, 0), (True, 0), (False, 3)],
                         ids=["pa_pt", "pa_pt_icm", "pa_pt_tuned"])
def test_pa_pt_solve_matches_oracle(ctx, use_icm, tune):
    print(f"[pa] icm={use_icm} tune={tune}", flush=True)
    qb = instances.random_qubo(90, 4, 82, offdiag=20, diag=30)
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    t_max, t_min = instances.suggested_temp_range(qb)
    args = dict(num_pops=2, num_temps=4, stages=10, sweeps_per_stage=3, temp_max=t_max,
                temp_min=t_min, ess_threshold=0.9, post_resample_relaxation=2, use_icm=use_icm,
                ladder_tuning_iters=tune, ladder_tuning_steps=5)
    res = pa_drv.solve(ctx, model, 4321, **args)
    st, family, ess_hist, events, inc, alphas = oracle_pa(
        csr, 4321, 2, 4, 10, 3, t_max, t_min, 0.9, 2, use_icm, tune, 5)
    np.testing.assert_array_equal(res.alphas, alphas)
    if tune:
        assert not np.array_equal(alphas, pa_drv.uniform_alphas(4))
    assert res.resample_events == events > 0
    np.testing.assert_allclose(res.ess_history, ess_hist, rtol=1e-12)
    np.testing.assert_array_equal(res.family, family)
    np.testing.assert_array_equal(res.state.spins(), st.words)
    np.testing.assert_array_equal(res.state.energies(), st.energy)
    np.testing.assert_array_equal(res.state.temp_of(), st.temp_of)
    xo = inc.x.copy()
    csr.polish_1opt(xo)
    np.testing.assert_array_equal(res.x, xo)
    assert res.energy == csr.energy(xo) <= inc.energy.value


def test_ladder_tuning_matches_the_oracle(ctx):
    # ige_pa_tune_ladder (ultimate.rs:599-660) against the oracle's loop: the same
    # per-pair swap counts every round, so the same alphas bit for bit, and the same state
    qb = instances.random_qubo(150, 6, 83, offdiag=30, diag=40)
    csr = ob.CSR.from_qubo(qb)
    t_max, t_min = instances.suggested_temp_range(qb)
    P, T = 3, 6
    st = _lib.State(make(ctx, qb), P * T, 99, kernel=_lib.KERNEL_GENERIC)
    st.set_populations(P)
    ost = ob.OracleState(csr, P * T, 99)
    ost.pops = P
    got = st.tune_ladder(t_max, t_min, 8, 10)
    want = pa.tune_ladder(ost, t_max, t_min, 8, 10)
    np.testing.assert_array_equal(got, want)
    assert got[0] == 0.0 and got[-1] == 1.0 and np.all(np.diff(got) > 0)
    np.testing.assert_array_equal(st.spins(), ost.words)
    np.testing.assert_array_equal(st.temp_of(), ost.temp_of)
    # fewer than 3 temperatures: unchanged, as in the reference
    st2 = _lib.State(make(ctx, qb), 4, 5)
    st2.set_populations(2)
    np.testing.assert_array_equal(st2.tune_ladder(t_max, t_min, 8, 10), [0.0, 1.0])


def test_device_swap_rates_on_the_two_level_system(ctx):
    # one spin, E = h x: the device's per-pair DEO acceptance against the equilibrium closed
    # form, and tuning rounds that track the closed-form iteration
    h = 2
    qb = instances.Qubo(n=1, offset=0, linear=np.array([h]), pi=np.zeros(0, np.uint32),
                        pj=np.zeros(0, np.uint32), q=np.zeros(0, np.int64), name="two_level")
    betas = 1.0 / np.array([4.0, 2.5, 1.6, 1.0, 0.6])
    st = _lib.State(make(ctx, qb), 5, 17, kernel=_lib.KERNEL_GENERIC)
    st.set_ladder(betas)
    st.run(200, exchange=True)
    _, a0, c0 = st.stats()
    st.run(4000, exchange=True)
    _, a1, c1 = st.stats()
    rates = (c1.astype(np.int64) - c0) / (a1.astype(np.int64) - a0)
    exact = pa.two_level_exchange_rates(h, betas)
    assert np.all(np.abs(rates - exact) < 0.012), (rates, exact)
    tune = _lib.State(make(ctx, qb), 20, 23, kernel=_lib.KERNEL_GENERIC)
    tune.set_populations(4)
    alphas = tune.tune_ladder(20.0, 0.1, 12, 200)
    from test_pa import exact_tuning
    closed, _ = exact_tuning(float(h), 20.0, 0.1, 5, 12)
    assert np.max(np.abs(alphas - closed)) < 0.006, (alphas, closed)
