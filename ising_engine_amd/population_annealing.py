"""Population annealing with parallel tempering on the device: the PA-PT hybrid of
UltimateSolver::solve_population_annealing (src/solver/ultimate.rs:559-852).

R = num_pops x 64 members; member m = p*64 + r is lane r of population p across that
population's ladder of num_temps levels (ige_set_populations).  Stage k's ladder spans
[T_cold(k), temp_max], T_t = temp_max (T_cold / temp_max)^alpha_t (alpha_t = t / (nt - 1)
before tuning), level nt - 1 the coldest, and T_cold anneals from temp_max to temp_min
across the stages.
After each stage but the first, member m gains log-weight -(beta_cold - beta_cold_prev) E_m
with E_m its coldest-level energy; when the effective sample size falls below
ess_threshold x R the population is systematically resampled (one offset u, a shuffled
parent list), whole ladders are copied on the device (ige_resample), the log-weights reset
and post_resample_relaxation PT steps follow.  Each stage then runs sweeps_per_stage PT
steps (a sweep and a DEO exchange inside every population), optionally an ICM sweep, and
an incumbent update; the best state seen is 1-opt polished at the end.

Before the stages, the ladder shape is tuned as LadderMode::AcceptanceUniform does it
(ultimate.rs:599-660, ige_pa_tune_ladder): on a throwaway state at the full span
[temp_min, temp_max], ladder_tuning_iters rounds of ladder_tuning_steps PT steps, each
followed by an acceptance-uniformizing update of the log-temperature positions alpha_t;
the stage ladders then use the tuned profile, T_t = temp_max (T_cold / temp_max)^alpha_t.
The throwaway state's seed is the solve's seed xor TUNE_SEED_XOR (the reference draws both
fields from one generator stream).

The library does the sweeps, exchanges, resampling copies, cluster moves, incumbent,
polish and the tuning loop; this module holds the stage loop and the PA bookkeeping the
reference keeps on the host.  The offset u of each resampling comes from a numpy PCG64
generator seeded with the solve's seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib

TUNE_SEED_XOR = 0x74756E65 << 32


def powi(a: float, b: int) -> float:
    """f64::powi as compiled (binary exponentiation, compiler-rt __powidf2)."""
    recip = b < 0
    r = 1.0
    while True:
        if b & 1:
            r *= a
        b = int(b / 2)
        if b == 0:
            break
        a *= a
    return 1.0 / r if recip else r


def uniform_alphas(num_temps: int) -> np.ndarray:
    """alpha_t = t / max(nt - 1, 1) (ultimate.rs:604-605): the geometric profile; one
    temperature gets alpha 0, i.e. temp_max, as in the reference."""
    return np.array([t / max(num_temps - 1, 1) for t in range(num_temps)])


def stage_temperatures(temp_max: float, temp_min: float, num_temps: int, stages: int,
                       alphas=None):
    """[(t_cold, temps[num_temps])] per stage, temps[t] = temp_max (t_cold /
    temp_max)^alpha_t (ultimate.rs:708-721), uniform alphas unless given."""
    k_stages = max(stages, 1)
    cold_factor = (temp_min / temp_max) ** (1.0 / (k_stages - 1)) if k_stages > 1 else 1.0
    alphas = uniform_alphas(num_temps) if alphas is None else alphas
    out = []
    for k in range(k_stages):
        t_cold = temp_max * powi(cold_factor, k) if k_stages > 1 else temp_min
        span = t_cold / temp_max
        out.append((t_cold, np.array([temp_max * span ** float(a) for a in alphas])))
    return out


@dataclass
class PaResult:
    x: np.ndarray
    energy: int
    free_energy: float = 0.0
    ess_history: list = field(default_factory=list)
    family_entropy_history: list = field(default_factory=list)
    resample_events: int = 0
    family: np.ndarray | None = None
    state: object = None  # the device state after the last stage
    alphas: np.ndarray | None = None  # the ladder profile the stages used


def coldest_energies(st: _lib.State, num_pops: int, num_temps: int) -> np.ndarray:
    """E of member p*64 + r at its coldest level num_temps - 1."""
    e, tof = st.energies(), st.temp_of()
    out = np.empty(num_pops * 64, dtype=np.int64)
    for s in range(tof.shape[0]):
        for r in range(64):
            g = int(tof[s, r])
            if g % num_temps == num_temps - 1:
                out[(g // num_temps) * 64 + r] = e[s, r]
    return out


def solve(ctx: _lib.Context, model: _lib.Model, seed: int, num_pops: int = 4,
          num_temps: int = 8, stages: int = 50, sweeps_per_stage: int = 10,
          temp_max: float = 10.0, temp_min: float = 0.1, ess_threshold: float = 0.5,
          post_resample_relaxation: int = 5, use_icm: bool = False,
          polish: bool = True, ladder_tuning_iters: int = 8,
          ladder_tuning_steps: int = 10) -> PaResult:
    P, T = num_pops, num_temps
    R = P * 64
    alphas = uniform_alphas(T)
    if T >= 3 and ladder_tuning_iters > 0 and ladder_tuning_steps > 0:
        tune = _lib.State(model, P * T, seed ^ TUNE_SEED_XOR)
        tune.set_populations(P)
        alphas = tune.tune_ladder(temp_max, temp_min, ladder_tuning_iters,
                                  ladder_tuning_steps, alphas)
        tune.close()
    st = _lib.State(model, P * T, seed)
    st.set_populations(P)
    rng = np.random.Generator(np.random.PCG64(seed))
    log_w = np.zeros(R)
    family = np.arange(R, dtype=np.uint32)
    res = PaResult(x=np.zeros(model.n, np.uint8), energy=0, alphas=alphas)
    prev = None
    for t_cold, temps in stage_temperatures(temp_max, temp_min, T, stages, alphas):
        st.set_ladder(np.tile(1.0 / temps, P))
        beta_cold = 1.0 / t_cold
        if prev is not None:
            log_w -= (beta_cold - prev) * coldest_energies(st, P, T)
            w, ess = _lib.pa_normalize(log_w)
            res.ess_history.append(ess)
            if ess < ess_threshold * R:
                res.free_energy += _lib.pa_free_energy_increment(log_w)
                counts = _lib.pa_systematic_resample(w, rng.random())
                parents = _lib.pa_parents(counts, seed, res.resample_events)
                st.resample(parents)
                family = family[parents]
                res.family_entropy_history.append(_lib.pa_family_entropy(family))
                res.resample_events += 1
                log_w[:] = 0.0
                if post_resample_relaxation:
                    st.run(post_resample_relaxation, exchange=True)
        prev = beta_cold
        st.run(sweeps_per_stage, exchange=True)
        if use_icm:
            st.icm(1)
        st.update_best()
    x, e, _ = st.best()
    if polish:
        xp, ep, _ = model.polish(x)
        x, e = xp[0], int(ep[0])
    res.x, res.energy, res.family = x, e, family
    res.state = st
    return res
