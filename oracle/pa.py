"""Population-annealing arithmetic restated in numpy, used ONLY as a checker.

TEST INFRASTRUCTURE. Only tests/ may load this module; the product computes these in
ising_engine_amd/csrc/pa.cpp.
  normalized_weights, effective_sample_size  src/solver/population_annealing.rs:45-60
  free_energy_increment                      src/solver/population_annealing.rs:62-75
  family_entropy                             src/solver/population_annealing.rs:77-101
  systematic_resample                        src/solver/population_annealing.rs:103-124
  parents (list + Fisher-Yates shuffle)      src/solver/population_annealing.rs:237-246,
      with j = floor(u (k + 1) / 2^32), u = Philox(seed; event, k, 0, 7 << 24)[0]
  adapt_spacings                             src/solver/population_annealing.rs:137-148
  tune_ladder (AcceptanceUniform tuning)     src/solver/ultimate.rs:599-660, on an
      OracleState with the swap statistics of or_run
"""
from __future__ import annotations

import math

import numpy as np

from . import binding as ob


def normalized_weights(log_w):
    lw = np.asarray(log_w, dtype=np.float64)
    e = np.exp(lw - lw.max())
    return e / e.sum()


def effective_sample_size(w):
    return 1.0 / float(np.sum(np.asarray(w) ** 2))


def free_energy_increment(log_w) -> float:
    lw = np.asarray(log_w, dtype=np.float64)
    m = lw.max()
    if not math.isfinite(m):
        return 0.0
    return float(m + math.log(np.exp(lw - m).sum()) - math.log(lw.size))


def systematic_resample(w, u: float):
    r_total = len(w)
    counts = [0] * r_total
    cum = 0.0
    j = 0
    for i, wi in enumerate(w):
        cum += wi
        while j < r_total and (j + u) / r_total < cum:
            counts[i] += 1
            j += 1
    if j < r_total:
        counts[-1] += r_total - j
    return np.array(counts, dtype=np.uint32)


def parents(counts, seed: int, event: int):
    par = [i for i, c in enumerate(counts) for _ in range(int(c))]
    key = [seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF]
    for k in range(len(par) - 1, 0, -1):
        u = int(ob.philox([event, k, 0, 7 << 24], key)[0])
        j = (u * (k + 1)) >> 32
        par[k], par[j] = par[j], par[k]
    return np.array(par, dtype=np.uint32)


def family_entropy(family):
    f = np.asarray(family)
    _, c = np.unique(f, return_counts=True)
    return float(np.sum((c / f.size) ** 2) * f.size), int(c.size)


def adapt_spacings(spacings, acceptance):
    """d_t sqrt(max(A_t, 0.02)) rescaled to the old total; sums in index order."""
    raw = [float(d) * math.sqrt(max(float(a), 0.02)) for d, a in zip(spacings, acceptance)]
    total_old = 0.0
    for d in spacings:
        total_old += float(d)
    total_new = 0.0
    for d in raw:
        total_new += d
    return np.array([d * total_old / total_new for d in raw])


def ladder_temperatures(temp_max: float, temp_min: float, alphas):
    """T_t = temp_max (temp_min / temp_max)^alpha_t."""
    return np.array([temp_max * (temp_min / temp_max) ** float(a) for a in alphas])


def pair_rates(stats_before, stats_after, pops: int, T: int):
    """Swap acceptance of pair (t, t + 1) summed over the populations."""
    att = stats_after[2::2].astype(np.int64) - stats_before[2::2].astype(np.int64)
    acc = stats_after[3::2].astype(np.int64) - stats_before[3::2].astype(np.int64)
    out = np.zeros(T - 1)
    for t in range(T - 1):
        na = sum(int(att[p * T + t]) for p in range(pops))
        nc = sum(int(acc[p * T + t]) for p in range(pops))
        out[t] = nc / na if na else 0.0
    return out


def tune_ladder(st, temp_max: float, temp_min: float, iters: int, steps: int, alphas=None):
    """The tuning phase on oracle state st (populations st.pops); returns the alphas."""
    T = st.cells // st.pops
    alphas = (np.array([t / max(T - 1, 1) for t in range(T)]) if alphas is None
              else np.array(alphas, dtype=np.float64))
    if T < 3 or iters == 0 or steps == 0:
        return alphas
    for _ in range(iters):
        temps = ladder_temperatures(temp_max, temp_min, alphas)
        ladder = np.tile(1.0 / temps, st.pops)
        before = st.stats.copy()
        st.run(steps, exchange=True, ladder=ladder)
        rates = pair_rates(before, st.stats, st.pops, T)
        new = adapt_spacings(np.diff(alphas), rates)
        acc = 0.0
        for t, d in enumerate(new):
            acc += d
            alphas[t + 1] = acc
        alphas[T - 1] = 1.0
    return alphas


def two_level_exchange_rates(h: float, betas):
    """Exact equilibrium acceptance of DEO pair (t, t + 1) for one spin of energy h x
    (x in {0, 1}) at betas[t] < betas[t + 1]: a swap fails only with the hot replica
    excited and the cold one not, and then is accepted with exp(-(b_{t+1} - b_t) h)."""
    b = np.asarray(betas, dtype=np.float64)
    p1 = np.exp(-b * h) / (1.0 + np.exp(-b * h))
    return 1.0 - p1[:-1] * (1.0 - p1[1:]) * (1.0 - np.exp(-(b[1:] - b[:-1]) * h))
