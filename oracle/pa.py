"""Population-annealing arithmetic restated in numpy, used ONLY as a checker.

TEST INFRASTRUCTURE. Only tests/ may load this module; the product computes these in
ising_engine_amd/csrc/pa.cpp.
  normalized_weights, effective_sample_size  src/solver/population_annealing.rs:45-60
  free_energy_increment                      src/solver/population_annealing.rs:62-75
  family_entropy                             src/solver/population_annealing.rs:77-101
  systematic_resample                        src/solver/population_annealing.rs:103-124
  parents (list + Fisher-Yates shuffle)      src/solver/population_annealing.rs:237-246,
      with j = floor(u (k + 1) / 2^32), u = Philox(seed; event, k, 0, 7 << 24)[0]
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
