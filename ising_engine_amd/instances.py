"""Synthetic instances for the sweep/exchange workloads (numpy, seeded).

Ising models H(sigma) = -sum_{i<j} J_ij sigma_i sigma_j are lowered to QUBO
with sigma = 1 - 2x: pair q_ij = -4 J_ij, linear_i = 2 sum_j J_ij,
offset = -sum_{i<j} J_ij, so E_QUBO(x) == H(sigma(x)) exactly and
temperatures stay in Ising units (src/bin/sk_parisi_benchmark.rs:56-110).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Qubo:
    """Upper-triangle integer QUBO: E = offset + lin.x + sum q_k x_pi x_pj."""
    n: int
    offset: int
    linear: np.ndarray  # int64 [n]
    pi: np.ndarray      # uint32 [m], pi < pj
    pj: np.ndarray      # uint32 [m]
    q: np.ndarray       # int64 [m]
    name: str = ""

    def energy(self, x: np.ndarray) -> np.ndarray:
        """Energy of configurations x ([..., n] of 0/1), exact int64."""
        x = np.asarray(x, dtype=np.int64)
        e = self.offset + x @ self.linear
        e = e + (self.q * x[..., self.pi.astype(np.int64)] * x[..., self.pj.astype(np.int64)]).sum(-1)
        return e


def ising_to_qubo(n: int, bi: np.ndarray, bj: np.ndarray, J: np.ndarray, name: str = "") -> Qubo:
    bi = np.asarray(bi, dtype=np.int64)
    bj = np.asarray(bj, dtype=np.int64)
    J = np.asarray(J, dtype=np.int64)
    lo = np.minimum(bi, bj)
    hi = np.maximum(bi, bj)
    if np.any(lo == hi):
        raise ValueError("self-bonds are not allowed")
    linear = np.zeros(n, dtype=np.int64)
    np.add.at(linear, lo, 2 * J)
    np.add.at(linear, hi, 2 * J)
    return Qubo(n=n, offset=int(-J.sum()), linear=linear, pi=lo.astype(np.uint32),
                pj=hi.astype(np.uint32), q=(-4 * J).astype(np.int64), name=name)


def hypercubic_pm_j(L: int, dim: int, seed: int) -> Qubo:
    """+-J Edwards-Anderson model on a periodic L^dim lattice, J uniform in {-1, +1}.

    Site index = x0 + L x1 + L^2 x2 ...; one bond per site and direction.
    """
    rng = np.random.default_rng(seed)
    n = L ** dim
    idx = np.arange(n, dtype=np.int64)
    coords = [(idx // L ** a) % L for a in range(dim)]
    bi, bj = [], []
    for a in range(dim):
        nb = idx + (((coords[a] + 1) % L) - coords[a]) * L ** a
        bi.append(idx)
        bj.append(nb)
    bi = np.concatenate(bi)
    bj = np.concatenate(bj)
    J = rng.choice(np.array([-1, 1], dtype=np.int64), size=bi.shape[0])
    return ising_to_qubo(n, bi, bj, J, name=f"pmJ{dim}d_L{L}")


def _er_pairs(n: int, mean_degree: float, rng) -> tuple[np.ndarray, np.ndarray]:
    """Distinct pairs i < j of G(n, p = mean_degree / (n - 1)), vectorised."""
    total = n * (n - 1) // 2
    m = int(rng.binomial(total, min(1.0, mean_degree / max(n - 1, 1))))
    keys = np.zeros(0, dtype=np.int64)
    while keys.size < m:
        a = rng.integers(0, n, size=2 * (m - keys.size) + 16, dtype=np.int64)
        b = rng.integers(0, n, size=a.size, dtype=np.int64)
        ok = a != b
        lo = np.minimum(a[ok], b[ok])
        hi = np.maximum(a[ok], b[ok])
        new = lo * n + hi
        keys = np.concatenate([keys, new[np.isin(new, keys, invert=True)]])
        keys = keys[np.sort(np.unique(keys, return_index=True)[1])]
    keys = keys[:m]
    return (keys // n).astype(np.uint32), (keys % n).astype(np.uint32)


def erdos_renyi_pm_j(n: int, mean_degree: float, seed: int) -> Qubo:
    """Erdos-Renyi G(n, p = mean_degree/(n-1)) with J uniform in {-1, +1}."""
    rng = np.random.default_rng(seed)
    bi, bj = _er_pairs(n, mean_degree, rng)
    J = rng.choice(np.array([-1, 1], dtype=np.int64), size=bi.size)
    return ising_to_qubo(n, bi, bj, J, name=f"ER_n{n}_d{mean_degree:g}")


def sk_int8(n: int, seed: int, gain: float = 32.0) -> Qubo:
    """Sherrington-Kirkpatrick: J_ij ~ N(0, 1/n) for every pair, quantised to int8 by
    J_q = clamp(round(J sqrt(n) gain), -127, 127) (zeros dropped by the CSR build), lowered
    to QUBO with sigma = 1 - 2x (src/bin/sk_parisi_benchmark.rs:56-110)."""
    rng = np.random.default_rng(seed)
    iu, ju = np.triu_indices(n, 1)
    J = np.clip(np.rint(rng.standard_normal(iu.size) * gain), -127, 127).astype(np.int64)
    keep = J != 0
    return ising_to_qubo(n, iu[keep], ju[keep], J[keep], name=f"SK_n{n}_int8")


def dense_qubo(n: int, seed: int, scale: int = 3, density: float = 0.6) -> Qubo:
    r"""Dense QUBO whose couplings are scale * int8: q_ij = scale * a_ij, a_ij uniform in
    [-127, 127] \ {0} on a fraction `density` of all pairs; linear terms uniform in
    [-2000, 2000]."""
    rng = np.random.default_rng(seed)
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(iu.size) < density
    a = rng.integers(-127, 128, size=int(keep.sum())).astype(np.int64)
    a[a == 0] = 1
    lin = rng.integers(-2000, 2001, size=n).astype(np.int64)
    return Qubo(n=n, offset=0, linear=lin, pi=iu[keep].astype(np.uint32),
                pj=ju[keep].astype(np.uint32), q=scale * a, name=f"denseQUBO_n{n}_s{scale}")


def random_qubo(n: int, mean_degree: float, seed: int, offdiag: int = 100, diag: int = 200) -> Qubo:
    """Sparse QUBO: ER graph, Q_ij int uniform in [-offdiag, offdiag] (zeros redrawn to +1),
    Q_ii uniform in [-diag, diag] as the linear term."""
    rng = np.random.default_rng(seed)
    pi, pj = _er_pairs(n, mean_degree, rng)
    q = rng.integers(-offdiag, offdiag + 1, size=pi.size).astype(np.int64)
    q[q == 0] = 1
    linear = rng.integers(-diag, diag + 1, size=n).astype(np.int64)
    return Qubo(n=n, offset=0, linear=linear, pi=pi, pj=pj, q=q,
                name=f"QUBO_n{n}_d{mean_degree:g}")


def suggested_temp_range(qb: Qubo) -> tuple[float, float]:
    """(T_max, T_min) of src/solver/schedule.rs:21-43: T_max = max_i local_i / ln 2 with
    local_i = |lin_i| + sum_j |q_ij|, T_min = min(min nonzero |coef| / ln 100, T_max / 2)."""
    local = np.abs(qb.linear).astype(np.float64)
    aq = np.abs(qb.q).astype(np.float64)
    np.add.at(local, qb.pi.astype(np.int64), aq)
    np.add.at(local, qb.pj.astype(np.int64), aq)
    nz = np.concatenate([np.abs(qb.linear[qb.linear != 0]), aq[aq != 0]])
    if local.max() == 0 or nz.size == 0:
        return 1.0, 0.1
    t_max = local.max() / np.log(2.0)
    t_min = min(nz.min() / np.log(100.0), t_max * 0.5)
    return float(t_max), float(t_min)


def workload_qubo_batch(count: int = 4096, sweeps: int = 2000, seed: int = 4,
                        n_lo: int = 64, n_hi: int = 8192) -> dict:
    """Independent sparse QUBO instances: N log-uniform in [n_lo, n_hi], mean degree 8,
    Q_ij uniform int in [-100, 100], Q_ii in [-200, 200], 64 replicas each, a geometric anneal
    between the instance's suggested temperatures."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), size=count))).astype(np.int64)
    qubos, b_lo, b_hi = [], [], []
    for i, n in enumerate(ns):
        qb = random_qubo(int(n), 8, int(rng.integers(1 << 62)))
        t_max, t_min = suggested_temp_range(qb)
        qubos.append(qb)
        b_lo.append(1.0 / t_max)
        b_hi.append(1.0 / t_min)
    seeds = rng.integers(1, 1 << 62, size=count)
    return dict(qubos=qubos, seeds=seeds, sweeps=[sweeps] * count, beta_lo=b_lo, beta_hi=b_hi,
                name=f"qubo_batch_{count}_n{n_lo}-{n_hi}_d8")


def geometric_betas(t_hi: float, t_lo: float, k: int) -> np.ndarray:
    """Inverse temperatures of a geometric temperature ladder T_t = t_hi (t_lo/t_hi)^(t/(k-1)),
    hottest first (src/solver/ultimate.rs:237-246 ordering)."""
    if k == 1:
        return np.array([1.0 / t_hi])
    t = t_hi * (t_lo / t_hi) ** (np.arange(k) / (k - 1))
    return 1.0 / t


def anneal_schedule(beta_lo: float, beta_hi: float, sweeps: int) -> np.ndarray:
    """beta_k = beta_lo (beta_hi/beta_lo)^(k/(sweeps-1)), one row per sweep."""
    if sweeps == 1:
        return np.array([[beta_hi]])
    k = np.arange(sweeps, dtype=np.float64)
    return (beta_lo * (beta_hi / beta_lo) ** (k / (sweeps - 1)))[:, None]


# --------------------------------------------------------------------------
# The named workloads.
# --------------------------------------------------------------------------
def workload_2d_ej(seed: int = 1) -> dict:
    """2D +-J L=32 periodic, 64 replicas, beta 0.1 -> 3.0 geometric over 1000 sweeps."""
    qb = hypercubic_pm_j(32, 2, seed)
    return dict(qubo=qb, cells=1, sweeps=1000, exchange=False,
                schedule=anneal_schedule(0.1, 3.0, 1000), name="2d_pmJ_L32_anneal")


def workload_3d_pt(L: int = 64, temps: int = 32, sweeps: int = 10000, seed: int = 2) -> dict:
    """3D +-J cubic periodic, `temps` temperatures geometric T 0.5 -> 2.0, 64 replicas per
    temperature, DEO swap attempt every sweep.  Ladder listed hottest first."""
    qb = hypercubic_pm_j(L, 3, seed)
    return dict(qubo=qb, cells=temps, sweeps=sweeps, exchange=True,
                ladder=geometric_betas(2.0, 0.5, temps), name=f"3d_pmJ_L{L}_pt{temps}")


def workload_er_anneal(n: int = 20000, degree: float = 6, sweeps: int = 5000,
                       seed: int = 3) -> dict:
    """Max-Cut-shaped Ising: Erdos-Renyi G(n, degree/(n-1)) with J = +-1, 64 replicas, a
    geometric anneal over the suggested temperature range."""
    qb = erdos_renyi_pm_j(n, degree, seed)
    t_max, t_min = suggested_temp_range(qb)
    return dict(qubo=qb, cells=1, sweeps=sweeps, beta_lo=1.0 / t_max, beta_hi=1.0 / t_min,
                name=f"er_pmJ_n{n}_d{degree:g}_anneal")


def workload_sk_anneal(n: int = 4096, sweeps: int = 5000, seed: int = 5) -> dict:
    """Dense SK with int8 couplings (gain 32), 64 replicas, a geometric anneal over the
    suggested temperature range."""
    qb = sk_int8(n, seed)
    t_max, t_min = suggested_temp_range(qb)
    return dict(qubo=qb, cells=1, sweeps=sweeps, beta_lo=1.0 / t_max, beta_hi=1.0 / t_min,
                name=f"sk_int8_n{n}_anneal")


def serving_request(rng, n_lo: int = 128, n_hi: int = 16384, w_max: int = 1000,
                    sweeps_lo: int = 500, sweeps_hi: int = 5000, num_temps: int = 10,
                    sweeps_per_exchange: int = 100) -> dict:
    """One solve request: 50/50 Ising or QUBO, N log-uniform in [n_lo, n_hi], mean degree
    uniform in 4..16, integer weights uniform in [-w_max, w_max] (nonzero), sweep budget
    uniform in [sweeps_lo, sweeps_hi].  As POST /api/v1/solve runs it
    (src/bin/server_api.rs:111-169): UltimateSolver with num_temps temperatures between the
    suggested temperatures, sweeps_per_exchange steps per exchange round and the budget
    rounded to whole rounds (`sweeps`); beta_lo / beta_hi give the same range for a
    one-cell anneal."""
    n = int(np.rint(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
    deg = int(rng.integers(4, 17))
    pi, pj = _er_pairs(n, deg, rng)
    w = rng.integers(-w_max, w_max + 1, size=pi.size).astype(np.int64)
    w[w == 0] = 1
    if rng.random() < 0.5:
        qb = ising_to_qubo(n, pi, pj, w, name=f"ising_n{n}_d{deg}")
    else:
        lin = rng.integers(-w_max, w_max + 1, size=n).astype(np.int64)
        qb = Qubo(n=n, offset=0, linear=lin, pi=pi, pj=pj, q=w, name=f"qubo_n{n}_d{deg}")
    t_max, t_min = suggested_temp_range(qb)
    budget = int(rng.integers(sweeps_lo, sweeps_hi + 1))
    spe = max(1, min(sweeps_per_exchange, budget))
    return dict(qubo=qb, seed=int(rng.integers(1, 1 << 62)),
                sweeps=max(1, round(budget / spe)) * spe, sweeps_per_exchange=spe,
                num_temps=num_temps, temp_max=t_max, temp_min=t_min,
                beta_lo=1.0 / t_max, beta_hi=1.0 / t_min)
