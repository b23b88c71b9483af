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


def erdos_renyi_pm_j(n: int, mean_degree: float, seed: int) -> Qubo:
    """Erdos-Renyi G(n, p = mean_degree/(n-1)) with J uniform in {-1, +1}."""
    rng = np.random.default_rng(seed)
    m_target = rng.binomial(n * (n - 1) // 2, mean_degree / (n - 1))
    # Sample distinct pairs by rejection on a set of encoded pairs.
    seen = set()
    bi, bj = [], []
    while len(bi) < m_target:
        a = rng.integers(0, n, size=2 * (m_target - len(bi)) + 16)
        b = rng.integers(0, n, size=a.shape[0])
        for u, v in zip(a.tolist(), b.tolist()):
            if u == v:
                continue
            lo, hi = (u, v) if u < v else (v, u)
            key = lo * n + hi
            if key in seen:
                continue
            seen.add(key)
            bi.append(lo)
            bj.append(hi)
            if len(bi) == m_target:
                break
    J = rng.choice(np.array([-1, 1], dtype=np.int64), size=len(bi))
    return ising_to_qubo(n, np.array(bi), np.array(bj), J, name=f"ER_n{n}_d{mean_degree:g}")


def random_qubo(n: int, mean_degree: float, seed: int, offdiag: int = 100, diag: int = 200) -> Qubo:
    """Sparse QUBO: ER graph, Q_ij int uniform in [-offdiag, offdiag] (zeros redrawn to +1),
    Q_ii uniform in [-diag, diag] as the linear term."""
    rng = np.random.default_rng(seed)
    er = erdos_renyi_pm_j(n, mean_degree, int(rng.integers(1 << 62)))
    q = rng.integers(-offdiag, offdiag + 1, size=er.q.shape[0]).astype(np.int64)
    q[q == 0] = 1
    linear = rng.integers(-diag, diag + 1, size=n).astype(np.int64)
    return Qubo(n=n, offset=0, linear=linear, pi=er.pi, pj=er.pj, q=q,
                name=f"QUBO_n{n}_d{mean_degree:g}")


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
