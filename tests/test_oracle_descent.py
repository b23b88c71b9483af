"""Pins the oracle's deterministic descent (SteepestDescent, steepest_descent_1opt) and the
incumbent update before they are trusted as GPU checkers.

The checks do not reuse the gain formula: single-flip energy changes come from full energy
evaluations (Qubo.energy, numpy) of the flipped configurations, so a wrong sign, a wrong
neighbour update or a wrong tie-break in the oracle shows up as a different trajectory.
"""
import numpy as np
import pytest

from oracle import binding as ob
from ising_engine_amd import instances


def pack(x):
    """[cells*64][n] bytes -> [cells][n] words (bit r = replica r of the cell)."""
    reps, n = x.shape
    cells = reps // 64
    b = x.reshape(cells, 64, n).astype(np.uint64)
    return (b << np.arange(64, dtype=np.uint64)[None, :, None]).sum(axis=1, dtype=np.uint64)


def unpack(words):
    cells, n = words.shape
    bits = (words[:, None, :] >> np.arange(64, dtype=np.uint64)[None, :, None]) & np.uint64(1)
    return bits.reshape(cells * 64, n).astype(np.uint8)


def flip_deltas(qb, x):
    """E(x with site i flipped) - E(x) for every i, by full energy evaluation."""
    flipped = np.repeat(x[None, :].astype(np.int64), qb.n, axis=0)
    idx = np.arange(qb.n)
    flipped[idx, idx] ^= 1
    return qb.energy(flipped) - qb.energy(x)


def brute_steepest(qb, x, clamped=None, max_steps=None):
    """Steepest descent by enumeration: flip the lowest site among the most negative deltas."""
    x = x.copy()
    path = []
    while max_steps is None or len(path) < max_steps:
        d = flip_deltas(qb, x).astype(np.float64)
        if clamped is not None:
            d[clamped.astype(bool)] = np.inf
        i = int(np.argmin(d))  # first occurrence: lowest site on ties
        if not d[i] < 0:
            break
        x[i] ^= 1
        path.append(i)
    return x, path


def ir7():
    """The seven-site instance of the reference's steepest_descent.rs:130-146 test."""
    pairs = [(0, 1, -1), (1, 2, 2), (2, 3, -3), (0, 3, 1), (3, 4, -1), (4, 5, 2), (5, 6, -2),
             (2, 6, 1)]
    pi, pj, q = (np.array(c) for c in zip(*pairs))
    return instances.Qubo(n=7, offset=-2, linear=np.array([1, -2, 3, 0, -1, 2, 1], dtype=np.int64),
                          pi=pi.astype(np.uint32), pj=pj.astype(np.uint32),
                          q=q.astype(np.int64), name="ir7")


def tie_qubo(n, seed):
    """Small weights on a dense graph: many equal deltas, so the tie-break matters."""
    rng = np.random.default_rng(seed)
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(iu.size) < 0.3
    q = rng.integers(-2, 3, size=int(keep.sum())).astype(np.int64)
    q[q == 0] = 1
    lin = rng.integers(-2, 3, size=n).astype(np.int64)
    return instances.Qubo(n=n, offset=5, linear=lin, pi=iu[keep].astype(np.uint32),
                          pj=ju[keep].astype(np.uint32), q=q, name=f"tie{n}")


def test_ir7_steepest_trajectory_step_by_step():
    qb = ir7()
    csr = ob.CSR.from_qubo(qb)
    init = np.array([1, 1, 0, 1, 0, 1, 1], dtype=np.uint8)
    words = pack(np.repeat(init[None, :], 64, axis=0))
    energies = csr.energies(words)
    e0 = int(qb.energy(init))
    assert (energies == e0).all()
    x_end, path = brute_steepest(qb, init)
    assert len(path) >= 1
    prev = init.copy()
    for site in path:
        steps, flips = csr.steepest_descent(words, energies, 1)
        assert steps == 1 and (flips == 1).all()
        x = unpack(words)
        assert (x == x[0]).all()
        assert np.flatnonzero(x[0] ^ prev).tolist() == [site]
        prev = x[0].copy()
    steps, flips = csr.steepest_descent(words, energies, 10)
    assert steps == 0 and (flips == 0).all()
    np.testing.assert_array_equal(unpack(words)[0], x_end)
    assert (energies == qb.energy(x_end)).all() and (energies <= e0).all()
    x = init.copy()
    imp, nflip = csr.polish_1opt(x)
    np.testing.assert_array_equal(x, x_end)
    assert imp == e0 - int(qb.energy(x_end)) and nflip == len(path)


def test_ir7_every_start_state():
    """All 2^7 configurations of the seven-site instance as the 128 replicas of two slots."""
    qb = ir7()
    csr = ob.CSR.from_qubo(qb)
    x0 = ((np.arange(128)[:, None] >> np.arange(7)[None, :]) & 1).astype(np.uint8)
    words = pack(x0)
    energies = csr.energies(words)
    steps, flips = csr.steepest_descent(words, energies, 1000)
    x = unpack(words)
    paths = [brute_steepest(qb, x0[r]) for r in range(128)]
    for r, (xe, path) in enumerate(paths):
        np.testing.assert_array_equal(x[r], xe)
        assert flips.reshape(-1)[r] == len(path)
    assert steps == max(len(p) for _, p in paths) >= 2
    np.testing.assert_array_equal(energies.reshape(-1), qb.energy(x))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_descent_matches_enumeration_with_ties(seed):
    qb = tie_qubo(24, seed)
    csr = ob.CSR.from_qubo(qb)
    rng = np.random.default_rng(100 + seed)
    x0 = rng.integers(0, 2, size=(128, qb.n), dtype=np.uint8)
    words = pack(x0)
    energies = csr.energies(words)
    steps, flips = csr.steepest_descent(words, energies, 10_000)
    x = unpack(words)
    longest = 0
    for r in range(128):
        xe, path = brute_steepest(qb, x0[r])
        np.testing.assert_array_equal(x[r], xe)
        assert flips.reshape(-1)[r] == len(path)
        longest = max(longest, len(path))
        assert (flip_deltas(qb, xe) >= 0).all()
        xp = x0[r].copy()
        imp, nflip = csr.polish_1opt(xp)
        np.testing.assert_array_equal(xp, xe)
        assert imp == int(qb.energy(x0[r]) - qb.energy(xe)) and nflip == len(path)
    assert steps == longest
    np.testing.assert_array_equal(energies.reshape(-1), qb.energy(x))


def test_descent_budget_stops_midway():
    qb = tie_qubo(30, 9)
    csr = ob.CSR.from_qubo(qb)
    x0 = np.random.default_rng(5).integers(0, 2, size=(64, qb.n), dtype=np.uint8)
    words = pack(x0)
    energies = csr.energies(words)
    steps, flips = csr.steepest_descent(words, energies, 2)
    assert steps == 2
    x = unpack(words)
    for r in range(64):
        xe, path = brute_steepest(qb, x0[r], max_steps=2)
        np.testing.assert_array_equal(x[r], xe)
        assert flips[0, r] == len(path)


def test_clamped_sites_never_move():
    qb = instances.random_qubo(40, 5, 11, offdiag=9, diag=9)
    csr = ob.CSR.from_qubo(qb)
    rng = np.random.default_rng(3)
    clamped = (rng.random(qb.n) < 0.3).astype(np.uint8)
    x0 = rng.integers(0, 2, size=(64, qb.n), dtype=np.uint8)
    words = pack(x0)
    energies = csr.energies(words)
    csr.steepest_descent(words, energies, 10_000, clamped=clamped)
    x = unpack(words)
    for r in range(64):
        xe, _ = brute_steepest(qb, x0[r], clamped=clamped)
        np.testing.assert_array_equal(x[r], xe)
        np.testing.assert_array_equal(x[r][clamped == 1], x0[r][clamped == 1])
        free = flip_deltas(qb, xe)[clamped == 0]
        assert (free >= 0).all()
        xp = x0[r].copy()
        csr.polish_1opt(xp, clamped=clamped)
        np.testing.assert_array_equal(xp, xe)


def test_linear_only_closed_form():
    """No couplings: the descent sets exactly the negative-linear sites, most negative first."""
    lin = np.array([5, -3, 0, -7, -3, 2, -1, 0], dtype=np.int64)
    qb = instances.Qubo(n=8, offset=4, linear=lin, pi=np.zeros(0, np.uint32),
                        pj=np.zeros(0, np.uint32), q=np.zeros(0, np.int64))
    csr = ob.CSR.from_qubo(qb)
    words = np.zeros((1, 8), dtype=np.uint64)
    energies = csr.energies(words)
    steps, flips = csr.steepest_descent(words, energies, 100)
    assert steps == 4 and (flips == 4).all()
    np.testing.assert_array_equal(unpack(words)[0], (lin < 0).astype(np.uint8))
    assert (energies == 4 + lin[lin < 0].sum()).all()
    # one step at a time: -7 (site 3), then the -3 tie resolved to site 1, then 4, then 6
    words[:] = 0
    energies = csr.energies(words)
    order = []
    for _ in range(4):
        before = unpack(words)[0].copy()
        csr.steepest_descent(words, energies, 1)
        order.append(int(np.flatnonzero(unpack(words)[0] ^ before)[0]))
    assert order == [3, 1, 4, 6]


def test_incumbent_update_by_enumeration():
    """update_incumbent: strict improvement only, ties to the lowest (temperature, lane)."""
    qb = instances.random_qubo(30, 4, 21, offdiag=3, diag=3)
    csr = ob.CSR.from_qubo(qb)
    cells = 3
    st = ob.OracleState(csr, cells, 77)
    inc = ob.Incumbent(qb.n)
    ladder = instances.geometric_betas(2.0, 0.2, cells)
    best_e, best_x = None, None
    for k in range(12):
        if k:
            st.run(1, exchange=True, ladder=ladder)
        x = unpack(st.words)
        e = qb.energy(x)
        np.testing.assert_array_equal(e, st.energy.reshape(-1))
        t = st.temp_of.reshape(-1).astype(np.int64)
        lane = np.tile(np.arange(64), cells)
        j = np.lexsort((lane, t, e))[0]
        changed = best_e is None or e[j] < best_e
        if changed:
            best_e, best_x = int(e[j]), x[j].copy()
        assert inc.update(st.words, st.temp_of, st.energy) == changed
        assert inc.energy.value == best_e
        np.testing.assert_array_equal(inc.x, best_x)
    # a tie with the incumbent never replaces it
    assert not inc.update(st.words, st.temp_of, np.full_like(st.energy, best_e))
