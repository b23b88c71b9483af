"""Serving loop: continuous batching with HBM-footprint admission.

A request's result must be the same alone and in any mix (bit-exact words and
energies), and equal to a single-state run of the same seed and schedule.
"""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances

pytestmark = pytest.mark.gpu


def requests(count, seed, **kw):
    rng = np.random.default_rng(seed)
    return [instances.serving_request(rng, **kw) for _ in range(count)]


def submit(srv, r):
    return srv.submit(r["qubo"], r["seed"], r["sweeps"], r["beta_lo"], r["beta_hi"])


def result(srv, rid):
    info = srv.wait(rid, 120000.0)
    assert info.status == _lib.REQ_DONE, info.status
    info, words, e = srv.poll(rid, release=True)
    return info, words, e


def test_same_result_alone_and_in_the_mix(ctx):
    reqs = requests(24, 1, n_lo=64, n_hi=2048, sweeps_lo=20, sweeps_hi=120)
    target = reqs[5]
    srv = _lib.Server(ctx, hbm_bytes=256 << 20, round_ms=2.0)
    alone = result(srv, submit(srv, target))
    ids = [submit(srv, r) for r in reqs]           # the mix, including the target again
    mixed = {i: result(srv, rid) for i, rid in enumerate(ids)}
    np.testing.assert_array_equal(mixed[5][1], alone[1])
    np.testing.assert_array_equal(mixed[5][2], alone[2])
    assert mixed[5][0].best_energy == alone[0].best_energy
    for i, r in enumerate(reqs):                   # every result is a true anneal end state
        qb = r["qubo"]
        x = ((mixed[i][1][None, :] >> np.arange(64, dtype=np.uint64)[:, None]) & 1).astype(np.int64)
        np.testing.assert_array_equal(qb.energy(x), mixed[i][2])
    st = srv.stats()
    assert st.completed == 25 and st.live == 0 and st.queued == 0 and st.arena_used == 0
    srv.close()


def test_matches_single_state_with_the_schedule(ctx):
    reqs = requests(3, 2, n_lo=100, n_hi=600, sweeps_lo=30, sweeps_hi=60)
    srv = _lib.Server(ctx, hbm_bytes=64 << 20, round_ms=1.0)
    ids = [submit(srv, r) for r in reqs]
    for r, rid in zip(reqs, ids):
        info, words, e = result(srv, rid)
        qb = r["qubo"]
        m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
        st = _lib.State(m, 1, r["seed"], kernel=_lib.KERNEL_GENERIC)
        sched = _lib.anneal_schedule(r["beta_lo"], r["beta_hi"], r["sweeps"])
        st.run(r["sweeps"], beta_schedule=sched[:, None])
        np.testing.assert_array_equal(st.spins()[0], words)
        np.testing.assert_array_equal(st.energies()[0], e)
        assert info.best_energy == e.min() and e[info.best_replica] == e.min()
    srv.close()


def test_admission_by_hbm_footprint(ctx):
    reqs = requests(12, 3, n_lo=1500, n_hi=2000, sweeps_lo=10, sweeps_hi=20)
    big = _lib.Server(ctx, hbm_bytes=512 << 20, round_ms=1.0)
    ref = [result(big, submit(big, r)) for r in reqs]
    big.close()
    small = _lib.Server(ctx, hbm_bytes=1 << 20, round_ms=1.0)   # room for a few at a time
    ids = [submit(small, r) for r in reqs]
    got = [result(small, rid) for rid in ids]
    st = small.stats()
    assert 1 <= st.max_live < len(reqs)
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_array_equal(a[2], b[2])
    huge = requests(1, 4, n_lo=40000, n_hi=40000, sweeps_lo=5, sweeps_hi=5)[0]
    with pytest.raises(_lib.IgeError) as ei:
        submit(small, huge)
    assert ei.value.code == _lib.IGE_ERR_NOMEM
    small.close()


def test_request_validation_mirrors_the_reference(ctx):
    srv = _lib.Server(ctx, hbm_bytes=16 << 20, round_ms=1.0)
    r = requests(1, 5, n_lo=100, n_hi=100, sweeps_lo=10, sweeps_hi=10)[0]
    with pytest.raises(_lib.IgeError):
        srv.submit(r["qubo"], 1, 10001, r["beta_lo"], r["beta_hi"])   # sweeps > 10000
    qb = instances.Qubo(n=50001, offset=0, linear=np.zeros(50001, np.int64),
                        pi=np.zeros(0, np.uint32), pj=np.zeros(0, np.uint32), q=np.zeros(0, np.int64))
    with pytest.raises(_lib.IgeError):
        srv.submit(qb, 1, 10, 0.1, 1.0)                               # num_vars > 50000
    srv.close()


def test_gangs_in_the_serving_loop(ctx):
    reqs = requests(10, 6, n_lo=500, n_hi=3000, sweeps_lo=20, sweeps_hi=60)
    plain = _lib.Server(ctx, hbm_bytes=256 << 20, round_ms=2.0)
    plain.configure(max_gang=1)
    ref = [result(plain, submit(plain, r)) for r in reqs]
    plain.close()
    gang = _lib.Server(ctx, hbm_bytes=256 << 20, round_ms=2.0)
    gang.configure(max_gang=8, gang_sites=256)
    ids = [submit(gang, r) for r in reqs]
    for (i, rid) in enumerate(ids):
        got = result(gang, rid)
        np.testing.assert_array_equal(got[1], ref[i][1])
        np.testing.assert_array_equal(got[2], ref[i][2])
    gang.close()


def reference_solution(ctx, r, track_every, polish):
    """UltimateSolver-style solution from a one-cell state: the anneal in chunks ending at
    every multiple of track_every and at the last sweep, an incumbent update after each chunk
    (or the best final replica), then the 1-opt polish of the model API."""
    qb = r["qubo"]
    m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    st = _lib.State(m, 1, r["seed"], kernel=_lib.KERNEL_GENERIC)
    sched = _lib.anneal_schedule(r["beta_lo"], r["beta_hi"], r["sweeps"])
    S = r["sweeps"]
    if track_every:
        k = 0
        while k < S:
            step = min(track_every - k % track_every, S - k)
            st.run(step, beta_schedule=sched[k:k + step, None])
            k += step
            st.update_best()
        x, e, sw = st.best()
    else:
        st.run(S, beta_schedule=sched[:, None])
        words, es = st.spins()[0], st.energies()[0]
        lane = int(np.argmin(es))
        x = ((words >> np.uint64(lane)) & np.uint64(1)).astype(np.uint8)
        e, sw = int(es[lane]), S
    flips = 0
    if polish:
        xp, ep, fp = m.polish(x)
        x, e, flips = xp[0], int(ep[0]), int(fp[0])
    return x, e, sw, flips


@pytest.mark.parametrize("gangs", [False, True], ids=["one_workgroup", "gangs"])
def test_solution_is_best_seen_then_polished(ctx, gangs):
    reqs = requests(6, 7, n_lo=200, n_hi=3000, sweeps_lo=20, sweeps_hi=70)
    modes = [(1, True), (7, True), (25, False), (0, True), (10, True), (0, False)]
    srv = _lib.Server(ctx, hbm_bytes=256 << 20, round_ms=1.0)
    if gangs:
        srv.configure(max_gang=4, gang_sites=256)
    print(f"[serving solution] gangs={gangs}", flush=True)
    ids = [srv.submit(r["qubo"], r["seed"], r["sweeps"], r["beta_lo"], r["beta_hi"],
                      track_every=t, polish=p) for r, (t, p) in zip(reqs, modes)]
    for r, (t, p), rid in zip(reqs, modes, ids):
        info = srv.wait(rid, 120000.0)
        assert info.status == _lib.REQ_DONE, info.status
        x, e = srv.solution(rid)
        xr, er, swr, fr = reference_solution(ctx, r, t, p)
        np.testing.assert_array_equal(x, xr)
        assert e == er == info.solution_energy == int(r["qubo"].energy(x))
        assert info.solution_sweep == swr and info.polish_flips == fr
        if t:
            assert e <= info.best_energy  # best seen is never worse than the final best
    srv.close()
