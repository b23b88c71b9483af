"""Serving loop: continuous batching with HBM-footprint admission.

A request's result must be the same alone and in any mix (bit-exact words and
energies), and equal to a single-state run of the same seed: the anneal schedule for a
one-cell request, and for a request that runs UltimateSolver's loop
(src/solver/ultimate.rs:234-374) the sequence ige_run(IGE_RUN_EXCHANGE) per exchange
round, ige_update_best after each, the reference's choice between the incumbent and the
best final replica, and ige_model_polish.
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
    """(info, words, energies); a one-cell request's as [n] and [64]."""
    info = srv.wait(rid, 120000.0)
    assert info.status == _lib.REQ_DONE, info.status
    info, words, e = srv.poll(rid, release=True)
    if words.shape[0] == 1:
        words, e = words[0], e[0]
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
    else:
        st.run(S, beta_schedule=sched[:, None])
    # the best final replica (lowest lane on a tie) unless the incumbent is strictly lower
    words, es = st.spins()[0], st.energies()[0]
    lane = int(np.argmin(es))
    x = ((words >> np.uint64(lane)) & np.uint64(1)).astype(np.uint8)
    e, sw = int(es[lane]), S
    if track_every:
        xb, eb, swb = st.best()
        if eb < e:
            x, e, sw = xb, eb, swb
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


# ---------------------------------------------------------------------------
# Requests that run UltimateSolver's loop (num_temps cells, DEO every sweep)
# ---------------------------------------------------------------------------
def submit_pt(srv, r, polish=True):
    return srv.submit(r["qubo"], r["seed"], r["sweeps"], track_every=r["sweeps_per_exchange"],
                      polish=polish, num_temps=r["num_temps"], temp_max=r["temp_max"],
                      temp_min=r["temp_min"])


def ultimate_sequence(ctx, r, polish=True):
    """The same solve on a single state: num_temps cells on ige_ultimate_ladder, per
    exchange round ige_run(sweeps_per_exchange, IGE_RUN_EXCHANGE) then ige_update_best;
    the best final replica by (energy, temperature index, lane), the incumbent instead
    when strictly lower (ultimate.rs:339-373); then the 1-opt polish."""
    qb = r["qubo"]
    m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    T = r["num_temps"]
    st = _lib.State(m, T, r["seed"], kernel=_lib.KERNEL_GENERIC)
    st.set_ladder(_lib.ultimate_ladder(r["temp_max"], r["temp_min"], T))
    spe = r["sweeps_per_exchange"]
    for _ in range(r["sweeps"] // spe):
        st.run(spe, exchange=T > 1)
        st.update_best()
    words, es, tof = st.spins(), st.energies(), st.temp_of()
    key = tof.astype(np.int64) * 64 + np.arange(64)[None, :]
    order = np.lexsort((key.reshape(-1), es.reshape(-1)))
    q = int(order[0])
    x = ((words[q // 64] >> np.uint64(q % 64)) & np.uint64(1)).astype(np.uint8)
    e, sw = int(es.reshape(-1)[q]), r["sweeps"]
    xb, eb, swb = st.best()
    if eb < e:
        x, e, sw = xb, eb, swb
    flips = 0
    if polish:
        xp, ep, fp = m.polish(x)
        x, e, flips = xp[0], int(ep[0]), int(fp[0])
    return words, es, x, e, sw, flips


# (max_gang, gang_sites): the defaults (cells owned one each, splits of large
# requests into idle CUs), fine splits, members owning several cells, one workgroup
GANG_CONFIGS = {"default": None, "split": (16, 256), "owners": (4, 256), "single": (1, 256)}


@pytest.mark.parametrize("gangs", list(GANG_CONFIGS), ids=list(GANG_CONFIGS))
def test_pt_request_is_the_ultimate_solver_loop(ctx, gangs):
    rng = np.random.default_rng(21)
    shapes = [(10, 25, 3), (10, 7, 4), (2, 13, 2), (1, 10, 3), (5, 1, 9), (16, 20, 2)]
    reqs = []
    for T, spe, rounds in shapes:
        r = instances.serving_request(rng, n_lo=150, n_hi=2500, num_temps=T,
                                      sweeps_per_exchange=spe, sweeps_lo=spe, sweeps_hi=spe)
        r["sweeps"] = spe * rounds
        reqs.append(r)
    # 2048 sites x 10 cells: split into 6 workgroups per cell (60 members) under
    # "split", into 2 under the defaults
    r = instances.serving_request(rng, n_lo=2048, n_hi=2048, num_temps=10,
                                  sweeps_per_exchange=10, sweeps_lo=10, sweeps_hi=10)
    r["sweeps"] = 30
    reqs.append(r)
    srv = _lib.Server(ctx, hbm_bytes=512 << 20, round_ms=1.0)
    if GANG_CONFIGS[gangs]:
        srv.configure(*GANG_CONFIGS[gangs])
    print(f"[serving pt] gangs={gangs}", flush=True)
    ids = [submit_pt(srv, r, polish=(i % 3 != 1)) for i, r in enumerate(reqs)]
    for i, (r, rid) in enumerate(zip(reqs, ids)):
        info = srv.wait(rid, 120000.0)
        assert info.status == _lib.REQ_DONE, info.status
        assert info.cells == r["num_temps"]
        x, e = srv.solution(rid)
        _, words, es = srv.poll(rid)
        w_ref, e_ref, xr, er, swr, fr = ultimate_sequence(ctx, r, polish=(i % 3 != 1))
        np.testing.assert_array_equal(words, w_ref, err_msg=f"request {i}")
        np.testing.assert_array_equal(es, e_ref, err_msg=f"request {i}")
        np.testing.assert_array_equal(x, xr, err_msg=f"request {i}")
        assert e == er == info.solution_energy == int(r["qubo"].energy(x))
        assert info.solution_sweep == swr and info.polish_flips == fr
    srv.close()


def test_pt_requests_same_alone_and_in_the_mix(ctx):
    rng = np.random.default_rng(22)
    reqs = [instances.serving_request(rng, n_lo=100, n_hi=3000, sweeps_lo=40, sweeps_hi=160,
                                      sweeps_per_exchange=20) for _ in range(16)]
    srv = _lib.Server(ctx, hbm_bytes=512 << 20, round_ms=2.0)
    alone = result(srv, submit_pt(srv, reqs[3]))
    ids = [submit_pt(srv, r) for r in reqs]
    for i, rid in enumerate(ids):
        srv.wait(rid, 120000.0)
    info, words, e = srv.poll(ids[3])
    np.testing.assert_array_equal(words, alone[1])
    np.testing.assert_array_equal(e, alone[2])
    assert info.solution_energy == alone[0].solution_energy
    for i, r in enumerate(reqs):
        info, words, e = srv.poll(ids[i])
        x = ((words[:, None, :] >> np.arange(64, dtype=np.uint64)[None, :, None]) & 1)
        np.testing.assert_array_equal(r["qubo"].energy(x.reshape(-1, r["qubo"].n).astype(np.int64)),
                                      e.reshape(-1))
    srv.close()


def test_pt_requests_at_the_serving_sizes(ctx):
    """The serving workload's own ranges (N log-uniform up to 16384, degree 4..16, weights
    in [-1000, 1000], 10 temperatures, 100 sweeps per exchange round) with the sweep budget
    cut to 500: alone, in the mix and against the single-state sequence."""
    rng = np.random.default_rng(23)
    reqs = [instances.serving_request(rng, sweeps_lo=500, sweeps_hi=500) for _ in range(5)]
    big = instances.serving_request(np.random.default_rng(24), n_lo=16384, n_hi=16384,
                                    sweeps_lo=500, sweeps_hi=500)
    reqs.append(big)
    print("[serving pt] sizes " + ", ".join(str(r["qubo"].n) for r in reqs), flush=True)
    srv = _lib.Server(ctx, hbm_bytes=2 << 30, round_ms=10.0)
    alone = result(srv, submit_pt(srv, big))
    ids = [submit_pt(srv, r) for r in reqs]
    for rid in ids:
        assert srv.wait(rid, 300000.0).status == _lib.REQ_DONE
    info, words, e = srv.poll(ids[-1])
    np.testing.assert_array_equal(words, alone[1])
    np.testing.assert_array_equal(e, alone[2])
    for i in (0, len(reqs) - 1):
        r = reqs[i]
        print(f"[serving pt] single-state check n={r['qubo'].n}", flush=True)
        x, en = srv.solution(ids[i])
        _, words, es = srv.poll(ids[i])
        w_ref, e_ref, xr, er, _, _ = ultimate_sequence(ctx, r)
        np.testing.assert_array_equal(words, w_ref)
        np.testing.assert_array_equal(es, e_ref)
        np.testing.assert_array_equal(x, xr)
        assert en == er == int(r["qubo"].energy(x))
    srv.close()


def test_pt_request_validation(ctx):
    srv = _lib.Server(ctx, hbm_bytes=64 << 20, round_ms=1.0)
    r = requests(1, 25, n_lo=100, n_hi=100)[0]
    qb = r["qubo"]
    with pytest.raises(_lib.IgeError):   # track_every must divide sweeps
        srv.submit(qb, 1, 250, track_every=100, num_temps=4, temp_max=5.0, temp_min=0.5)
    with pytest.raises(_lib.IgeError):   # more than 10000 sweeps per exchange
        srv.submit(qb, 1, 20002, track_every=10001, num_temps=4, temp_max=5.0, temp_min=0.5)
    with pytest.raises(_lib.IgeError):   # temperatures must be > 0
        srv.submit(qb, 1, 100, track_every=10, num_temps=4, temp_max=5.0, temp_min=0.0)
    with pytest.raises(_lib.IgeError):   # at most 64 temperature cells
        srv.submit(qb, 1, 100, track_every=10, num_temps=65, temp_max=5.0, temp_min=0.5)
    srv.close()
