"""Benchmark: spin-flip attempts/s of the sweep + replica-exchange hot path on one MI355X.

Workload (default): 3D +-J Edwards-Anderson cubic lattice L=64 periodic
(N=262144, 6 neighbours), couplings +-1, 32 temperatures geometric T 0.5 -> 2.0,
64 replicas per temperature, one DEO swap attempt after every sweep.
A step = one Metropolis sweep of all 32 x 64 replicas + one exchange.

Prints ONE JSON line.  Inputs are resident in HBM before timing starts.
The roofline figure uses HIP events recorded on the state's own stream
around the sweep launches (ige_profile_*), averaged per launch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_attempt(cells: int, deg: float, w_b: int = 1, w_l: int = 1) -> float:
    """Algorithmic HBM bytes per attempt: per cell-site read+write of the u64 word (16 B),
    plus the model stream once per site shared by all cells (row pointer 4 B, deg x
    (4 B column + w_b coupling), w_l linear)."""
    return (16.0 * cells + 4 + deg * (4 + w_b) + w_l) / (64.0 * cells)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads(cells: int) -> int:
    """Host threads for a CPU baseline: the reference parallelises over cells only, so at
    most one per cell, and at most the 16 host threads one GPU's share of the box has."""
    return max(1, min(cells, 16, os.cpu_count() or 1))


def measured_counters(kernel: str, workload: str) -> dict:
    """The committed counter summary of `kernel` on `workload` (HBM bytes per launch from
    FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads for a 1 GiB copy; VALU
    busy), or {} when none matches."""
    for name in sorted(os.listdir(PROFILES)) if os.path.isdir(PROFILES) else []:
        if not (name.startswith("traffic_") and name.endswith(".json")):
            continue
        with open(os.path.join(PROFILES, name)) as f:
            t = json.load(f)
        if t.get("workload") == workload and t.get("kernel", "").startswith(f"{kernel}<"):
            return t
    return {}


def counter_probe(_lib, instances, L: int, temps: int, sweeps: int, workload: str = "pt",
                  waves: int = 8):
    """The program the counter passes run: `sweeps` PT steps of the workload (for workload
    "sk", sweeps of the dense SK N=4096 anneal; for "er", the first sweeps of the ER
    N=20000 anneal in one call (one chained launch); for "batch", sweeps [0, sweeps) of the
    4096-instance QUBO batch in one persistent launch of `waves` waves per workgroup), then
    one 1 GiB device-to-device copy whose FETCH_SIZE / WRITE_SIZE readings calibrate the
    counters' units."""
    import torch
    ctx = _lib.Context(0)
    if workload == "batch":
        wl = instances.workload_qubo_batch()
        st = _lib.Batch(ctx, wl["qubos"], wl["seeds"], wl["sweeps"], wl["beta_lo"],
                        wl["beta_hi"])
        st.run(_lib.BATCH_PERSISTENT, waves=waves, k0=0, k1=sweeps)
    elif workload == "er":
        wl = instances.workload_er_anneal()
        qb = wl["qubo"]
        m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
        st = _lib.State(m, 1, 3, kernel=_lib.KERNEL_MULTISPIN)
        sched = _lib.anneal_schedule(wl["beta_lo"], wl["beta_hi"], wl["sweeps"])
        st.run(sweeps, beta_schedule=sched[:sweeps, None])
    elif workload == "sk":
        wl = instances.workload_sk_anneal()
        qb = wl["qubo"]
        m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
        st = _lib.State(m, 1, 3)
        sched = _lib.anneal_schedule(wl["beta_lo"], wl["beta_hi"], wl["sweeps"])
        k0 = wl["sweeps"] // 2
        st.run(sweeps, beta_schedule=sched[k0:k0 + sweeps, None])
    else:
        wl = instances.workload_3d_pt(L=L, temps=temps)
        qb = wl["qubo"]
        m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
        st = _lib.State(m, wl["cells"], 0x5EED)
        st.set_ladder(wl["ladder"])
        st.run(sweeps, exchange=True)
    st.synchronize()
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] counter probe: {sweeps} steps of {wl['name']} and a 1 GiB copy done")


def cpu_baseline(qb, cells, ladder, budget_s: float):
    """The reference's production sweep (engine.rs::step port, OpenMP over cells) on a
    bounded sample of the same workload."""
    import ctypes
    from ctypes import POINTER, c_double, c_int, c_int8, c_uint32, c_uint64
    from oracle import binding as ob

    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "build", "libcpu_baseline.so"))
    lib.cb_run.argtypes = [c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double),
                           POINTER(c_double), c_uint32, POINTER(c_double), c_uint32, c_uint64,
                           c_int, POINTER(c_int8), POINTER(c_double), POINTER(c_int)]
    lib.cb_run.restype = c_int
    csr = ob.CSR.from_qubo(qb)
    n = qb.n
    rp = csr.row_ptr
    ci = csr.col_idx
    w = csr.w.astype(np.float64)
    lin = csr.linear.astype(np.float64)
    temps = np.ascontiguousarray(1.0 / np.asarray(ladder, dtype=np.float64))
    rng = np.random.default_rng(5)
    spins = rng.integers(0, 2, size=(cells, n, 64), dtype=np.int8)
    energies = np.zeros((cells, 64), dtype=np.float64)
    phase = c_int(0)
    threads = cpu_threads(cells)
    P = lambda a, t: a.ctypes.data_as(POINTER(t))
    args = lambda steps: (n, P(rp, c_uint32), P(ci, c_uint32), P(w, c_double), P(lin, c_double),
                          cells, P(temps, c_double), steps, 7, threads, P(spins, c_int8),
                          P(energies, c_double), ctypes.byref(phase))
    lib.cb_run(*args(1))  # warm-up step
    t0 = time.perf_counter()
    lib.cb_run(*args(1))
    one = time.perf_counter() - t0
    steps = max(1, int(budget_s / max(one, 1e-6)))
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        lib.cb_run(*args(max(1, steps // 3)))
        times.append((time.perf_counter() - t0) / max(1, steps // 3))
    per_step = float(np.median(times))
    attempts = n * 64 * cells
    return dict(value=attempts / per_step, unit="spin-flip attempts/s", cores=threads,
                kind="port", nproc=os.cpu_count(), cpu_model=cpu_model(),
                sample=f"engine.rs::step port (i8 lanes, f64 dE, xoshiro256++x8, fast_exp, DEO), "
                       f"same 3D L={round(n ** (1 / 3))} {cells}-temperature workload, "
                       f"{3 * max(1, steps // 3)} steps, median of 3 repetitions")


def cpu_sample(qb, cells, temps, budget_s: float) -> float:
    """engine.rs::step port on one thread per cell: attempts/s over about budget_s
    (None when budget_s is 0)."""
    if budget_s <= 0:
        return None
    import ctypes
    from ctypes import POINTER, c_double, c_int, c_int8, c_uint32, c_uint64
    from oracle import binding as ob

    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "build", "libcpu_baseline.so"))
    lib.cb_run.argtypes = [c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double),
                           POINTER(c_double), c_uint32, POINTER(c_double), c_uint32, c_uint64,
                           c_int, POINTER(c_int8), POINTER(c_double), POINTER(c_int)]
    lib.cb_run.restype = c_int
    csr = ob.CSR.from_qubo(qb)
    w = csr.w.astype(np.float64)
    lin = csr.linear.astype(np.float64)
    t = np.ascontiguousarray(temps, dtype=np.float64)
    spins = np.random.default_rng(5).integers(0, 2, size=(cells, qb.n, 64), dtype=np.int8)
    e = np.zeros((cells, 64), dtype=np.float64)
    phase = c_int(0)
    P = lambda a, ty: a.ctypes.data_as(POINTER(ty))
    steps, dt = 1, 0.0
    while True:
        t0 = time.perf_counter()
        lib.cb_run(qb.n, P(csr.row_ptr, c_uint32), P(csr.col_idx, c_uint32), P(w, c_double),
                   P(lin, c_double), cells, P(t, c_double), steps, 9, min(cells, 16),
                   P(spins, c_int8), P(e, c_double), ctypes.byref(phase))
        dt = time.perf_counter() - t0
        if dt >= budget_s / 4 or steps >= 1 << 20:
            break
        steps *= 4
    return steps * qb.n * 64 * cells / dt


def anneal_once(_lib, model, seed, sched, kernel=0, eager=False):
    """One 64-replica anneal over `sched` (one beta per sweep), timed on the device; the
    sweep loop replays the state's captured graph unless `eager`."""
    st = _lib.State(model, 1, seed, kernel=kernel)
    st.synchronize()
    t0 = time.perf_counter()
    st.run(len(sched), beta_schedule=sched[:, None], eager=eager)
    st.synchronize()
    dt = time.perf_counter() - t0
    if not np.array_equal(st.energies(), st.recompute_energies()):
        raise SystemExit("tracked energies drifted from recompute")
    return dt, st


def bench_workloads(_lib, instances, ctx, tts_runs: int, sk_sweeps: int, cpu_s: float) -> dict:
    """The single-cell anneal workloads: 2D +-J L=32 (1000 sweeps), Erdos-Renyi N=20000
    d=6 +-J (5000 sweeps, with time-to-target), dense SK N=4096 int8 (matrix-core fields)."""
    kname = {1: "generic", 2: "multispin", 3: "dense"}
    out = {}
    # 2D +-J L=32, beta 0.1 -> 3.0 over 1000 sweeps
    wl = instances.workload_2d_ej()
    qb = wl["qubo"]
    m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    sched = _lib.anneal_schedule(0.1, 3.0, 1000)
    anneal_once(_lib, m, 1, sched)
    dt, st = anneal_once(_lib, m, 2, sched)
    dt_eager, _ = anneal_once(_lib, m, 2, sched, eager=True)
    out[wl["name"]] = {"value": 1000 * qb.n * 64 / dt, "unit": "spin-flip attempts/s",
                       "seconds": dt, "seconds_eager_launches": dt_eager,
                       "kernel": kname[st.stats()[0].kernel],
                       "best_energy_per_spin": float(st.energies().min()) / qb.n,
                       "cpu_1thread": cpu_sample(qb, 1, [1.0], cpu_s)}
    log(f"[bench] {wl['name']}: {out[wl['name']]['value']:.3e} attempts/s ({dt * 1e3:.1f} ms)")
    # Erdos-Renyi N=20000, mean degree 6, +-J: one anneal (64 replicas), then TTS(0.99)
    # from R independent restarts run as R cells of one state (the uniforms are keyed by
    # the cell, so the restarts are independent chains), each cell following the schedule.
    wl = instances.workload_er_anneal()
    qb = wl["qubo"]
    m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    sched = _lib.anneal_schedule(wl["beta_lo"], wl["beta_hi"], wl["sweeps"])
    anneal_once(_lib, m, 1, sched[:50])
    dt1, st = anneal_once(_lib, m, 7, sched)
    dt1_eager, _ = anneal_once(_lib, m, 7, sched, eager=True)
    single = wl["sweeps"] * qb.n * 64 / dt1
    R = tts_runs
    st = _lib.State(m, R, 1000)
    st.synchronize()
    t0 = time.perf_counter()
    st.run(wl["sweeps"], beta_schedule=np.repeat(sched[:, None], R, axis=1))
    st.synchronize()
    dtR = time.perf_counter() - t0
    if not np.array_equal(st.energies(), st.recompute_energies()):
        raise SystemExit("tracked energies drifted from recompute")
    bests = st.energies().min(axis=1)
    t_run = dtR / R
    e_best = int(bests.min())
    target = e_best + int(0.005 * abs(e_best))
    p_s = float(np.mean(bests <= target))
    p_best = float(np.mean(bests <= e_best))
    tts = lambda p: (t_run if p >= 1.0 else (float("inf") if p <= 0.0 else
                                              t_run * np.log(0.01) / np.log(1.0 - p)))
    out[wl["name"]] = {"value": single, "unit": "spin-flip attempts/s",
                       "seconds_per_anneal": dt1, "seconds_eager_launches": dt1_eager,
                       "kernel": kname[st.stats()[0].kernel],
                       "restarts": R, "restarts_seconds": dtR,
                       "restarts_value": R * wl["sweeps"] * qb.n * 64 / dtR,
                       "t_run_per_restart_s": t_run, "best_energy": e_best,
                       "target_energy": target, "target": "within 0.5% of the best of all restarts",
                       "p_success": p_s, "tts99_s": tts(p_s),
                       "p_success_best": p_best, "tts99_best_s": tts(p_best),
                       "temperature_range": [1.0 / wl["beta_lo"], 1.0 / wl["beta_hi"]],
                       "cpu_1thread": cpu_sample(qb, 1, [1.0], cpu_s)}
    log(f"[bench] {wl['name']}: {single:.3e} attempts/s single anneal ({dt1:.3f} s); "
        f"{R} restarts in {dtR:.3f} s, TTS99 {tts(p_s):.3f} s (p_s {p_s:.2f}), "
        f"to best {tts(p_best):.3f} s")
    # Dense SK N=4096, int8 couplings: the matrix-core field recompute
    wl = instances.workload_sk_anneal()
    qb = wl["qubo"]
    m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    sched = _lib.anneal_schedule(wl["beta_lo"], wl["beta_hi"], wl["sweeps"])
    st = _lib.State(m, 1, 3)
    st.run(10, beta_schedule=sched[:10, None])
    st.synchronize()
    k0 = wl["sweeps"] // 2
    win = sched[k0:k0 + sk_sweeps, None]
    st.profile(True)
    t0 = time.perf_counter()
    st.run(len(win), beta_schedule=win)
    st.synchronize()
    dt = time.perf_counter() - t0
    ms, launches, _, _ = st.profile_read()
    st.profile(False)
    if not np.array_equal(st.energies(), st.recompute_energies()):
        raise SystemExit("tracked energies drifted from recompute")
    per_sweep = dt / len(win)
    # end-of-run polish (1-opt) of every replica of the window's last state, 64 states at once
    x64 = np.zeros((64, qb.n), dtype=np.uint8)
    words = st.spins()[0]
    for r in range(64):
        x64[r] = (words >> np.uint64(r)) & np.uint64(1)
    m.polish(x64[:1])
    t0 = time.perf_counter()
    _, e_pol, f_pol = m.polish(x64)
    dt_pol = time.perf_counter() - t0
    ops = 2.0 * qb.n * qb.n * 64  # int8 multiply-adds of one full field recompute, x2
    mfma_tops = ops / (ms / 1e3 / launches) / 1e12
    # The same window on R independent restarts (R cells of one state, one workgroup each,
    # one per CU): a sweep is one chain per slot, so restarts fill the other CUs.
    R = 256
    stR = _lib.State(m, R, 11)
    winR = np.repeat(win, R, axis=1)
    stR.run(2, beta_schedule=winR[:2])
    stR.synchronize()
    t0 = time.perf_counter()
    stR.run(len(winR), beta_schedule=winR)
    stR.synchronize()
    dtR = time.perf_counter() - t0
    restarts_value = R * qb.n * 64 * len(winR) / dtR
    restarts_tops = R * ops * len(winR) / dtR / 1e12
    stR.close()
    out[wl["name"]] = {"value": qb.n * 64 / per_sweep, "unit": "spin-flip attempts/s",
                       "ms_per_sweep": per_sweep * 1e3,
                       "projected_anneal_s": per_sweep * wl["sweeps"],
                       "sample": f"{len(win)} sweeps from the middle of the "
                                 f"{wl['sweeps']}-sweep schedule",
                       "kernel": kname[st.stats()[0].kernel],
                       "polish_64_states_s": dt_pol,
                       "polish_flips_max": int(f_pol.max()),
                       "polish_best_energy": int(e_pol.min()),
                       "best_energy_before_polish": int(st.energies().min()),
                       "roofline": {"bound": "mfma", "achieved": round(mfma_tops, 3),
                                    "peak": 5000.0, "unit": "TOP/s (int8)",
                                    "frac": mfma_tops / 5000.0,
                                    "frac_of_one_cu": mfma_tops / (5000.0 / 256),
                                    "ops_per_sweep": ops},
                       "restarts": R, "restarts_ms_per_sweep": dtR / len(winR) * 1e3,
                       "restarts_value": restarts_value,
                       "restarts_field_tops": restarts_tops,
                       "cpu_1thread": cpu_sample(qb, 1, [1.0], cpu_s)}
    log(f"[bench] {wl['name']}: {per_sweep * 1e3:.2f} ms/sweep, {mfma_tops:.2f} TOP/s int8; "
        f"{R} restarts {dtR / len(winR) * 1e3:.2f} ms/sweep, {restarts_value:.3e} attempts/s, "
        f"{restarts_tops:.1f} TOP/s int8")
    return out


def bench_descent(_lib, st, qb, cells: int, sweeps: int, steps: int = 200) -> dict:
    """SteepestDescent on the timed parallel-tempering state: `steps` steps, each one scan of
    every (slot, site, replica) and one flip per replica that can improve; and the cost per
    sweep of the incumbent update (one argmin launch after each sweep)."""
    n, nnz = qb.n, 2 * len(qb.q)
    st.synchronize()
    t0 = time.perf_counter()
    st.run(sweeps, exchange=True)
    st.synchronize()
    plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    st.run(sweeps, exchange=True, track_best=True)
    st.synchronize()
    tracked = time.perf_counter() - t0
    _, e_best, _ = st.best()
    before = st.energies().copy()
    st.steepest_descent(2)  # first-call allocations outside the timed region
    t0 = time.perf_counter()
    rep = st.steepest_descent(steps)
    dt = time.perf_counter() - t0
    after = st.energies()
    if not np.array_equal(after, st.recompute_energies()):
        raise SystemExit("tracked energies drifted from recompute after descent")
    done = rep.steps + rep.converged
    # one scan reads every slot's words once and the CSR rows once (rp, lin, order; ci, wt)
    scan_bytes = cells * n * 8 + n * 12 + nnz * 8
    log(f"[bench] descent: {rep.steps} steps, {rep.flips} flips in {dt * 1e3:.1f} ms "
        f"({dt / done * 1e6:.1f} us/step); incumbent update "
        f"{(tracked - plain) / sweeps * 1e6:.1f} us/sweep")
    return {"value": rep.proposed / dt, "unit": "site-replica scans/s (Report.proposed)",
            "steps": rep.steps, "converged": rep.converged, "flips": rep.flips,
            "seconds": dt, "us_per_step": dt / done * 1e6,
            "scan_bytes_per_step": scan_bytes,
            "mean_improvement_per_replica": float((before - after).mean()),
            "incumbent_energy": e_best,
            "incumbent_update_us_per_sweep": (tracked - plain) / sweeps * 1e6,
            "sweeps_timed_with_and_without_tracking": sweeps}


def bench_icm_pa(_lib, instances, ctx, st, qb, cells: int) -> dict:
    """ICM sweeps (cells x 32 Houdayer moves each) on the timed parallel-tempering state, and
    a PA-PT solve (4 populations x 8 temperatures, 50 stages of 10 PT steps, ICM after every
    stage) on the Erdos-Renyi N=20000 d=6 +-J workload."""
    from ising_engine_amd import population_annealing as pa_drv
    out = {}
    st.icm(1)
    st.synchronize()
    t0 = time.perf_counter()
    flipped = st.icm(10)
    dt = time.perf_counter() - t0
    if not np.array_equal(st.energies(), st.recompute_energies()):
        raise SystemExit("tracked energies drifted from recompute after ICM")
    out["icm_3d"] = {"ms_per_icm_sweep": dt / 10 * 1e3, "moves_per_sweep": cells * 32,
                     "mean_cluster_sites": flipped / (10 * cells * 32)}
    log(f"[bench] ICM: {dt / 10 * 1e3:.2f} ms per sweep of {cells * 32} moves, "
        f"mean cluster {flipped / (10 * cells * 32):.0f} sites")
    wl = instances.workload_er_anneal()
    q = wl["qubo"]
    m = _lib.Model(ctx, q.n, q.offset, q.linear, q.pi, q.pj, q.q)
    args = dict(num_pops=4, num_temps=8, stages=50, sweeps_per_stage=10,
                temp_max=1.0 / wl["beta_lo"], temp_min=1.0 / wl["beta_hi"], use_icm=True)
    pa_drv.solve(ctx, m, 1, **dict(args, stages=3))
    t0 = time.perf_counter()
    res = pa_drv.solve(ctx, m, 2, **args)
    dt = time.perf_counter() - t0
    if res.energy != int(q.energy(res.x)):
        raise SystemExit("PA-PT solution energy does not match its state")
    pt_steps = 50 * 10 + res.resample_events * 5
    out["pa_pt_er_n20000"] = {"seconds": dt, "best_energy": res.energy,
                              "resample_events": res.resample_events,
                              "min_ess": float(min(res.ess_history)),
                              "free_energy": res.free_energy, "pt_steps": pt_steps,
                              "attempts_per_s": pt_steps * q.n * 64 * 32 / dt}
    log(f"[bench] PA-PT ER N=20000: {dt:.2f} s, best {res.energy}, "
        f"{res.resample_events} resamplings")
    return out


def bench_gset(_lib, ctx, cpu_s: float) -> dict:
    """solve_v2 (rudy file -> MaxCut QUBO -> heat-bath anneal of 256 all-zero replicas over
    200 temperatures, best state tracked every sweep) on the four G-Set instances the
    reference ships; wall time of the whole solve (parse, upload, sweeps, result)."""
    from ising_engine_amd import solve_v2
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "gset")
    out = {}
    for name in ("G1", "G22", "G39", "G55"):
        text = open(os.path.join(gdir, f"{name}.txt")).read()
        solve_v2.solve(ctx, text, 1, 5, 256)
        t0 = time.perf_counter()
        x, e, done = solve_v2.solve(ctx, text, 1, 200, 256)
        dt = time.perf_counter() - t0
        prob = _lib.Problem(text, _lib.FORMAT_RUDY)
        if e != prob.energy(x):
            raise SystemExit(f"{name}: solve_v2 energy does not match its state")
        out[name] = {"n": prob.n, "edges": int(prob.info.npairs), "seconds": dt, "cut": -e,
                     "sweeps": done, "replicas": 256,
                     "attempts_per_s": done * prob.n * 256 / dt}
        log(f"[bench] solve_v2 {name}: cut {-e} in {dt * 1e3:.1f} ms")
    if cpu_s > 0:
        from oracle import binding as ob
        prob = _lib.Problem(open(os.path.join(gdir, "G1.txt")).read(), _lib.FORMAT_RUDY)
        off, lin, pi, pj, q = prob.pairs()
        csr = ob.CSR(prob.n, off, lin, pi, pj, q)
        st = ob.OracleState(csr, 4, 1, words=np.zeros((4, prob.n), np.uint64),
                            view=ob.VIEW_HEATBATH)
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_s:
            st.run(1, exchange=False, schedule=np.full((1, 4), 0.5))
            k += 1
        dt = time.perf_counter() - t0
        out["cpu_1thread_G1"] = {"value": k * prob.n * 256 / dt, "unit": "spin-flip attempts/s",
                                 "kind": "port", "cores": 1,
                                 "sample": f"{k} heat-bath sweeps of G1, 256 replicas"}
    return out


def cpu_batch_baseline(wl, budget_s: float) -> dict:
    """The reference's production sweep (engine.rs::step port) on the batch workload with
    instances spread across host threads, one instance per thread at a time (BASELINE's
    harness choice for the batch: the reference parallelises over cells only, and a batch
    instance is one cell).  A bounded sample: random instances of the batch, each run for a
    fixed number of sweeps at its mid-anneal temperature, until about budget_s of wall time."""
    import ctypes
    import threading
    from ctypes import POINTER, c_double, c_int, c_int8, c_uint32, c_uint64
    from oracle import binding as ob

    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "build", "libcpu_baseline.so"))
    lib.cb_run.argtypes = [c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double),
                           POINTER(c_double), c_uint32, POINTER(c_double), c_uint32, c_uint64,
                           c_int, POINTER(c_int8), POINTER(c_double), POINTER(c_int)]
    lib.cb_run.restype = c_int
    P = lambda a, t: a.ctypes.data_as(POINTER(t))
    threads = cpu_threads(10 ** 6)
    sweeps = 100
    rng = np.random.default_rng(29)
    order = rng.permutation(len(wl["qubos"]))
    lock = threading.Lock()
    done = {"attempts": 0, "instances": 0, "next": 0}
    t_end = time.perf_counter() + budget_s

    def worker(seed):
        while time.perf_counter() < t_end:
            with lock:
                if done["next"] >= len(order):
                    return
                i = int(order[done["next"]])
                done["next"] += 1
            qb = wl["qubos"][i]
            csr = ob.CSR.from_qubo(qb)
            w = csr.w[: csr.nnz].astype(np.float64)
            lin = csr.linear.astype(np.float64)
            temp = np.array([2.0 / (wl["beta_lo"][i] + wl["beta_hi"][i])])
            spins = np.random.default_rng(seed + i).integers(0, 2, size=(1, qb.n, 64),
                                                             dtype=np.int8)
            e = np.zeros((1, 64), dtype=np.float64)
            phase = c_int(0)
            lib.cb_run(qb.n, P(csr.row_ptr, c_uint32), P(csr.col_idx, c_uint32), P(w, c_double),
                       P(lin, c_double), 1, P(temp, c_double), sweeps, 13 + i, 1,
                       P(spins, c_int8), P(e, c_double), ctypes.byref(phase))
            with lock:
                done["attempts"] += qb.n * 64 * sweeps
                done["instances"] += 1

    t0 = time.perf_counter()
    pool = [threading.Thread(target=worker, args=(k,)) for k in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    dt = time.perf_counter() - t0
    return dict(value=done["attempts"] / dt, unit="spin-flip attempts/s", cores=threads,
                kind="port", nproc=os.cpu_count(), cpu_model=cpu_model(),
                sample=f"engine.rs::step port, {done['instances']} random instances of the "
                       f"batch x {sweeps} sweeps at the mid-anneal temperature, one instance "
                       f"per thread at a time, {dt:.1f} s wall")


def bench_batch(_lib, instances, ctx, count: int, cpu_s: float) -> dict:
    """Independent sparse QUBO instances (N log-uniform 64..8192, degree 8, int weights),
    2000-sweep anneals: one persistent launch vs plain per-instance launches.  After the
    timed launch every instance's tracked energies must equal a recompute from its words."""
    wl = instances.workload_qubo_batch(count=count)
    b = _lib.Batch(ctx, wl["qubos"], wl["seeds"], wl["sweeps"], wl["beta_lo"], wl["beta_hi"])
    waves, probe_ms = b.autotune(probe_sweeps=10)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(_lib.BATCH_PERSISTENT, waves=0)
    b.synchronize()
    dt = time.perf_counter() - t0
    persistent = b.total_attempts / dt
    launch_s = b.launch_ms() / 1e3
    tracked, recomputed = b.energies()
    if not np.array_equal(tracked, recomputed):
        bad = int(np.sum(np.any(tracked != recomputed, axis=1)))
        raise SystemExit(f"batch: tracked energies of {bad} instances differ from a recompute")
    inst_ms = b.timing_ms()
    ns = np.array([q.n for q in wl["qubos"]], dtype=np.int64)
    nnz = np.array([2 * len(q.q) for q in wl["qubos"]], dtype=np.int64)
    sweeps = wl["sweeps"][0]
    # algorithmic bytes: per (instance, site, sweep) the word in and out (16 B), the row
    # pointer (4 B), deg x (4 B column + 2 B int16 coupling), the int16 linear term (2 B)
    alg_bytes = sweeps * (22 * int(ns.sum()) + 6 * int(nnz.sum()))
    bpa = alg_bytes / b.total_attempts
    achieved = alg_bytes / launch_s / 1e9
    counters = measured_counters("persistent_anneal_kernel", wl["name"])
    # plain launches: the first 3 sweeps of every instance (the per-sweep cost of the plain
    # path does not depend on the sweep index), scaled to attempts/s
    b.reset()
    b.synchronize()
    t0 = time.perf_counter()
    b.run(_lib.BATCH_PLAIN, k0=0, k1=3)
    b.synchronize()
    dp = time.perf_counter() - t0
    plain = b.total_attempts * 3 / wl["sweeps"][0] / dp
    log(f"[bench] batch {count}: persistent {persistent:.4e} attempts/s ({dt:.2f} s, waves={waves}),"
        f" plain {plain:.4e} attempts/s -> {persistent / plain:.1f}x; energies == recompute "
        f"for all {count}; launch {launch_s * 1e3:.1f} ms, largest instance "
        f"{inst_ms[ns.argmax()]:.1f} ms, slowest {inst_ms.max():.1f} ms")
    b.close()
    cpu = cpu_batch_baseline(wl, cpu_s) if cpu_s > 0 else None
    if cpu:
        log(f"[bench] batch cpu baseline {cpu['value']:.3e} attempts/s on {cpu['cores']} threads")
    return {"workload": wl["name"], "instances": count, "sweeps": sweeps,
            "value": persistent, "unit": "spin-flip attempts/s", "seconds": dt,
            "plain_value": plain, "plain_sample": "first 3 sweeps of every instance",
            "ratio_vs_plain": persistent / plain, "waves_per_instance": waves,
            "autotune_ms": probe_ms,
            "energies_verified": f"tracked == device recompute for all {count} instances",
            "critical_path": {"largest_instance_ms": float(inst_ms[ns.argmax()]),
                              "slowest_instance_ms": float(inst_ms.max()),
                              "median_instance_ms": float(np.median(inst_ms)),
                              "sum_instance_ms": float(inst_ms.sum()),
                              "launch_ms": launch_s * 1e3},
            "roofline": dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS,
                             unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 5),
                             traffic=counters.get("traffic_bytes_per_launch"),
                             valu_busy=counters.get("valu_busy"),
                             algorithmic_bytes_per_launch=alg_bytes,
                             kernel=f"persistent_anneal_kernel<{waves}>",
                             avg_launch_us=round(launch_s * 1e6, 1),
                             bytes_per_attempt=round(bpa, 4)),
            "cpu_baseline": cpu}


def _serving_submit(srv, r, polish=True):
    return srv.submit(r["qubo"], r["seed"], r["sweeps"], track_every=r["sweeps_per_exchange"],
                      polish=polish, num_temps=r["num_temps"], temp_max=r["temp_max"],
                      temp_min=r["temp_min"])


def _serving_check(srv, reqs, ids, infos):
    """Every result's solution energy is exact and never worse than its best final replica."""
    for k in (0, len(ids) // 2, len(ids) - 1):
        x, e = srv.solution(ids[k])
        if e != int(reqs[k]["qubo"].energy(x)) or e > infos[k].best_energy:
            raise SystemExit("serving solution energy check failed")


def bench_serving(_lib, instances, ctx, rates, burst: int, count: int) -> dict:
    """Solve requests as POST /api/v1/solve runs them (UltimateSolver: 10 temperatures
    between the suggested temperatures, a DEO exchange every sweep, the incumbent every
    100-sweep exchange round, 1-opt polish; N log-uniform 128..16384, degree 4..16,
    weights in [-1000, 1000], 500..5000 sweeps), into the continuous-batching loop.

    Capacity: a burst of `burst` requests submitted at once; completions per second over
    the burst (its drain, the largest solves finishing on few workgroups, included, so
    this is a lower bound).  Latency: a Poisson stream of `count` requests at 80 % of that
    capacity, where the queue stays bounded: p50 / p99 and the largest queue seen.  Each
    rate in `rates` is reported as saturated (above capacity: the queue grows without
    bound, so no p99 exists) or run like the 80 % stream."""
    rng = np.random.default_rng(11)
    pool = [instances.serving_request(rng) for _ in range(max(burst, count))]
    per_req = np.array([r["qubo"].n * 64 * r["num_temps"] * r["sweeps"] for r in pool])
    srv = _lib.Server(ctx, hbm_bytes=32 << 30, round_ms=20.0)
    # warm-up: a few requests, outside the measurements
    for rid in [_serving_submit(srv, r) for r in pool[:4]]:
        srv.wait(rid, 600000.0)
    t0 = time.perf_counter()
    ids = [_serving_submit(srv, r) for r in pool[:burst]]
    infos = [srv.wait(rid, 600000.0) for rid in ids]
    t_burst = time.perf_counter() - t0
    if any(i.status != _lib.REQ_DONE for i in infos):
        raise SystemExit("a serving request did not complete")
    _serving_check(srv, pool[:burst], ids, infos)
    capacity = burst / t_burst
    cap_attempts = float(per_req[:burst].sum()) / t_burst
    log(f"[bench] serving capacity: {burst} requests in {t_burst:.2f} s -> {capacity:.1f} req/s, "
        f"{cap_attempts:.3e} attempts/s")

    def stream(rate, n):
        arrivals = np.cumsum(np.random.default_rng(12).exponential(1.0 / rate, size=n))
        reqs = pool[:n]
        ids, max_q = [], 0
        t0 = time.perf_counter()
        for a, r in zip(arrivals, reqs):
            while time.perf_counter() - t0 < a:
                time.sleep(0.0005)
            ids.append(_serving_submit(srv, r))
            if len(ids) % 16 == 0:
                max_q = max(max_q, int(srv.stats().queued))
        infos = [srv.wait(rid, 600000.0) for rid in ids]
        wall = time.perf_counter() - t0
        if any(i.status != _lib.REQ_DONE for i in infos):
            raise SystemExit("a serving request did not complete")
        _serving_check(srv, reqs, ids, infos)
        lat = np.array([i.latency_ms for i in infos])
        return {"rate_req_s": rate, "requests": n, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
                "max_queued": max_q, "wall_s": wall,
                "attempts_per_s": float(per_req[:n].sum()) / wall,
                "mean_solution_gain": float(np.mean([i.best_energy - i.solution_energy
                                                     for i in infos]))}

    held = stream(0.8 * capacity, count)
    log(f"[bench] serving at {held['rate_req_s']:.1f} req/s x {count}: p50 {held['p50_ms']:.0f} ms "
        f"p99 {held['p99_ms']:.0f} ms, max queued {held['max_queued']}")
    at_rates = []
    for rate in rates:
        if rate > capacity:
            at_rates.append({"rate_req_s": rate, "saturated": True,
                             "note": f"above the measured capacity {capacity:.1f} req/s: the "
                                     f"queue grows without bound"})
        else:
            at_rates.append(dict(stream(rate, count), saturated=False))
    st = srv.stats()
    srv.close()
    return {"requests": "UltimateSolver solves: 10 temperatures, DEO every sweep, incumbent "
                        "per 100-sweep round, 1-opt polish; N 128..16384, degree 4..16, "
                        "weights +-1000, 500..5000 sweeps",
            "mean_attempts_per_request": float(per_req.mean()),
            "capacity_req_s": capacity, "capacity_attempts_per_s": cap_attempts,
            "capacity_sample": f"burst of {burst} requests, {t_burst:.2f} s",
            "held": held, "rates": at_rates, "rounds": int(st.rounds),
            "max_live": int(st.max_live)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--prewarm-seconds", type=float, default=0.3,
                    help="untimed sweeps before the warmup steps, so that the GPU clocks have "
                         "ramped up from idle when a short timed region starts")
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--temps", type=int, default=32)
    ap.add_argument("--kernel", choices=["auto", "generic", "multispin"], default="auto")
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-workloads", action="store_true")
    ap.add_argument("--no-descent", action="store_true")
    ap.add_argument("--no-gset", action="store_true")
    ap.add_argument("--no-icm-pa", action="store_true")
    ap.add_argument("--tts-runs", type=int, default=32)
    ap.add_argument("--sk-sweeps", type=int, default=200)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--serving-rates", type=str, default="100,250,500,1000,2000",
                    help="comma-separated Poisson request rates (req/s); those above the "
                         "measured capacity are reported as saturated")
    ap.add_argument("--serving-burst", type=int, default=200,
                    help="requests in the capacity burst")
    ap.add_argument("--serving-requests", type=int, default=300,
                    help="requests per Poisson stream")
    ap.add_argument("--counter-probe", type=int, default=0, metavar="SWEEPS",
                    help="only run SWEEPS steps of --probe-workload and a 1 GiB copy "
                         "(for rocprofv3 --pmc passes)")
    ap.add_argument("--probe-workload", choices=["pt", "sk", "er", "batch"], default="pt")
    ap.add_argument("--probe-waves", type=int, default=8,
                    help="waves per workgroup of the batch counter probe")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    from ising_engine_amd import _lib, instances

    if args.counter_probe:
        counter_probe(_lib, instances, args.L, args.temps, args.counter_probe, args.probe_workload,
                      args.probe_waves)
        return
    kernel = {"auto": _lib.KERNEL_AUTO, "generic": _lib.KERNEL_GENERIC,
              "multispin": _lib.KERNEL_MULTISPIN}[args.kernel]
    wl = instances.workload_3d_pt(L=args.L, temps=args.temps)
    qb, cells, ladder = wl["qubo"], wl["cells"], wl["ladder"]
    log(f"[bench] building {wl['name']}: n={qb.n} bonds={len(qb.q)} cells={cells}")
    ctx = _lib.Context(0)
    model = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    st = _lib.State(model, cells, 0x5EED, kernel=kernel)
    st.set_ladder(ladder)
    kname = {1: "generic", 2: "multispin"}[st.stats()[0].kernel]
    log(f"[bench] kernel={kname} colours={model.info.num_colors}")

    # An idle GPU starts at low clocks: 20 steps right after 5 warmup steps measured
    # 0.124 ms a step against 0.109 ms once about 0.2 s of sweeps had run.
    prewarm_steps = 0
    t_pre = time.perf_counter()
    while time.perf_counter() - t_pre < args.prewarm_seconds:
        st.run(64, exchange=True)
        st.synchronize()
        prewarm_steps += 64
    st.run(args.warmup, exchange=True)
    st.synchronize()
    t0 = time.perf_counter()
    st.run(args.steps, exchange=True)
    st.synchronize()
    dt = time.perf_counter() - t0
    attempts = args.steps * qb.n * 64 * cells
    value = attempts / dt
    log(f"[bench] {args.steps} steps in {dt:.3f} s -> {value:.4e} attempts/s")

    # Parity guard on the timed state: tracked energies equal a recompute.
    if not np.array_equal(st.energies(), st.recompute_energies()):
        raise SystemExit("tracked energies drifted from recompute")

    # Per-launch kernel timing with HIP events on the state's stream.
    st.profile(True)
    st.run(args.profile_steps, exchange=True)
    st.synchronize()
    ms, launches, ms_x, n_x = st.profile_read()
    st.profile(False)
    avg_launch_s = ms / 1e3 / launches
    attempts_per_launch = qb.n * 64 * cells / model.info.num_colors
    deg = 2.0 * len(qb.q) / qb.n
    bpa = bytes_per_attempt(cells, deg)
    achieved = attempts_per_launch * bpa / avg_launch_s / 1e9
    log(f"[bench] sweep launch avg {avg_launch_s * 1e6:.1f} us over {launches} launches; "
        f"exchange avg {ms_x / max(n_x, 1) * 1e3:.1f} us; {bpa:.4f} B/attempt")
    counters = measured_counters(f"sweep_{kname}_kernel", wl["name"])
    roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 5),
                    traffic=counters.get("traffic_bytes_per_launch"),
                    valu_busy=counters.get("valu_busy"),
                    algorithmic_bytes_per_launch=round(attempts_per_launch * bpa),
                    kernel=f"sweep_{kname}_kernel",
                    avg_launch_us=round(avg_launch_s * 1e6, 2),
                    bytes_per_attempt=round(bpa, 4))

    descent = None
    if not args.no_descent:
        descent = bench_descent(_lib, st, qb, cells, args.profile_steps)

    icm_pa = None
    if not args.no_icm_pa:
        icm_pa = bench_icm_pa(_lib, instances, ctx, st, qb, cells)

    batched = None
    if not args.no_batch:
        batched = bench_batch(_lib, instances, ctx, args.batch_count,
                              cpu_s=0.0 if args.no_cpu_baseline else args.cpu_seconds)
    workloads = None
    if not args.no_workloads:
        workloads = bench_workloads(_lib, instances, ctx, args.tts_runs, args.sk_sweeps,
                                    cpu_s=0.0 if args.no_cpu_baseline else 2.0)
    gset = None
    if not args.no_gset:
        gset = bench_gset(_lib, ctx, cpu_s=0.0 if args.no_cpu_baseline else 2.0)
    serving = None
    if not args.no_serving:
        serving = bench_serving(_lib, instances, ctx,
                                [float(r) for r in args.serving_rates.split(",") if r],
                                args.serving_burst, args.serving_requests)

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(qb, cells, ladder, args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.4e} attempts/s on {cpu['cores']} threads")

    out = {
        "metric": "spin-flip attempts/s, batched vs plain on 1 MI355X; p50/p99 solve latency",
        "value": value,
        "unit": "spin-flip attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "prewarm_steps": prewarm_steps,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": wl["name"], "L": args.L, "n": qb.n, "temperatures": cells,
                   "replicas_per_temperature": 64, "exchange": "DEO every sweep",
                   "kernel": kname, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
        "workloads": workloads,
        "descent": descent,
        "gset_solve_v2": gset,
        "icm_pa": icm_pa,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
