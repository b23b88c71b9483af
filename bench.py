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


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads for a 1 GiB copy), or
    None when no summary matches."""
    for name in sorted(os.listdir(PROFILES)) if os.path.isdir(PROFILES) else []:
        if not (name.startswith("traffic_") and name.endswith(".json")):
            continue
        with open(os.path.join(PROFILES, name)) as f:
            t = json.load(f)
        if t.get("workload") == workload and t.get("kernel", "").startswith(f"{kernel}<"):
            return int(t["traffic_bytes_per_launch"])
    return None


def counter_probe(_lib, instances, L: int, temps: int, sweeps: int):
    """The program the HBM counter passes run: `sweeps` PT steps of the workload, then one
    1 GiB device-to-device copy whose FETCH_SIZE / WRITE_SIZE readings calibrate the
    counters' units."""
    import torch
    wl = instances.workload_3d_pt(L=L, temps=temps)
    qb = wl["qubo"]
    ctx = _lib.Context(0)
    m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    st = _lib.State(m, wl["cells"], 0x5EED)
    st.set_ladder(wl["ladder"])
    st.run(sweeps, exchange=True)
    st.synchronize()
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] counter probe: {sweeps} PT steps of {wl['name']} and a 1 GiB copy done")


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
    threads = max(1, min(cells, 16, os.cpu_count() or 1))
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
                kind="port",
                sample=f"engine.rs::step port (i8 lanes, f64 dE, xoshiro256++x8, fast_exp, DEO), "
                       f"same 3D L={round(n ** (1 / 3))} {cells}-temperature workload, "
                       f"{3 * max(1, steps // 3)} steps, median of 3 repetitions")


def bench_batch(_lib, instances, ctx, count: int) -> dict:
    """Independent sparse QUBO instances (N log-uniform 64..8192, degree 8, int weights),
    2000-sweep anneals: one persistent launch vs plain per-instance launches."""
    wl = instances.workload_qubo_batch(count=count)
    b = _lib.Batch(ctx, wl["qubos"], wl["seeds"], wl["sweeps"], wl["beta_lo"], wl["beta_hi"])
    waves, probe_ms = b.autotune(probe_sweeps=10)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(_lib.BATCH_PERSISTENT, waves=0)
    b.synchronize()
    dt = time.perf_counter() - t0
    persistent = b.total_attempts / dt
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
        f" plain {plain:.4e} attempts/s -> {persistent / plain:.1f}x")
    b.close()
    return {"workload": wl["name"], "instances": count, "sweeps": wl["sweeps"][0],
            "value": persistent, "unit": "spin-flip attempts/s", "seconds": dt,
            "plain_value": plain, "plain_sample": "first 3 sweeps of every instance",
            "ratio_vs_plain": persistent / plain, "waves_per_instance": waves}


def bench_serving(_lib, instances, ctx, rate: float, count: int) -> dict:
    """Poisson stream of mixed Ising/QUBO solve requests (N log-uniform 128..16384, degree
    4..16, weights in [-1000, 1000], 500..5000 sweeps) into the continuous-batching loop."""
    rng = np.random.default_rng(11)
    reqs = [instances.serving_request(rng) for _ in range(count)]
    srv = _lib.Server(ctx, hbm_bytes=16 << 30, round_ms=10.0)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    t0 = time.perf_counter()
    ids = []
    for a, r in zip(arrivals, reqs):
        while time.perf_counter() - t0 < a:
            time.sleep(0.0005)
        ids.append(srv.submit(r["qubo"], r["seed"], r["sweeps"], r["beta_lo"], r["beta_hi"]))
    lat = np.array([srv.wait(rid, 600000.0).latency_ms for rid in ids])
    wall = time.perf_counter() - t0
    st = srv.stats()
    attempts = sum(r["qubo"].n * 64 * r["sweeps"] for r in reqs)
    srv.close()
    out = {"rate_req_s": rate, "requests": count, "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
           "attempts_per_s": attempts / wall, "wall_s": wall, "rounds": int(st.rounds),
           "max_live": int(st.max_live)}
    log(f"[bench] serving {rate}/s x {count}: p50 {out['p50_ms']:.1f} ms p99 {out['p99_ms']:.1f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--temps", type=int, default=32)
    ap.add_argument("--kernel", choices=["auto", "generic", "multispin"], default="auto")
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--serving-rate", type=float, default=20.0)
    ap.add_argument("--serving-requests", type=int, default=300)
    ap.add_argument("--counter-probe", type=int, default=0, metavar="SWEEPS",
                    help="only run SWEEPS PT steps and a 1 GiB copy (for rocprofv3 --pmc passes)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    from ising_engine_amd import _lib, instances

    if args.counter_probe:
        counter_probe(_lib, instances, args.L, args.temps, args.counter_probe)
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
    traffic = measured_traffic(f"sweep_{kname}_kernel", wl["name"])
    roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic,
                    algorithmic_bytes_per_launch=round(attempts_per_launch * bpa),
                    kernel=f"sweep_{kname}_kernel",
                    avg_launch_us=round(avg_launch_s * 1e6, 2),
                    bytes_per_attempt=round(bpa, 4))

    batched = None
    if not args.no_batch:
        batched = bench_batch(_lib, instances, ctx, args.batch_count)
    serving = None
    if not args.no_serving:
        serving = bench_serving(_lib, instances, ctx, args.serving_rate, args.serving_requests)

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
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
