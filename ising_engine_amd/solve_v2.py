"""solve_v2: rudy MaxCut file -> QUBO -> heat-bath anneal on the GPU -> best state as JSON.

The command-line contract of src/bin/solve_v2.rs:100-197:

    python -m ising_engine_amd.solve_v2 --file G1.txt --format rudy --seed 1 \\
        [--sweeps 200] [--replicas 256] [--wall-ms T]

prints {"state":"0101...","energy":E}.  Every replica starts from all zeros; one heat-bath
(Gibbs) colour sweep runs per temperature of the geometric ladder 4.0 -> 0.08 (QUBO energy
units, t <- t * ratio after each sweep, as the reference steps it); after every sweep the
lowest-energy replica (first on a tie) replaces the best state when strictly lower.  The
parse, the sweeps and the tracking run in the library (frontend.cpp, the heat-bath kernel,
the incumbent kernels); this module only drives them.  Replicas come in cells of 64, so
--replicas must be a multiple of 64.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import _lib


def ladder(sweeps: int, t_hi: float = 4.0, t_lo: float = 0.08) -> np.ndarray:
    """Temperatures of solve_v2.rs:151-158 and :187: t_0 = t_hi, t_{k+1} = t_k * ratio."""
    levels = max(sweeps, 1)
    ratio = (t_lo / t_hi) ** (1.0 / (levels - 1.0)) if levels > 1 else 1.0
    out = np.empty(levels)
    t = t_hi
    for k in range(levels):
        out[k] = t
        t *= ratio
    return out


def solve(ctx: _lib.Context, text: str, seed: int, sweeps: int = 200, replicas: int = 256,
          wall_ms: float | None = None):
    """Returns (state bytes in site order, best energy, sweeps run)."""
    if replicas < 64 or replicas % 64:
        raise ValueError("--replicas must be a positive multiple of 64")
    problem = _lib.Problem(text, _lib.FORMAT_RUDY)
    if not problem.info.integral:
        raise ValueError("instance is non-integral; the integer engine is not eligible")
    model = problem.model(ctx)
    cells = replicas // 64
    st = _lib.State(model, cells, seed, init_words=np.zeros((cells, problem.n), np.uint64),
                    kernel=_lib.KERNEL_GENERIC)
    betas = 1.0 / ladder(sweeps)
    sched = np.repeat(betas[:, None], cells, axis=1)
    if wall_ms is None:
        st.run(len(betas), beta_schedule=sched, track_best=True, heatbath=True)
        done = len(betas)
    else:
        t0 = time.perf_counter()
        done = 0
        for k in range(len(betas)):
            st.run(1, beta_schedule=sched[k:k + 1], track_best=True, heatbath=True)
            st.synchronize()
            done += 1
            if (time.perf_counter() - t0) * 1000.0 >= wall_ms:
                break
    x, e, _ = st.best()
    return x, e, done


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="solve_v2")
    ap.add_argument("--file")
    ap.add_argument("--format", default="rudy")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--sweeps", type=int, default=200)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--wall-ms", type=float, default=None)
    a = ap.parse_args(argv)
    if not a.file:
        print("usage: solve_v2 --file <path> --format rudy --seed N [--sweeps N] [--replicas N]",
              file=sys.stderr)
        return 2
    if a.format != "rudy":
        print(f"unsupported format: {a.format}", file=sys.stderr)
        return 2
    try:
        text = open(a.file).read()
    except OSError as e:
        print(f"cannot read {a.file}: {e}", file=sys.stderr)
        return 1
    ctx = _lib.Context(0)
    try:
        x, e, _ = solve(ctx, text, a.seed, a.sweeps, a.replicas, a.wall_ms)
    except (ValueError, _lib.IgeError) as err:
        print(str(err), file=sys.stderr)
        return 1
    print('{"state":"%s","energy":%d}' % ("".join("1" if b else "0" for b in x), e))
    return 0


if __name__ == "__main__":
    sys.exit(main())
