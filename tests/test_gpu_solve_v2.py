"""solve_v2 on the device against the oracle: the heat-bath anneal over the reference's
temperature ladder from all-zero replicas, best state tracked after every sweep; the
state string and its energy must match bit for bit."""
import json
import os

import numpy as np
import pytest

from ising_engine_amd import _lib, solve_v2
from oracle import binding as ob
from oracle import frontends as fr

pytestmark = pytest.mark.gpu
GSET = os.path.join(os.path.dirname(__file__), "golden", "gset")


def oracle_solve(text, seed, sweeps, replicas):
    m = fr.parse_rudy_engine(text)
    keys = list(m["pairs"].keys())
    csr = ob.CSR(m["n"], int(m["offset"]), [int(v) for v in m["linear"]],
                 [k[0] for k in keys], [k[1] for k in keys], [int(m["pairs"][k]) for k in keys])
    cells = replicas // 64
    st = ob.OracleState(csr, cells, seed, words=np.zeros((cells, m["n"]), np.uint64),
                        view=ob.VIEW_HEATBATH)
    inc = ob.Incumbent(m["n"])
    for t in solve_v2.ladder(sweeps):
        st.run(1, exchange=False, schedule=np.full((1, cells), 1.0 / t))
        inc.update(st.words, st.temp_of, st.energy)
    return inc.x, inc.energy.value


def random_rudy(n, m, seed):
    rng = np.random.default_rng(seed)
    lines = [f"{n} {m}"]
    for _ in range(m):
        u, v = rng.choice(n, size=2, replace=False) + 1
        lines.append(f"{u} {v} {int(rng.integers(-3, 4)) or 1}")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("name,text_fn,seed,sweeps,replicas", [
    ("rudy_n60_weighted", lambda: random_rudy(60, 240, 1), 7, 200, 128),
    ("G1_40_sweeps", lambda: open(os.path.join(GSET, "G1.txt")).read(), 3, 40, 256),
])
def test_solve_v2_matches_oracle(ctx, name, text_fn, seed, sweeps, replicas):
    print(f"[solve_v2] {name}", flush=True)
    text = text_fn()
    x, e, done = solve_v2.solve(ctx, text, seed, sweeps, replicas)
    xo, eo = oracle_solve(text, seed, sweeps, replicas)
    assert done == sweeps
    np.testing.assert_array_equal(x, xo)
    assert e == eo == _lib.Problem(text, _lib.FORMAT_RUDY).energy(x)


def test_solve_v2_cli(ctx, capsys):
    path = os.path.join(GSET, "G1.txt")
    assert solve_v2.main(["--file", path, "--format", "rudy", "--seed", "5", "--sweeps", "60"]) == 0
    out = json.loads(capsys.readouterr().out)
    x = np.array([int(c) for c in out["state"]], dtype=np.uint8)
    p = _lib.Problem(open(path).read(), _lib.FORMAT_RUDY)
    assert x.size == 800 and out["energy"] == p.energy(x) < -10000  # a cut above 10000
