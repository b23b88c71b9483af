"""Runs the C++ operator-registry tests (tests/cpp/test_engine_v2.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "test_engine_v2")


def _run(*args):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (build() compiles it)")
    p = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr


def test_registry_and_capability_queries():
    _run("--cpu")


@pytest.mark.gpu
def test_gpu_operators_through_the_registry():
    _run()


def _fnv_digest(words_logical, n, cells):
    """state.rs:55-67 over logical words [cells][n] (bit b of word c = replica c*64+b)."""
    h = 0xCBF29CE484222325
    for i in range(n):
        for c in range(cells):
            v = int(words_logical[c, i])
            for k in range(8):
                h ^= (v >> (8 * k)) & 0xFF
                h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.gpu
def test_operator_layer_matches_the_oracle(tmp_path):
    """The registered GPU operators and the host per-site calls of GpuBitSlice, in the
    reference's operator-test shape (frustrated 7-site instance, 128 copies of 1010010,
    geometric_ladder(64, 3.0, 0.3) read as temps[rep % 64]): Gibbs 5, exchange 6,
    Metropolis 30, exchange 3, apply_flips / swap_replicas / copy_replica on the host,
    Metropolis 4.  Every Report, configuration, energy and the digest equal the oracle's."""
    import numpy as np
    from oracle import binding as ob

    nrep, seed = 128, 0x5EED1234
    temps = ob.geometric_ladder(64, 3.0, 0.3)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.uint64(seed).tobytes() + np.uint32(nrep).tobytes()
                + np.uint32(temps.size).tobytes() + temps.astype(np.float64).tobytes())
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (build() compiles it)")
    p = subprocess.run([BIN, "--dump", str(inp), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = out.read_bytes()
    reps = np.frombuffer(raw[:80], dtype=np.uint64).reshape(5, 2)
    x = np.frombuffer(raw[80:80 + nrep * 7], dtype=np.uint8).reshape(nrep, 7)
    e = np.frombuffer(raw[80 + nrep * 7:80 + nrep * 15], dtype=np.float64)
    digest, = np.frombuffer(raw[80 + nrep * 15:88 + nrep * 15], dtype=np.uint64)
    audit, = np.frombuffer(raw[88 + nrep * 15:96 + nrep * 15], dtype=np.float64)

    csr = ob.CSR(7, -2, [1, -2, 3, 0, -1, 2, 1], [0, 1, 2, 0, 3, 4, 5, 2],
                 [1, 2, 3, 3, 4, 5, 6, 6], [-1, 2, -3, 1, -1, 2, -2, 1])
    cells = nrep // 64
    init = np.array([1, 0, 1, 0, 0, 1, 0], dtype=np.uint64)
    st = ob.OracleState(csr, cells, seed, words=np.tile(init * np.uint64(0xFFFFFFFFFFFFFFFF),
                                                        (cells, 1)))
    st.set_replica_temperatures(temps)
    expect = []

    def step(fn):
        a0, p0 = int(st.stats[0]), int(st.stats[1])
        x0, c0 = int(st.stats[2::2].sum()), int(st.stats[3::2].sum())
        fn()
        da, dp = int(st.stats[0]) - a0, int(st.stats[1]) - p0
        dx, dc = int(st.stats[2::2].sum()) - x0, int(st.stats[3::2].sum()) - c0
        expect.append((da, dp) if dp else (dc, dx))

    def sweeps(view, k):
        st.view = view
        st.run(k, exchange=False)

    step(lambda: sweeps(ob.VIEW_HEATBATH, 5))
    step(lambda: st.exchange_replicas(6))
    step(lambda: sweeps(ob.VIEW_WORD, 30))
    step(lambda: st.exchange_replicas(3))

    def loc():
        lo = np.empty(nrep, dtype=np.int64)
        lo[st.temp_of.reshape(-1)] = np.arange(nrep)
        return lo

    def get(q):
        return (st.words[q // 64] >> np.uint64(q % 64)) & np.uint64(1)

    def put(q, bits):
        m = np.uint64(1) << np.uint64(q % 64)
        st.words[q // 64] = (st.words[q // 64] & ~m) | (bits.astype(np.uint64) << np.uint64(q % 64))

    lo = loc()
    for r in (0, 5, nrep - 1):  # apply_flips(site 2, {0, 5, last})
        st.words[lo[r] // 64, 2] ^= np.uint64(1) << np.uint64(lo[r] % 64)
    a, b = get(lo[3]).copy(), get(lo[4]).copy()  # swap_replicas(3, 4)
    put(lo[3], b)
    put(lo[4], a)
    put(lo[11], get(lo[10]).copy())  # copy_replica(10, 11)
    st.energy = csr.energies(st.words)
    step(lambda: sweeps(ob.VIEW_WORD, 4))

    assert [tuple(int(v) for v in r) for r in reps] == expect
    lo = loc()
    xo = np.array([[int(get(lo[r])[i]) for i in range(7)] for r in range(nrep)], dtype=np.uint8)
    np.testing.assert_array_equal(x, xo)
    np.testing.assert_array_equal(e, st.energy.reshape(-1)[lo].astype(np.float64))
    logical = np.zeros((cells, 7), dtype=np.uint64)
    for r in range(nrep):
        logical[r // 64] |= xo[r].astype(np.uint64) << np.uint64(r % 64)
    assert int(digest) == _fnv_digest(logical, 7, cells)
    assert audit == 0.0
