"""Instance frontends: the C ABI parsers (frontend.cpp) against the Python restatement
(oracle/frontends.py), the reference's own parser tests by brute force
(tests/test_benchmark.rs:24-123, src/engine_v2/frontend.rs:93-101), and closed forms on the
four G-Set fixtures the reference ships (gset/G1, G22, G39, G55).

Host only: parsing needs no device.
"""
import itertools
import os

import numpy as np
import pytest

from ising_engine_amd import _lib
from oracle import frontends as fr

GSET = os.path.join(os.path.dirname(__file__), "golden", "gset")


def brute_min(energy, n):
    return min(energy(np.array(bits, dtype=np.uint8))
               for bits in itertools.product([0, 1], repeat=n))


def same(p: _lib.Problem, m: dict):
    """The C ABI problem equals the restatement: n, offset, linear terms, merged pairs."""
    assert p.n == m["n"] and p.info.npairs == len(m["pairs"])
    assert p.info.maximize == int(m["maximize"]) and p.info.offset == m["offset"]
    if p.info.integral:
        off, lin, pi, pj, q = p.pairs()
        assert off == m["offset"]
        np.testing.assert_array_equal(lin, np.array(m["linear"], dtype=np.float64))
        keys = list(m["pairs"].keys())
        np.testing.assert_array_equal(pi, [k[0] for k in keys])
        np.testing.assert_array_equal(pj, [k[1] for k in keys])
        np.testing.assert_array_equal(q, [m["pairs"][k] for k in keys])
    rng = np.random.default_rng(p.n)
    for _ in range(5):
        x = rng.integers(0, 2, size=p.n, dtype=np.uint8)
        assert p.energy(x) == fr.energy(m, x)


# ---- the reference's own parser tests -------------------------------------------------
def test_rudy_triangle_engine_frontend():
    # frontend.rs:93-101: x = (0,1,0) cuts two edges -> energy -2; x = 0 -> 0
    p = _lib.Problem("3 3\n1 2\n2 3\n1 3\n", _lib.FORMAT_RUDY)
    assert p.energy(np.array([0, 1, 0], np.uint8)) == -2.0
    assert p.energy(np.zeros(3, np.uint8)) == 0.0
    same(p, fr.parse_rudy_engine("3 3\n1 2\n2 3\n1 3\n"))


@pytest.mark.parametrize("text,fmt,minimum,maximize", [
    ("3 3\n1 2 1\n2 3 1\n1 3 1\n", _lib.FORMAT_GSET, -2.0, True),       # triangle: cut 2
    ("4 4\n1 2 1\n2 3 1\n3 4 1\n4 1 1\n", _lib.FORMAT_GSET, -4.0, True),  # 4-cycle: cut 4
    ("1\n2 3\n1 1 1\n2 2 1\n1 2 3\n", _lib.FORMAT_ORLIB, -8.0, True),     # max x'Qx = 8
    ("2 3\n1 1 1\n2 2 1\n1 2 -3\n", _lib.FORMAT_BIQMAC, -4.0, False),     # min x'Qx = -4
    ("toy name\nQBN type\nminimize sense\n2 n\n1 nq\n1 2 -6 q12\n0 b_default\n2 nb\n"
     "1 1 b1\n2 1 b2\n0 c\n", _lib.FORMAT_QPLIB, -1.0, False),            # x1 + x2 - 3 x1x2
])
def test_reference_parser_cases_by_brute_force(text, fmt, minimum, maximize):
    p = _lib.Problem(text, fmt)
    assert brute_min(p.energy, p.n) == minimum
    assert bool(p.info.maximize) == maximize
    m = {_lib.FORMAT_GSET: fr.parse_gset, _lib.FORMAT_BIQMAC: fr.parse_biqmac,
         _lib.FORMAT_QPLIB: fr.parse_qplib,
         _lib.FORMAT_ORLIB: lambda t: fr.parse_orlib(t)[0]}[fmt](text)
    same(p, m)
    assert brute_min(lambda x: fr.energy(m, x), p.n) == minimum


@pytest.mark.parametrize("text,msg", [
    ("c name\nQBL type\nminimize sense\n2 n\n", "constrained"),       # test_benchmark.rs:112
    ("c name\nQCN type\nminimize sense\n2 n\n", "not binary-variable"),  # test_benchmark.rs:119
])
def test_qplib_refusals(text, msg):
    with pytest.raises(_lib.IgeError) as ei:
        _lib.Problem(text, _lib.FORMAT_QPLIB)
    assert ei.value.code == _lib.IGE_ERR_INVALID and msg in str(ei.value)
    with pytest.raises(fr.ParseError, match=msg):
        fr.parse_qplib(text)


# ---- the two rudy readers differ where the reference's do ----------------------------
def test_rudy_readers_differ_as_the_reference():
    text = "4 3\n1 2\n2 3 5\n3 9 1\n"  # a missing weight and an out-of-range vertex
    eng = _lib.Problem(text, _lib.FORMAT_RUDY)          # w = 1 default, bad vertex skipped
    same(eng, fr.parse_rudy_engine(text))
    assert eng.info.npairs == 2
    with pytest.raises(_lib.IgeError, match="vertex out of range"):
        _lib.Problem(text, _lib.FORMAT_GSET)
    commented = "% a comment\n3 2  # header\n1 2 1\n2 3 2 % trailing\n"
    g = _lib.Problem(commented, _lib.FORMAT_GSET)
    same(g, fr.parse_gset(commented))
    with pytest.raises(_lib.IgeError, match="bad"):
        _lib.Problem(commented, _lib.FORMAT_RUDY)         # the engine reader has no comments


def test_orlib_problem_index_and_duplicates():
    text = "2\n3 3\n1 1 2\n1 2 4\n2 1 1\n2 1\n2 2 -7\n"
    both = fr.parse_orlib(text)
    for k in range(2):
        p = _lib.Problem(text, _lib.FORMAT_ORLIB, index=k)
        assert p.info.problems_in_file == 2
        same(p, both[k])
    p0 = _lib.Problem(text, _lib.FORMAT_ORLIB, index=0)
    assert p0.info.npairs == 1 and p0.pairs()[4][0] == -10  # (1,2) and (2,1) summed: -2(4+1)
    with pytest.raises(_lib.IgeError):
        _lib.Problem(text, _lib.FORMAT_ORLIB, index=2)


def test_non_integral_instance_is_refused_for_the_integer_engine():
    text = "t name\nQBN type\nminimize sense\n2 n\n1 nq\n1 2 3 q\n0 b\n0 nb\n0 c\n"  # pair 1.5
    p = _lib.Problem(text, _lib.FORMAT_QPLIB)
    assert p.info.integral == 0
    same(p, fr.parse_qplib(text))
    with pytest.raises(_lib.IgeError) as ei:
        p.pairs()
    assert ei.value.code == _lib.IGE_ERR_RANGE


# ---- the G-Set fixtures ---------------------------------------------------------------
@pytest.mark.parametrize("name,n,m", [("G1", 800, 19176), ("G22", 2000, 19990),
                                      ("G39", 2000, 11778), ("G55", 5000, 12498)])
def test_gset_fixtures(name, n, m):
    text = open(os.path.join(GSET, f"{name}.txt")).read()
    head = text.split()[:2]
    assert [int(head[0]), int(head[1])] == [n, m]
    p = _lib.Problem(text, _lib.FORMAT_GSET)
    ref = fr.parse_gset(text)
    same(p, ref)
    assert p.n == n and p.info.npairs == m and p.info.maximize and p.info.integral
    off, lin, pi, pj, q = p.pairs()
    # closed forms: pair 2w, linear -sum of incident weights, so E(0) = E(1) = 0 and
    # E(e_i) = -(weighted degree of i) = minus the cut of one vertex
    w = q / 2
    deg = np.zeros(n)
    np.add.at(deg, pi, w)
    np.add.at(deg, pj, w)
    np.testing.assert_array_equal(lin, -deg)
    assert p.energy(np.zeros(n, np.uint8)) == 0.0 and p.energy(np.ones(n, np.uint8)) == 0.0
    for i in (0, n // 2, n - 1):
        x = np.zeros(n, np.uint8)
        x[i] = 1
        assert p.energy(x) == -deg[i]
    # a random bipartition: -E equals the cut weight counted edge by edge
    x = np.random.default_rng(3).integers(0, 2, size=n, dtype=np.uint8)
    cut = float(np.sum(w * (x[pi] != x[pj])))
    assert p.energy(x) == -cut
    # the engine_v2 reader sees the same instance (weights present, no comments)
    same(_lib.Problem(text, _lib.FORMAT_RUDY), ref)


def test_solve_v2_argument_errors(tmp_path, capsys):
    from ising_engine_amd import solve_v2
    assert solve_v2.main([]) == 2
    f = tmp_path / "t.rudy"
    f.write_text("3 1\n1 2 1\n")
    assert solve_v2.main(["--file", str(f), "--format", "gset"]) == 2
    assert solve_v2.main(["--file", str(tmp_path / "missing.rudy")]) == 1
    assert "unsupported format" in capsys.readouterr().err


def test_solve_v2_ladder():
    from ising_engine_amd import solve_v2
    t = solve_v2.ladder(200)
    assert t[0] == 4.0 and abs(t[-1] - 0.08) < 1e-12 and len(t) == 200
    assert np.all(np.diff(t) < 0)
    assert list(solve_v2.ladder(1)) == [4.0]
