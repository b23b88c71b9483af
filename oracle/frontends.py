"""Python restatement of the reference's instance parsers, used ONLY as a checker.

TEST INFRASTRUCTURE. Only tests/ may load this module; the product parses with
ising_engine_amd/csrc/frontend.cpp.

Every parser returns a dict: n, offset, linear (list of float), pairs (dict
(i, j) -> summed coefficient, i < j, zero sums dropped), maximize, and the
energy convention E = offset + sum lin_i x_i + sum_{i<j} q_ij x_i x_j.
  parse_rudy_engine  src/engine_v2/frontend.rs:40-84
  parse_gset         src/benchmark/instances.rs:103-140 (tokens: :89-97)
  parse_orlib        src/benchmark/instances.rs:142-186
  parse_biqmac       src/benchmark/instances.rs:188-223
  parse_qplib        src/benchmark/instances.rs:225-344
"""
from __future__ import annotations

import re


class ParseError(ValueError):
    pass


def _uint(tok: str, what: str) -> int:
    if not re.fullmatch(r"\+?[0-9]+", tok):
        raise ParseError(what)
    return int(tok)


def _f64(tok: str, what: str) -> float:
    if "x" in tok.lower():
        raise ParseError(what)
    try:
        return float(tok)
    except ValueError:
        raise ParseError(what) from None


def _model(n, linear, edges, offset=0.0, maximize=False):
    pairs: dict = {}
    for i, j, w in edges:
        a, b = min(i, j), max(i, j)
        pairs[(a, b)] = pairs.get((a, b), 0.0) + w
    pairs = {k: v for k, v in sorted(pairs.items()) if v != 0.0}
    return dict(n=n, offset=offset, linear=linear, pairs=pairs, maximize=maximize)


def parse_rudy_engine(text: str):
    lines = [l.split() for l in text.splitlines()]
    lines = [t for t in lines if t]
    if not lines:
        raise ParseError("empty file")
    n = _uint(lines[0][0], "bad n")
    linear = [0.0] * n
    edges = []
    for t in lines[1:]:
        if len(t) < 2:
            continue
        u, v = _uint(t[0], "bad u"), _uint(t[1], "bad v")
        w = _f64(t[2], "bad w") if len(t) > 2 else 1.0
        if u == v or u < 1 or v < 1 or u > n or v > n:
            continue
        a, b = min(u, v) - 1, max(u, v) - 1
        linear[a] -= w
        linear[b] -= w
        edges.append((a, b, 2.0 * w))
    return _model(n, linear, edges, maximize=True)


def _tokens(line: str):
    return re.split(r"[#%]", line, maxsplit=1)[0].split()


def parse_gset(text: str):
    lines = [l for l in text.splitlines() if _tokens(l)]
    if not lines:
        raise ParseError("empty MaxCut file")
    ht = _tokens(lines[0])
    if len(ht) < 2:
        raise ParseError("bad MaxCut header")
    n = _uint(ht[0], "bad n")
    _uint(ht[1], "bad m")
    linear = [0.0] * n
    edges = []
    for l in lines[1:]:
        t = _tokens(l)
        if len(t) < 3:
            continue
        i, j = _uint(t[0], "bad edge i"), _uint(t[1], "bad edge j")
        w = _f64(t[2], "bad edge w")
        if i == 0 or j == 0 or i > n or j > n:
            raise ParseError("vertex out of range")
        if i == j:
            continue
        edges.append((i - 1, j - 1, 2.0 * w))
        linear[i - 1] -= w
        linear[j - 1] -= w
    return _model(n, linear, edges, maximize=True)


def parse_orlib(text: str):
    toks = iter(text.split())

    def nxt():
        try:
            return next(toks)
        except StopIteration:
            raise ParseError("unexpected EOF in BQP") from None

    out = []
    for _ in range(_uint(nxt(), "bad problem count")):
        n = _uint(nxt(), "bad n")
        nnz = _uint(nxt(), "bad nnz")
        linear = [0.0] * n
        edges = []
        for _ in range(nnz):
            i, j = _uint(nxt(), "bad i"), _uint(nxt(), "bad j")
            v = _f64(nxt(), "bad q")
            if i == 0 or j == 0 or i > n or j > n:
                raise ParseError("index out of range")
            if i == j:
                linear[i - 1] -= v
            else:
                edges.append((i - 1, j - 1, -2.0 * v))
        out.append(_model(n, linear, edges, maximize=True))
    return out


def parse_biqmac(text: str):
    toks = iter(text.split())

    def nxt():
        try:
            return next(toks)
        except StopIteration:
            raise ParseError("unexpected EOF in Biq Mac sparse") from None

    n = _uint(nxt(), "bad n")
    nnz = _uint(nxt(), "bad nnz")
    linear = [0.0] * n
    edges = []
    for _ in range(nnz):
        i, j = _uint(nxt(), "bad i"), _uint(nxt(), "bad j")
        v = _f64(nxt(), "bad q")
        if i == 0 or j == 0 or i > n or j > n:
            raise ParseError("index out of range")
        if i == j:
            linear[i - 1] += v
        else:
            edges.append((i - 1, j - 1, 2.0 * v))
    return _model(n, linear, edges, maximize=False)


def _as_count(v: float) -> int:
    if not v > 0.0:
        return 0
    return int(min(v, 2.0 ** 64 - 1))


def parse_qplib(text: str):
    lines = [l.split() for l in text.splitlines()]
    it = iter([t for t in lines if t])

    def nl(what):
        try:
            return next(it)
        except StopIteration:
            raise ParseError(f"qplib: missing {what}") from None

    nl("name")
    ptype = nl("type")[0]
    if len(ptype) < 2 or ptype[1].upper() != "B":
        raise ParseError(f"QPLIB type {ptype} not binary-variable (unsupported)")
    if len(ptype) < 3 or ptype[2].upper() != "N":
        raise ParseError(f"QPLIB type {ptype} is constrained; only unconstrained instances supported")
    maximize = nl("sense")[0].lower() == "maximize"
    sign = -1.0 if maximize else 1.0
    n = _as_count(_f64(nl("n")[0], "qplib: bad n"))
    nq = _as_count(_f64(nl("nq")[0], "qplib: bad nq"))
    linear = [0.0] * n
    edges = []
    for _ in range(nq):
        t = nl("quad term")
        if len(t) < 3:
            raise ParseError("short quad line")
        i, j = _as_count(_f64(t[0], "qplib: bad qi")), _as_count(_f64(t[1], "qplib: bad qj"))
        q = _f64(t[2], "qplib: bad qv")
        if i == 0 or j == 0 or i > n or j > n:
            raise ParseError("quad index out of range")
        if i == j:
            linear[i - 1] += sign * 0.5 * q
        else:
            edges.append((i - 1, j - 1, sign * 0.5 * q))
    b_default = _f64(nl("b_default")[0], "qplib: bad b_default")
    linear = [v + sign * b_default for v in linear]
    for _ in range(_as_count(_f64(nl("nb")[0], "qplib: bad nb"))):
        t = nl("linear term")
        if len(t) < 2:
            raise ParseError("short linear line")
        i = _as_count(_f64(t[0], "qplib: bad bi"))
        v = _f64(t[1], "qplib: bad bv")
        if i == 0 or i > n:
            raise ParseError("linear index out of range")
        linear[i - 1] += sign * (v - b_default)
    c = 0.0
    try:
        t = next(it)
        try:
            c = _f64(t[0], "c")
        except ParseError:
            c = 0.0
    except StopIteration:
        pass
    return _model(n, linear, edges, offset=sign * c, maximize=maximize)


def energy(m, x) -> float:
    e = m["offset"]
    for i, v in enumerate(m["linear"]):
        if x[i]:
            e += v
    for (i, j), w in m["pairs"].items():
        if x[i] and x[j]:
            e += w
    return e
