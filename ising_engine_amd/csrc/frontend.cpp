// frontend.cpp -- instance formats lowered to the engine's QUBO form.
//
//   E(x) = offset + sum_i linear[i] x_i + sum_{i<j} q_ij x_i x_j
//
// Each parser follows the reference's rules for one format, kept in f64 as the
// reference keeps them; ige_problem_pairs then hands out the integer form the
// engine runs on, after the integrality check of ProblemIR::is_integral
// (src/engine_v2/ir.rs:143-148).
//   IGE_FORMAT_RUDY    parse_rudy + maxcut_to_qubo   src/engine_v2/frontend.rs:40-84
//   IGE_FORMAT_GSET    parse_rudy_maxcut             src/benchmark/instances.rs:103-140
//   IGE_FORMAT_ORLIB   parse_orlib_bqp               src/benchmark/instances.rs:142-186
//   IGE_FORMAT_BIQMAC  parse_biqmac_sparse           src/benchmark/instances.rs:188-223
//   IGE_FORMAT_QPLIB   parse_qplib (binary, unconstrained only)
//                                                    src/benchmark/instances.rs:225-344
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "abi_util.hpp"

using ige::abi::Failure;
using ige::abi::guarded;
using ige::abi::require;

struct ige_problem {
    uint32_t n = 0;
    double offset = 0.0;
    std::vector<double> linear;
    std::vector<uint32_t> pi, pj;  // merged pairs, i < j, sorted, zero sums dropped
    std::vector<double> q;
    int32_t maximize = 0;
    uint32_t problems_in_file = 1;
    std::string family;
};

namespace {

struct Edge {
    uint32_t i, j;
    double w;
};

[[noreturn]] void bad(const std::string& msg) { throw Failure(IGE_ERR_INVALID, msg); }

std::vector<std::string> split_ws(const std::string& line) {
    std::vector<std::string> t;
    size_t k = 0;
    while (k < line.size()) {
        while (k < line.size() && std::isspace(static_cast<unsigned char>(line[k]))) ++k;
        const size_t b = k;
        while (k < line.size() && !std::isspace(static_cast<unsigned char>(line[k]))) ++k;
        if (k > b) t.push_back(line.substr(b, k - b));
    }
    return t;
}

std::vector<std::string> lines_of(const char* text, uint64_t len) {
    std::vector<std::string> out;
    std::string cur;
    for (uint64_t k = 0; k < len; ++k) {
        if (text[k] == '\n') {
            out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(text[k]);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    for (auto& l : out)
        if (!l.empty() && l.back() == '\r') l.pop_back();
    return out;
}

// Unsigned decimal (an optional leading '+'), as usize::from_str accepts.
bool parse_uint(const std::string& s, uint64_t* v) {
    size_t k = s.size() && s[0] == '+' ? 1 : 0;
    if (k >= s.size()) return false;
    uint64_t r = 0;
    for (; k < s.size(); ++k) {
        if (s[k] < '0' || s[k] > '9') return false;
        if (r > (std::numeric_limits<uint64_t>::max() - 9) / 10) return false;
        r = r * 10 + static_cast<uint64_t>(s[k] - '0');
    }
    *v = r;
    return true;
}

// Decimal floating point as f64::from_str accepts (no hexadecimal forms).
bool parse_f64(const std::string& s, double* v) {
    if (s.empty() || s.find_first_of("xX") != std::string::npos) return false;
    char* end = nullptr;
    const double r = std::strtod(s.c_str(), &end);
    if (end != s.c_str() + s.size()) return false;
    *v = r;
    return true;
}

uint64_t need_uint(const std::string& s, const char* what) {
    uint64_t v = 0;
    if (!parse_uint(s, &v)) bad(what);
    return v;
}

double need_f64(const std::string& s, const char* what) {
    double v = 0;
    if (!parse_f64(s, &v)) bad(what);
    return v;
}

uint32_t checked_n(uint64_t n) {
    if (n > 0x7FFFFFFFull) bad("n too large");
    return static_cast<uint32_t>(n);
}

// Sorts the upper-triangle edges, sums duplicates and drops zero sums
// (the CSR rules of ProblemIR::from_pairs, src/engine_v2/ir.rs:30-69).
void finish(ige_problem& p, std::vector<Edge> edges) {
    for (auto& e : edges)
        if (e.i > e.j) std::swap(e.i, e.j);
    std::stable_sort(edges.begin(), edges.end(), [](const Edge& a, const Edge& b) {
        return a.i != b.i ? a.i < b.i : a.j < b.j;
    });
    for (size_t k = 0; k < edges.size();) {
        size_t e = k;
        double w = 0.0;
        for (; e < edges.size() && edges[e].i == edges[k].i && edges[e].j == edges[k].j; ++e)
            w += edges[e].w;
        if (w != 0.0) {
            p.pi.push_back(edges[k].i);
            p.pj.push_back(edges[k].j);
            p.q.push_back(w);
        }
        k = e;
    }
}

// frontend.rs:40-84: header n; lines `u v [w]`, 1-indexed, w = 1 when absent;
// short lines, self-loops and out-of-range vertices skipped; then MaxCut ->
// QUBO: linear[u] -= w, linear[v] -= w, pair 2w.
void parse_rudy_engine(const std::vector<std::string>& lines, ige_problem& p) {
    size_t k = 0;
    std::vector<std::string> t;
    while (k < lines.size() && (t = split_ws(lines[k])).empty()) ++k;
    if (k == lines.size()) bad("empty file");
    const uint32_t n = checked_n(need_uint(t[0], "bad n"));
    p.n = n;
    p.linear.assign(n, 0.0);
    std::vector<Edge> edges;
    for (++k; k < lines.size(); ++k) {
        t = split_ws(lines[k]);
        if (t.size() < 2) continue;
        const uint64_t u = need_uint(t[0], "bad u");
        const uint64_t v = need_uint(t[1], "bad v");
        const double w = t.size() > 2 ? need_f64(t[2], "bad w") : 1.0;
        if (u == v || u < 1 || v < 1 || u > n || v > n) continue;
        const uint32_t a = static_cast<uint32_t>(std::min(u, v) - 1);
        const uint32_t b = static_cast<uint32_t>(std::max(u, v) - 1);
        p.linear[a] -= w;
        p.linear[b] -= w;
        edges.push_back({a, b, 2.0 * w});
    }
    p.maximize = 1;
    p.family = "maxcut";
    finish(p, std::move(edges));
}

// instances.rs:89-97: a line's tokens after a '#' or '%' comment is cut off.
std::vector<std::string> tokens_nocomment(const std::string& line) {
    const size_t c = line.find_first_of("#%");
    return split_ws(c == std::string::npos ? line : line.substr(0, c));
}

// instances.rs:103-140: header `n m`; lines `i j w` (fewer than three tokens
// skipped), vertices out of range an error, self-loops skipped.
void parse_gset(const std::vector<std::string>& lines, ige_problem& p) {
    size_t k = 0;
    std::vector<std::string> t;
    while (k < lines.size() && (t = tokens_nocomment(lines[k])).empty()) ++k;
    if (k == lines.size()) bad("empty MaxCut file");
    if (t.size() < 2) bad("bad MaxCut header");
    const uint32_t n = checked_n(need_uint(t[0], "bad n"));
    (void)need_uint(t[1], "bad m");
    p.n = n;
    p.linear.assign(n, 0.0);
    std::vector<Edge> edges;
    for (++k; k < lines.size(); ++k) {
        t = tokens_nocomment(lines[k]);
        if (t.size() < 3) continue;
        const uint64_t i = need_uint(t[0], "bad edge i");
        const uint64_t j = need_uint(t[1], "bad edge j");
        const double w = need_f64(t[2], "bad edge w");
        if (i == 0 || j == 0 || i > n || j > n) bad("vertex out of range");
        if (i == j) continue;
        edges.push_back({static_cast<uint32_t>(i - 1), static_cast<uint32_t>(j - 1), 2.0 * w});
        p.linear[i - 1] -= w;
        p.linear[j - 1] -= w;
    }
    p.maximize = 1;
    p.family = "maxcut";
    finish(p, std::move(edges));
}

struct TokenStream {
    std::vector<std::string> t;
    size_t k = 0;
    const std::string& next(const char* what) {
        if (k >= t.size()) bad(what);
        return t[k++];
    }
};

TokenStream all_tokens(const char* text, uint64_t len) {
    TokenStream s;
    s.t = split_ws(std::string(text, len));
    return s;
}

// instances.rs:142-186: `p`, then per problem `n nnz` and nnz triples
// `i j q` (one triangle, 1-indexed).  E = -x'Qx: diagonal -> linear -= q,
// off-diagonal -> pair -2q.  `index` picks the problem.
void parse_orlib(const char* text, uint64_t len, uint32_t index, ige_problem& p) {
    TokenStream s = all_tokens(text, len);
    const uint64_t count = need_uint(s.next("unexpected EOF in BQP"), "bad problem count");
    if (index >= count) bad("problem index out of range");
    for (uint64_t k = 0; k < count; ++k) {  // every problem must parse, as the reference's Vec
        const uint32_t n = checked_n(need_uint(s.next("unexpected EOF in BQP"), "bad n"));
        const uint64_t nnz = need_uint(s.next("unexpected EOF in BQP"), "bad nnz");
        ige_problem cur;
        cur.n = n;
        cur.linear.assign(n, 0.0);
        std::vector<Edge> edges;
        for (uint64_t e = 0; e < nnz; ++e) {
            const uint64_t i = need_uint(s.next("unexpected EOF in BQP"), "bad i");
            const uint64_t j = need_uint(s.next("unexpected EOF in BQP"), "bad j");
            const double v = need_f64(s.next("unexpected EOF in BQP"), "bad q");
            if (i == 0 || j == 0 || i > n || j > n) bad("index out of range");
            if (i == j)
                cur.linear[i - 1] -= v;
            else
                edges.push_back({static_cast<uint32_t>(i - 1), static_cast<uint32_t>(j - 1),
                                 -2.0 * v});
        }
        if (k == index) {
            p.n = cur.n;
            p.linear = std::move(cur.linear);
            finish(p, std::move(edges));
        }
    }
    p.maximize = 1;
    p.family = "orlib_bqp";
    p.problems_in_file = static_cast<uint32_t>(std::min<uint64_t>(count, 0xFFFFFFFFull));
}

// instances.rs:188-223: `n nnz`, nnz triples; MINIMIZE x'Qx with one triangle
// listed: diagonal -> linear += q, off-diagonal -> pair 2q.
void parse_biqmac(const char* text, uint64_t len, ige_problem& p) {
    TokenStream s = all_tokens(text, len);
    const char* eof = "unexpected EOF in Biq Mac sparse";
    const uint32_t n = checked_n(need_uint(s.next(eof), "bad n"));
    const uint64_t nnz = need_uint(s.next(eof), "bad nnz");
    p.n = n;
    p.linear.assign(n, 0.0);
    std::vector<Edge> edges;
    for (uint64_t e = 0; e < nnz; ++e) {
        const uint64_t i = need_uint(s.next(eof), "bad i");
        const uint64_t j = need_uint(s.next(eof), "bad j");
        const double v = need_f64(s.next(eof), "bad q");
        if (i == 0 || j == 0 || i > n || j > n) bad("index out of range");
        if (i == j)
            p.linear[i - 1] += v;
        else
            edges.push_back({static_cast<uint32_t>(i - 1), static_cast<uint32_t>(j - 1), 2.0 * v});
    }
    p.maximize = 0;
    p.family = "biqmac_sparse";
    finish(p, std::move(edges));
}

// instances.rs:225-344: line-structured; name, type (the second character
// must be B, the third N), sense, n, nq, nq lines `i j q` (every listed entry
// carries 1/2: diagonal -> linear, off-diagonal -> pair), b_default (added to
// every linear term), nb, nb lines `i v` (v replaces the default), then an
// optional c.  A maximising objective is negated so E is always minimised.
void parse_qplib(const std::vector<std::string>& raw, ige_problem& p) {
    std::vector<std::vector<std::string>> lines;
    for (const auto& l : raw) {
        auto t = split_ws(l);
        if (!t.empty()) lines.push_back(std::move(t));
    }
    size_t k = 0;
    auto next_line = [&](const char* what) -> const std::vector<std::string>& {
        if (k >= lines.size()) bad(std::string("qplib: missing ") + what);
        return lines[k++];
    };
    auto first_f64 = [&](const std::vector<std::string>& t, const char* what) {
        return need_f64(t[0], (std::string("qplib: bad ") + what).c_str());
    };
    auto as_count = [](double v) -> uint64_t {
        if (!(v > 0.0)) return 0;  // f64 -> usize saturates at 0
        return v >= 1.8e19 ? std::numeric_limits<uint64_t>::max() : static_cast<uint64_t>(v);
    };
    (void)next_line("name");
    const std::string ptype = next_line("type")[0];
    if (ptype.size() < 2 || std::toupper(static_cast<unsigned char>(ptype[1])) != 'B')
        bad("QPLIB type " + ptype + " not binary-variable (unsupported)");
    if (ptype.size() < 3 || std::toupper(static_cast<unsigned char>(ptype[2])) != 'N')
        bad("QPLIB type " + ptype + " is constrained; only unconstrained instances supported");
    std::string sense = next_line("sense")[0];
    for (auto& c : sense) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    const bool maximize = sense == "maximize";
    const double sign = maximize ? -1.0 : 1.0;
    const uint32_t n = checked_n(as_count(first_f64(next_line("n"), "n")));
    const uint64_t nq = as_count(first_f64(next_line("nq"), "nq"));
    p.n = n;
    p.linear.assign(n, 0.0);
    std::vector<Edge> edges;
    for (uint64_t e = 0; e < nq; ++e) {
        const auto& t = next_line("quad term");
        if (t.size() < 3) bad("short quad line");
        const uint64_t i = as_count(need_f64(t[0], "qplib: bad qi"));
        const uint64_t j = as_count(need_f64(t[1], "qplib: bad qj"));
        const double q = need_f64(t[2], "qplib: bad qv");
        if (i == 0 || j == 0 || i > n || j > n) bad("quad index out of range");
        if (i == j)
            p.linear[i - 1] += sign * 0.5 * q;
        else
            edges.push_back({static_cast<uint32_t>(i - 1), static_cast<uint32_t>(j - 1),
                             sign * 0.5 * q});
    }
    const double b_default = first_f64(next_line("b_default"), "b_default");
    for (auto& v : p.linear) v += sign * b_default;
    const uint64_t nb = as_count(first_f64(next_line("nb"), "nb"));
    for (uint64_t e = 0; e < nb; ++e) {
        const auto& t = next_line("linear term");
        if (t.size() < 2) bad("short linear line");
        const uint64_t i = as_count(need_f64(t[0], "qplib: bad bi"));
        const double v = need_f64(t[1], "qplib: bad bv");
        if (i == 0 || i > n) bad("linear index out of range");
        p.linear[i - 1] += sign * (v - b_default);
    }
    double c = 0.0;
    if (k < lines.size()) {
        double v = 0.0;
        if (parse_f64(lines[k][0], &v)) c = v;
        ++k;
    }
    p.offset = sign * c;
    p.maximize = maximize ? 1 : 0;
    p.family = "qplib";
    finish(p, std::move(edges));
}

bool intish(double v) { return std::isfinite(v) && v == std::trunc(v); }

bool integral(const ige_problem& p, bool* in_range) {
    bool ok = intish(p.offset);
    bool range = std::fabs(p.offset) < 9.2e18;
    for (double v : p.linear) {
        ok = ok && intish(v);
        range = range && std::fabs(v) <= 2147483647.0;
    }
    for (double v : p.q) {
        ok = ok && intish(v);
        range = range && std::fabs(v) <= 2147483647.0;
    }
    if (in_range) *in_range = range;
    return ok;
}

}  // namespace

extern "C" {

int32_t ige_problem_parse(const char* text, uint64_t len, uint32_t format, uint32_t index,
                          ige_problem** out) {
    return guarded([&] {
        require(out != nullptr && (text != nullptr || len == 0), "NULL argument");
        require(format <= IGE_FORMAT_QPLIB, "unknown format");
        require(index == 0 || format == IGE_FORMAT_ORLIB, "index applies to OR-Library files");
        auto p = std::make_unique<ige_problem>();
        switch (format) {
            case IGE_FORMAT_RUDY: parse_rudy_engine(lines_of(text, len), *p); break;
            case IGE_FORMAT_GSET: parse_gset(lines_of(text, len), *p); break;
            case IGE_FORMAT_ORLIB: parse_orlib(text, len, index, *p); break;
            case IGE_FORMAT_BIQMAC: parse_biqmac(text, len, *p); break;
            default: parse_qplib(lines_of(text, len), *p); break;
        }
        require(p->n >= 1, "instance needs n >= 1");
        *out = p.release();
    });
}

int32_t ige_problem_destroy(ige_problem* p) {
    delete p;
    return IGE_OK;
}

int32_t ige_problem_get_info(const ige_problem* p, ige_problem_info* info) {
    return guarded([&] {
        require(p && info, "NULL argument");
        std::memset(info, 0, sizeof(*info));
        info->n = p->n;
        info->problems_in_file = p->problems_in_file;
        info->npairs = p->pi.size();
        info->maximize = p->maximize;
        bool range = false;
        info->integral = integral(*p, &range) ? 1 : 0;
        info->offset = p->offset;
    });
}

int32_t ige_problem_pairs(const ige_problem* p, int64_t* offset, int32_t* linear, uint32_t* pi,
                          uint32_t* pj, int32_t* q) {
    return guarded([&] {
        require(p && offset && linear && (p->pi.empty() || (pi && pj && q)), "NULL argument");
        bool range = false;
        if (!integral(*p, &range))
            throw Failure(IGE_ERR_RANGE,
                          "instance is non-integral; the integer engine is not eligible");
        if (!range) throw Failure(IGE_ERR_RANGE, "a coefficient exceeds int32");
        *offset = static_cast<int64_t>(p->offset);
        for (uint32_t i = 0; i < p->n; ++i) linear[i] = static_cast<int32_t>(p->linear[i]);
        for (size_t e = 0; e < p->pi.size(); ++e) {
            pi[e] = p->pi[e];
            pj[e] = p->pj[e];
            q[e] = static_cast<int32_t>(p->q[e]);
        }
    });
}

int32_t ige_problem_energy(const ige_problem* p, const uint8_t* x, double* energy) {
    return guarded([&] {
        require(p && x && energy, "NULL argument");
        double e = p->offset;
        for (uint32_t i = 0; i < p->n; ++i)
            if (x[i]) e += p->linear[i];
        for (size_t k = 0; k < p->pi.size(); ++k)
            if (x[p->pi[k]] && x[p->pj[k]]) e += p->q[k];
        *energy = e;
    });
}

}  // extern "C"
