// problem_ir.cpp -- host construction of the canonical model; see problem_ir.hpp.
#include "problem_ir.hpp"

#include <algorithm>
#include <numeric>
#include <cstdlib>
#include <limits>

namespace ige {

ProblemIR ProblemIR::from_pairs(uint32_t n, int64_t offset, const int32_t* linear,
                                uint64_t npairs, const uint32_t* pi, const uint32_t* pj,
                                const int32_t* q) {
    ProblemIR ir;
    ir.n = n;
    ir.offset = offset;
    ir.linear.assign(linear, linear + n);
    // Counting sort into rows (both triangles), then sort + merge each row:
    // the same CSR the reference's per-row Vec + sort_unstable_by_key builds.
    std::vector<uint64_t> start(static_cast<size_t>(n) + 1, 0);
    for (uint64_t k = 0; k < npairs; ++k) {
        if (!(pi[k] < pj[k] && pj[k] < n))
            throw std::invalid_argument("pairs must satisfy i < j < n (pair " +
                                        std::to_string(k) + ")");
        start[pi[k] + 1]++;
        start[pj[k] + 1]++;
    }
    for (uint32_t i = 0; i < n; ++i) start[i + 1] += start[i];
    std::vector<std::pair<uint32_t, int64_t>> adj(start[n]);
    std::vector<uint64_t> fill(start.begin(), start.end() - 1);
    for (uint64_t k = 0; k < npairs; ++k) {
        adj[fill[pi[k]]++] = {pj[k], q[k]};
        adj[fill[pj[k]]++] = {pi[k], q[k]};
    }
    ir.row_ptr.assign(static_cast<size_t>(n) + 1, 0);
    ir.col_idx.reserve(adj.size());
    ir.weights.reserve(adj.size());
    for (uint32_t i = 0; i < n; ++i) {
        auto b = adj.begin() + static_cast<std::ptrdiff_t>(start[i]);
        auto e = adj.begin() + static_cast<std::ptrdiff_t>(start[i + 1]);
        std::sort(b, e, [](const auto& a, const auto& c) { return a.first < c.first; });
        for (auto it = b; it != e;) {
            const uint32_t j = it->first;
            int64_t acc = 0;
            for (; it != e && it->first == j; ++it) acc += it->second;
            if (acc == 0) continue;
            if (acc > std::numeric_limits<int32_t>::max() ||
                acc < std::numeric_limits<int32_t>::min())
                throw RangeError("accumulated pair coefficient leaves int32");
            ir.col_idx.push_back(j);
            ir.weights.push_back(static_cast<int32_t>(acc));
        }
        if (ir.col_idx.size() > std::numeric_limits<uint32_t>::max())
            throw RangeError("more than 2^32-1 CSR entries");
        ir.row_ptr[i + 1] = static_cast<uint32_t>(ir.col_idx.size());
    }
    return ir;
}

int64_t ProblemIR::energy(const uint8_t* x) const {
    int64_t lin = 0, pair = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!x[i]) continue;
        lin += linear[i];
        for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
            if (x[col_idx[e]]) pair += weights[e];
    }
    return offset + lin + pair / 2;
}

std::vector<uint32_t> ProblemIR::greedy_coloring() const {
    constexpr uint32_t kNone = std::numeric_limits<uint32_t>::max();
    std::vector<uint32_t> colors(n, kNone);
    // stamp[c] == i marks colour c as used by a coloured neighbour of site i.
    std::vector<uint32_t> stamp(static_cast<size_t>(n) + 2, kNone);
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
            const uint32_t c = colors[col_idx[e]];
            if (c != kNone) stamp[c] = i;
        }
        uint32_t c = 0;
        while (stamp[c] == i) ++c;
        colors[i] = c;
    }
    return colors;
}

int64_t ProblemIR::max_local_field() const {
    int64_t best = 0;
    for (uint32_t i = 0; i < n; ++i) {
        int64_t s = std::llabs(static_cast<long long>(linear[i]));
        for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
            s += std::llabs(static_cast<long long>(weights[e]));
        best = std::max(best, s);
    }
    return best;
}

uint32_t ProblemIR::max_degree() const {
    uint32_t d = 0;
    for (uint32_t i = 0; i < n; ++i) d = std::max(d, row_ptr[i + 1] - row_ptr[i]);
    return d;
}

ColorOrder color_order(const std::vector<uint32_t>& colors) {
    ColorOrder co;
    const uint32_t n = static_cast<uint32_t>(colors.size());
    for (uint32_t c : colors) co.num_colors = std::max(co.num_colors, c + 1);
    std::vector<uint32_t> count(co.num_colors + 1, 0);
    for (uint32_t c : colors) count[c + 1]++;
    for (uint32_t c = 0; c < co.num_colors; ++c) count[c + 1] += count[c];
    co.class_start = count;
    co.order.resize(n);
    co.pos_of.resize(n);
    std::vector<uint32_t> fill(count.begin(), count.end() - 1);
    for (uint32_t i = 0; i < n; ++i) {  // ascending sites within each class
        const uint32_t p = fill[colors[i]]++;
        co.order[p] = i;
        co.pos_of[i] = p;
    }
    return co;
}

MultispinShape multispin_shape(const ProblemIR& ir) {
    MultispinShape s;
    if (ir.n == 0 || ir.nnz() == 0) return s;
    const int64_t w0 = std::llabs(static_cast<long long>(ir.weights[0]));
    for (int32_t w : ir.weights)
        if (std::llabs(static_cast<long long>(w)) != w0) return s;
    for (uint32_t i = 0; i < ir.n; ++i) {
        int64_t sum = 0;
        for (uint32_t e = ir.row_ptr[i]; e < ir.row_ptr[i + 1]; ++e) sum += ir.weights[e];
        // lin_i must equal -sum/2 exactly (zero Ising field).
        if (2 * static_cast<int64_t>(ir.linear[i]) != -sum) return s;
    }
    if (ir.max_degree() > 31) return s;
    // the kernel addresses words and CSR entries with 32-bit byte offsets
    if (ir.n >= (1u << 29) || ir.nnz() >= (1ull << 30)) return s;
    s.eligible = true;
    s.w0 = static_cast<int32_t>(w0);
    return s;
}

DenseShape dense_shape(const ProblemIR& ir) {
    DenseShape d;
    if (ir.n == 0 || ir.n > kDenseMaxN) return d;
    int64_t g = 0;
    for (int32_t w : ir.weights) g = std::gcd(g, static_cast<int64_t>(std::llabs(w)));
    if (g == 0) g = 1;  // no couplings: any scale works
    for (int32_t w : ir.weights)
        if (std::llabs(static_cast<long long>(w)) / g > 127) return d;
    d.eligible = true;
    d.scale = static_cast<int32_t>(g);
    const uint64_t full = static_cast<uint64_t>(ir.n) * (ir.n - 1);
    d.dense = ir.n >= 64 && 4 * ir.nnz() >= full;
    return d;
}

std::vector<int8_t> dense_matrix(const ProblemIR& ir, const ColorOrder& co, uint32_t n_pad,
                                 int32_t scale) {
    std::vector<int8_t> a(static_cast<size_t>(n_pad) * n_pad, 0);
    for (uint32_t i = 0; i < ir.n; ++i) {
        const size_t row = static_cast<size_t>(co.pos_of[i]) * n_pad;
        for (uint32_t e = ir.row_ptr[i]; e < ir.row_ptr[i + 1]; ++e)
            a[row + co.pos_of[ir.col_idx[e]]] = static_cast<int8_t>(ir.weights[e] / scale);
    }
    return a;
}

}  // namespace ige
