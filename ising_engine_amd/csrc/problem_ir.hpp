// problem_ir.hpp -- the canonical integer QUBO model on the host.
//
//   E(x) = offset + sum_i linear[i] x_i + sum_{i<j} q_ij x_i x_j,  x in {0,1}
//
// Stored as symmetric CSR with both triangles, rows sorted by column,
// duplicate pairs summed and zero sums dropped, exactly as
// ProblemIR::from_pairs builds it (src/engine_v2/ir.rs:30-69).  The
// reference keeps f64 containers and checks is_integral() (ir.rs:143-148)
// before choosing its integer backend; this engine only accepts integer
// input, so integrality holds by construction and range checks replace it.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ige {

struct RangeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct ProblemIR {
    uint32_t n = 0;
    int64_t offset = 0;
    std::vector<int32_t> linear;
    std::vector<uint32_t> row_ptr;  // n+1
    std::vector<uint32_t> col_idx;  // 2m
    std::vector<int32_t> weights;   // 2m, full q_ij per entry

    // ir.rs:30-69.  Throws std::invalid_argument unless i < j < n for every
    // pair, RangeError if an accumulated coefficient leaves int32.
    static ProblemIR from_pairs(uint32_t n, int64_t offset, const int32_t* linear,
                                uint64_t npairs, const uint32_t* pi, const uint32_t* pj,
                                const int32_t* q);

    uint64_t nnz() const { return col_idx.size(); }

    // ir.rs:85-103, exact in int64 (the both-triangle pair sum is even).
    int64_t energy(const uint8_t* x) const;

    // ir.rs:114-139: ascending sites, smallest colour unused by an already
    // coloured neighbour.
    std::vector<uint32_t> greedy_coloring() const;

    // max_i |linear_i| + sum_j |q_ij|: bounds every local field and |dE|.
    int64_t max_local_field() const;
    uint32_t max_degree() const;
};

// Sweep order of MetropolisSweep::ensure_order
// (src/engine_v2/operators/metropolis_sweep.rs:40-58): colour 0 sites in
// ascending order, then colour 1, ...  order[pos] = site, pos_of[site] = pos.
struct ColorOrder {
    uint32_t num_colors = 0;
    std::vector<uint32_t> order;
    std::vector<uint32_t> pos_of;
    std::vector<uint32_t> class_start;  // num_colors + 1
};
ColorOrder color_order(const std::vector<uint32_t>& colors);

// +-J structure the bit-sliced kernel needs: every |q_ij| equals w0 and the
// linear terms are those of a zero-field Ising model, lin_i = -1/2 sum_j q_ij
// (the sigma = 1 - 2x lowering of src/bin/sk_parisi_benchmark.rs:56-110 with
// pair -4J and linear 2 sum_j J_ij).  Then dE_i = w0 (k_i - d_i / 2), k_i the
// number of bonds with x_i ^ x_j ^ [q_ij < 0] set.
struct MultispinShape {
    bool eligible = false;
    int32_t w0 = 0;
};
MultispinShape multispin_shape(const ProblemIR& ir);

// Dense-coupling structure for the int8 matrix-core field recompute: every
// q_ij = scale * a_ij with a_ij an int8 (|a_ij| <= 127), scale = gcd of |q_ij|,
// and n small enough for the replica-major spin bits of one slot to sit in
// LDS.  `dense` marks models whose both-triangle CSR fills at least a quarter
// of the n x n matrix (fully connected SK, dense QUBO), where the automatic
// kernel choice takes this path.
constexpr uint32_t kDenseMaxN = 8192;
struct DenseShape {
    bool eligible = false;
    bool dense = false;
    int32_t scale = 0;
};
DenseShape dense_shape(const ProblemIR& ir);

// a[p * n_pad + p'] = q(order[p], order[p']) / scale in position space,
// zero on the diagonal and in the padding up to n_pad (a multiple of 64).
std::vector<int8_t> dense_matrix(const ProblemIR& ir, const ColorOrder& co, uint32_t n_pad,
                                 int32_t scale);

}  // namespace ige
