// batch_internal.hpp -- lowering of independent instances to position space
// and the persistent-kernel launcher, shared by the batch API and the
// serving loop.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/ising_gpu.h"
#include "abi_util.hpp"
#include "ige_math.h"
#include "persistent.hpp"
#include "problem_ir.hpp"

namespace ige {

static __global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

}  // namespace ige


using ige::abi::check;
using ige::abi::Failure;
using ige::abi::require;

namespace ige {

// UltimateSolver's ladder (ultimate.rs:237-246): T_t = temp_max * f^t with
// f = (temp_min / temp_max)^(1 / (k - 1)); f^t by binary exponentiation, the
// way Rust's powi lowers it.  beta[t] = 1 / T_t.
inline double powi_binary(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

inline std::vector<double> ultimate_ladder(double temp_max, double temp_min, uint32_t k) {
    const double f = k > 1 ? std::pow(temp_min / temp_max, 1.0 / static_cast<double>(k - 1)) : 1.0;
    std::vector<double> beta(k);
    for (uint32_t t = 0; t < k; ++t) beta[t] = 1.0 / (temp_max * powi_binary(f, static_cast<int>(t)));
    return beta;
}

// Host-side description of one instance once lowered to position space.
struct LoweredInstance {
    uint32_t n = 0, ncol = 0, sweeps = 0, dtab = 0;
    uint32_t track_every = 0, polish = 0;  // serving: incumbent cadence, 1-opt polish
    uint32_t cells = 1, pt = 0;            // PT solve: cells on a fixed ladder
    std::vector<double> ladder;            // its betas
    uint64_t seed = 0;
    double beta_hi = 0.0, c = 0.0;
    int64_t offset = 0;
    std::vector<uint32_t> rp;   // n+1, relative to the instance's first edge
    std::vector<uint32_t> ci;   // neighbour positions
    std::vector<int32_t> wt;
    std::vector<int32_t> lin;
    std::vector<uint32_t> cs;   // ncol+1
    std::vector<uint32_t> order;  // position -> site
    std::vector<uint32_t> pos_of; // site -> position
    uint64_t work() const { return (static_cast<uint64_t>(n) + ci.size()) * sweeps * cells; }
};

inline void check_schedule(double beta_lo, double beta_hi, const std::string& who) {
    if (!(std::isfinite(beta_lo) && std::isfinite(beta_hi) && beta_lo > 0 && beta_hi >= beta_lo &&
          std::log(beta_hi / beta_lo) <= 40.0))
        throw Failure(IGE_ERR_INVALID, who + ": need 0 < beta_lo <= beta_hi <= e^40 beta_lo");
}

// ProblemIR::from_pairs (ir.rs:30-69), greedy colouring (ir.rs:114-139) and the
// colour-major position order (metropolis_sweep.rs:40-58) of one request.
inline LoweredInstance lower_instance(const ige_instance_desc& d, const std::string& who) {
    require(d.n >= 1 && d.linear && (d.npairs == 0 || (d.pi && d.pj && d.q)),
            who + ": bad model arrays");
    require(d.sweeps >= 1, who + ": sweeps must be >= 1");
    if (d.num_temps == 0) {
        check_schedule(d.beta_lo, d.beta_hi, who);
    } else {
        require(d.num_temps <= kMaxPtCells, who + ": num_temps must be in [1, 64]");
        require(std::isfinite(d.temp_max) && std::isfinite(d.temp_min) && d.temp_max > 0.0 &&
                    d.temp_min > 0.0,
                who + ": temperatures must be finite and > 0");
        require(d.track_every >= 1 && d.sweeps % d.track_every == 0,
                who + ": track_every (sweeps per exchange round) must divide sweeps");
    }
    ProblemIR ir = ProblemIR::from_pairs(d.n, d.offset, d.linear, d.npairs, d.pi, d.pj, d.q);
    const int64_t ml = ir.max_local_field();
    if (ml > 0x7FFFFFFFll) throw RangeError(who + ": max local field exceeds int32");
    const ColorOrder co = color_order(ir.greedy_coloring());
    LoweredInstance L;
    L.n = d.n;
    L.ncol = co.num_colors;
    L.sweeps = d.sweeps;
    L.dtab = static_cast<uint32_t>(ml);
    L.seed = d.seed;
    L.beta_hi = d.beta_hi;
    L.c = d.sweeps > 1 ? (std::log(d.beta_hi) - std::log(d.beta_lo)) / (d.sweeps - 1) : 0.0;
    L.offset = d.offset;
    L.rp.reserve(d.n + 1);
    L.rp.push_back(0);
    for (uint32_t p = 0; p < d.n; ++p) {
        const uint32_t site = co.order[p];
        L.lin.push_back(ir.linear[site]);
        for (uint32_t e = ir.row_ptr[site]; e < ir.row_ptr[site + 1]; ++e) {
            L.ci.push_back(co.pos_of[ir.col_idx[e]]);
            L.wt.push_back(ir.weights[e]);
        }
        L.rp.push_back(static_cast<uint32_t>(L.ci.size()));
    }
    L.cs = co.class_start;
    L.order = co.order;
    L.pos_of = co.pos_of;
    L.track_every = d.track_every;
    L.polish = d.polish;
    if (d.num_temps) {
        L.pt = 1;
        L.cells = d.num_temps;
        L.ladder = ultimate_ladder(d.temp_max, d.temp_min, d.num_temps);
        for (double b : L.ladder)
            require(std::isfinite(b) && b > 0.0, who + ": ladder temperatures out of range");
    }
    return L;
}

// Everything a persistent launch needs besides the descriptors.
struct PersistentLaunch {
    const InstDesc* insts = nullptr;
    const uint32_t* queue = nullptr;  // tickets: instance << kGangBits | member
    uint32_t tickets = 0;
    unsigned int* next = nullptr;
    uint32_t* bars = nullptr;         // one counter per instance
    uint32_t nbars = 0;
    uint32_t* err = nullptr;
    uint32_t round = 1;
    uint32_t spin_limit = 1u << 24;
    int num_cus = 1;
};

template <int WAVES, bool kPacked>
void launch_persistent_waves(const PersistentLaunch& L, hipStream_t st) {
    int per_cu = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &per_cu, persistent_anneal_kernel<WAVES, kPacked>, WAVES * 64, 0),
          "occupancy");
    if (per_cu < 1) throw Failure(IGE_ERR_HIP, "persistent kernel does not fit on a CU");
    // Every workgroup of the launch must be resident (gang waits rely on it).
    const uint32_t blocks = std::max(
        1u, std::min<uint32_t>(L.tickets, static_cast<uint32_t>(per_cu * L.num_cus)));
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, L.next, 0u, 1u);
    if (L.nbars)
        hipLaunchKernelGGL(fill_u32_kernel, dim3((L.nbars + 255) / 256), dim3(256), 0, st, L.bars,
                           0u, L.nbars);
    hipLaunchKernelGGL((persistent_anneal_kernel<WAVES, kPacked>), dim3(blocks), dim3(WAVES * 64),
                       0, st, L.insts, L.queue, L.tickets, L.next, L.round, L.spin_limit, L.err);
    check(hipGetLastError(), "persistent launch");
}

// Workgroups of persistent_anneal_kernel<waves, kPacked> resident per CU.
template <bool kPacked>
int persistent_per_cu(uint32_t waves) {
    int per_cu = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (waves) {
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_anneal_kernel<1, kPacked>, 64, 0); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_anneal_kernel<2, kPacked>, 128, 0); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_anneal_kernel<4, kPacked>, 256, 0); break;
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_anneal_kernel<8, kPacked>, 512, 0); break;
        case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_anneal_kernel<16, kPacked>, 1024, 0); break;
        default: break;
    }
    return e == hipSuccess && per_cu > 0 ? per_cu : 1;
}

template <bool kPacked>
void launch_persistent(uint32_t waves, const PersistentLaunch& L, hipStream_t st) {
    switch (waves) {
        case 1: launch_persistent_waves<1, kPacked>(L, st); break;
        case 2: launch_persistent_waves<2, kPacked>(L, st); break;
        case 4: launch_persistent_waves<4, kPacked>(L, st); break;
        case 8: launch_persistent_waves<8, kPacked>(L, st); break;
        case 16: launch_persistent_waves<16, kPacked>(L, st); break;
        default: throw Failure(IGE_ERR_INVALID, "waves per instance must be 1, 2, 4, 8 or 16");
    }
}

// Tickets for instances in the given order: gang[i] consecutive tickets each.
inline void build_tickets(const std::vector<uint32_t>& order, const std::vector<uint32_t>& gang,
                          std::vector<uint32_t>& tickets) {
    tickets.clear();
    for (uint32_t id : order)
        for (uint32_t m = 0; m < gang[id]; ++m) tickets.push_back((id << kGangBits) | m);
}

}  // namespace ige

