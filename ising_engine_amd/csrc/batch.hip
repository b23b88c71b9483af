// batch.hip -- thousands of independent Ising/QUBO instances of uneven size
// in one persistent launch over a packed CSR arena.
//
// Each instance is one cell of 64 replicas annealed over its own geometric
// beta schedule.  A workgroup takes the next instance from a device work
// queue (instances ordered largest first), runs all of its sweeps with
// workgroup barriers between colour classes, and returns for the next one:
// no grid-wide synchronisation is ever needed.  The per-sweep Metropolis
// thresholds are built in LDS by the workgroup itself.
//
// The uniforms are keyed by (instance seed, position, slot 0, sweep), so an
// instance's trajectory is identical to a single state (ige_state_create +
// ige_run with the schedule of ige_anneal_schedule), whichever workgroup,
// co-resident instance or launch mode (persistent or plain) runs it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ising_gpu.h"
#include "abi_util.hpp"
#include "ige_math.h"
#include "kernels.hpp"
#include "problem_ir.hpp"
#include "sweep_generic.h"

namespace ige {

struct BatchInst {
    uint32_t n;
    uint32_t ncol;
    uint32_t pos_base;  // spins / lin / order arenas
    uint32_t rp_base;   // rp arena (n + 1 entries, absolute edge indices)
    uint32_t cs_base;   // class-start arena (ncol + 1 entries)
    uint32_t sweeps;
    uint32_t dtab;      // max local field: bound on dE
    uint32_t key0, key1;
    uint32_t pad;
    double beta_hi;
    double c;           // beta_k = beta_hi * det_exp(-(sweeps - 1 - k) * c)
    int64_t offset;
};

__host__ __device__ __forceinline__ double anneal_beta(double beta_hi, double c, uint32_t sweeps,
                                                       uint32_t k) {
    return sweeps > 1 ? beta_hi * det_exp(-(double)(sweeps - 1 - k) * c) : beta_hi;
}

namespace {

constexpr uint32_t kWave = 64;
constexpr uint32_t kLdsEntries = 4096;  // threshold table entries kept in LDS

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void batch_persistent_kernel(
    const BatchInst* __restrict__ insts, const uint32_t* __restrict__ queue, uint32_t count,
    unsigned int* next, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin,
    const uint32_t* __restrict__ cs, const uint64_t* spins_rd, uint64_t* spins_wr,
    int64_t* energy, unsigned long long* accepted, uint32_t k0, uint32_t k1) {
    __shared__ uint32_t thr_lds[kLdsEntries];
    __shared__ int64_t part[WAVES][kWave];
    __shared__ uint32_t s_item;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = uniform(threadIdx.x >> 6);
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(next, 1u);
        __syncthreads();
        const uint32_t item = uniform(s_item);
        __syncthreads();
        if (item >= count) break;
        const uint32_t id = queue[item];
        const BatchInst I = insts[id];
        const bool table = I.dtab < kLdsEntries;
        const uint32_t* irp = rp + I.rp_base;
        const int32_t* ilin = lin + I.pos_base;
        const uint64_t* S = spins_rd + I.pos_base;
        uint64_t* Wr = spins_wr + I.pos_base;
        const uint32_t kend = min(k1, I.sweeps);
        int64_t eacc = 0;
        uint32_t nacc = 0;
        for (uint32_t k = k0; k < kend; ++k) {
            const double beta = anneal_beta(I.beta_hi, I.c, I.sweeps, k);
            if (table) {
                for (uint32_t d = threadIdx.x; d <= I.dtab; d += WAVES * 64)
                    thr_lds[d] = d == 0 ? 0xFFFFFFFFu : metropolis_threshold(beta, (int64_t)d);
                __syncthreads();
            }
            auto thr = [&](int32_t de) {
                return table ? thr_lds[de] : metropolis_threshold(beta, (int64_t)de);
            };
            for (uint32_t c = 0; c < I.ncol; ++c) {
                const uint32_t cb = cs[I.cs_base + c], ce = cs[I.cs_base + c + 1];
                if (ce > cb) {
                    for (uint32_t g = (cb >> 2) + wave; g <= ((ce - 1) >> 2); g += WAVES) {
                        const uint32_t p0 = max(g * 4u, cb), p1 = min(g * 4u + 4u, ce);
                        generic_positions(irp, ci, wt, ilin, S, Wr, p0, p1, 0u, k, I.key0,
                                          I.key1, lane, thr, eacc, nacc);
                    }
                }
                __syncthreads();
            }
        }
        part[wave][lane] = eacc;
        for (uint32_t off = 32; off > 0; off >>= 1) nacc += __shfl_xor(nacc, (int)off, 64);
        if (lane == 0 && nacc) atomicAdd(accepted + id, (unsigned long long)nacc);
        __syncthreads();
        if (wave == 0) {
            int64_t tot = 0;
#pragma unroll
            for (int w = 0; w < WAVES; ++w) tot += part[w][lane];
            energy[(uint64_t)id * kWave + lane] += tot;
        }
        __syncthreads();
    }
}

// Default initial words and exact energies, one workgroup per instance.
__global__ __launch_bounds__(256) void batch_init_kernel(const BatchInst* __restrict__ insts,
                                                         const uint32_t* __restrict__ order,
                                                         uint64_t* spins) {
    const BatchInst I = insts[blockIdx.x];
    for (uint32_t p = threadIdx.x; p < I.n; p += blockDim.x) {
        const U32x4 o = philox4x32_10(order[I.pos_base + p], 0u, 0u, kTagInit << 24, I.key0,
                                      I.key1);
        spins[I.pos_base + p] = (uint64_t)o.v[0] | ((uint64_t)o.v[1] << 32);
    }
}

__global__ __launch_bounds__(256) void batch_energy_kernel(
    const BatchInst* __restrict__ insts, const uint32_t* __restrict__ rp,
    const uint32_t* __restrict__ ci, const int32_t* __restrict__ wt,
    const int32_t* __restrict__ lin, const uint64_t* __restrict__ spins, int64_t* energy) {
    __shared__ int64_t part[4][kWave];
    const BatchInst I = insts[blockIdx.x];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wv = uniform(threadIdx.x >> 6);
    const uint64_t* S = spins + I.pos_base;
    const uint32_t* irp = rp + I.rp_base;
    int64_t a = 0;
    for (uint32_t p = wv; p < I.n; p += 4) {
        int64_t f = 2 * (int64_t)lin[I.pos_base + p];
        for (uint32_t e = irp[p]; e < irp[p + 1]; ++e) f += lane_bit(S[ci[e]]) ? wt[e] : 0;
        a += lane_bit(S[p]) ? f : 0;
    }
    part[wv][lane] = a;
    __syncthreads();
    if (threadIdx.x < kWave)
        energy[(uint64_t)blockIdx.x * kWave + lane] =
            I.offset + (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]) / 2;
}

__global__ void thr_value_kernel(double beta, uint32_t* thr, uint32_t dtab) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d <= dtab) thr[d] = d == 0 ? 0xFFFFFFFFu : metropolis_threshold(beta, (int64_t)d);
}

}  // namespace
}  // namespace ige

using ige::abi::check;
using ige::abi::DevBuf;
using ige::abi::Failure;
using ige::abi::guarded;
using ige::abi::require;

struct ige_batch {
    ige_ctx* ctx = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t count = 0;
    std::vector<ige::BatchInst> host_insts;
    std::vector<std::vector<uint32_t>> orders;  // per instance: position -> site
    std::vector<uint32_t> queue;                 // instances, largest work first
    std::vector<uint32_t> host_cs;               // class-start arena
    uint64_t total_sites = 0, total_attempts = 0;
    uint32_t max_ncol = 0, max_dtab = 0;
    uint32_t tuned_waves = 0;
    DevBuf<ige::BatchInst> insts;
    DevBuf<uint32_t> q, rp, ci, cs, order, zero_temp, thr;
    DevBuf<int32_t> wt, lin;
    DevBuf<uint64_t> spins;
    DevBuf<int64_t> energy;
    DevBuf<unsigned long long> accepted;
    DevBuf<unsigned int> next;
    int num_cus = 0;
    ~ige_batch() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace {

void batch_reset(ige_batch* b) {
    hipLaunchKernelGGL(ige::batch_init_kernel, dim3(b->count), dim3(256), 0, b->stream,
                       b->insts.p, b->order.p, b->spins.p);
    check(hipGetLastError(), "batch init");
    hipLaunchKernelGGL(ige::batch_energy_kernel, dim3(b->count), dim3(256), 0, b->stream,
                       b->insts.p, b->rp.p, b->ci.p, b->wt.p, b->lin.p, b->spins.p, b->energy.p);
    check(hipGetLastError(), "batch energy");
    std::vector<unsigned long long> z(b->count, 0);
    b->accepted.upload(z.data(), z.size(), b->stream);
    check(hipStreamSynchronize(b->stream), "batch reset");
}

template <int WAVES>
void launch_persistent(ige_batch* b, uint32_t k0, uint32_t k1) {
    int per_cu = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &per_cu, ige::batch_persistent_kernel<WAVES>, WAVES * 64, 0),
          "occupancy");
    const uint32_t blocks =
        std::max(1u, std::min<uint32_t>(b->count, static_cast<uint32_t>(per_cu * b->num_cus)));
    const unsigned int zero = 0;
    b->next.upload(&zero, 1, b->stream);
    hipLaunchKernelGGL(ige::batch_persistent_kernel<WAVES>, dim3(blocks), dim3(WAVES * 64), 0,
                       b->stream, b->insts.p, b->q.p, b->count, b->next.p, b->rp.p, b->ci.p,
                       b->wt.p, b->lin.p, b->cs.p, b->spins.p, b->spins.p, b->energy.p,
                       b->accepted.p, k0, k1);
    check(hipGetLastError(), "batch persistent launch");
}

void run_persistent(ige_batch* b, uint32_t waves, uint32_t k0, uint32_t k1) {
    switch (waves) {
        case 1: launch_persistent<1>(b, k0, k1); break;
        case 2: launch_persistent<2>(b, k0, k1); break;
        case 4: launch_persistent<4>(b, k0, k1); break;
        case 8: launch_persistent<8>(b, k0, k1); break;
        default: throw Failure(IGE_ERR_INVALID, "waves per instance must be 1, 2, 4 or 8");
    }
}

// The same sweeps as the persistent kernel, issued the way one would without
// batching: per instance, per sweep, one threshold launch and one sweep
// launch per colour class (the single-state kernels).
void run_plain(ige_batch* b, uint32_t k0, uint32_t k1) {
    for (uint32_t id = 0; id < b->count; ++id) {
        const ige::BatchInst& I = b->host_insts[id];
        ige::DevModel m;
        m.n = I.n;
        m.rp = b->rp.p + I.rp_base;
        m.ci = b->ci.p;
        m.wt = b->wt.p;
        m.lin = b->lin.p + I.pos_base;
        ige::DevState s;
        s.cells = 1;
        s.spins = b->spins.p + I.pos_base;
        s.energy = b->energy.p + static_cast<size_t>(id) * 64;
        s.temp_of = b->zero_temp.p;
        s.accepted = b->accepted.p + id;
        const uint64_t seed = (uint64_t)I.key0 | ((uint64_t)I.key1 << 32);
        const std::vector<uint32_t> csv(b->host_cs.begin() + I.cs_base,
                                        b->host_cs.begin() + I.cs_base + I.ncol + 1);
        for (uint32_t k = k0; k < std::min(k1, I.sweeps); ++k) {
            const double beta = ige::anneal_beta(I.beta_hi, I.c, I.sweeps, k);
            hipLaunchKernelGGL(ige::thr_value_kernel, dim3((I.dtab + 256) / 256), dim3(256), 0,
                               b->stream, beta, b->thr.p, I.dtab);
            check(hipGetLastError(), "thr");
            for (uint32_t c = 0; c < I.ncol; ++c)
                check(ige::launch_sweep_generic(m, s, csv[c], csv[c + 1], k, seed, b->thr.p,
                                                I.dtab, b->stream),
                      "plain sweep");
        }
    }
}

}  // namespace

extern "C" {

int32_t ige_anneal_schedule(double beta_lo, double beta_hi, uint32_t sweeps, double* out) {
    return guarded([&] {
        require(out != nullptr && sweeps >= 1, "bad arguments");
        require(std::isfinite(beta_lo) && std::isfinite(beta_hi) && beta_lo > 0 &&
                    beta_hi >= beta_lo && std::log(beta_hi / beta_lo) <= 40.0,
                "need 0 < beta_lo <= beta_hi <= e^40 beta_lo");
        const double c = sweeps > 1 ? (std::log(beta_hi) - std::log(beta_lo)) / (sweeps - 1) : 0.0;
        for (uint32_t k = 0; k < sweeps; ++k) out[k] = ige::anneal_beta(beta_hi, c, sweeps, k);
    });
}

int32_t ige_batch_create(ige_ctx* ctx, uint32_t count, const ige_instance_desc* desc,
                         ige_batch** out) {
    return guarded([&] {
        require(ctx && desc && out && count >= 1, "bad arguments");
        auto b = std::make_unique<ige_batch>();
        b->ctx = ctx;
        b->device = ige::abi::context_device(ctx);
        check(hipSetDevice(b->device), "hipSetDevice");
        check(hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount,
                                    b->device),
              "device attribute");
        check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
        b->count = count;
        std::vector<uint32_t> rp, ci, cs, order;
        std::vector<int32_t> wt, lin;
        std::vector<uint64_t> work(count);
        for (uint32_t i = 0; i < count; ++i) {
            const ige_instance_desc& d = desc[i];
            require(d.n >= 1 && d.linear && (d.npairs == 0 || (d.pi && d.pj && d.q)),
                    "instance " + std::to_string(i) + ": bad model arrays");
            require(d.sweeps >= 1, "instance " + std::to_string(i) + ": sweeps must be >= 1");
            ige::ProblemIR ir =
                ige::ProblemIR::from_pairs(d.n, d.offset, d.linear, d.npairs, d.pi, d.pj, d.q);
            const int64_t ml = ir.max_local_field();
            if (ml > 0x7FFFFFFFll) throw ige::RangeError("instance max local field exceeds int32");
            const auto colors = ir.greedy_coloring();
            const ige::ColorOrder co = ige::color_order(colors);
            ige::BatchInst I{};
            I.n = d.n;
            I.ncol = co.num_colors;
            I.pos_base = static_cast<uint32_t>(lin.size());
            I.rp_base = static_cast<uint32_t>(rp.size());
            I.cs_base = static_cast<uint32_t>(cs.size());
            I.sweeps = d.sweeps;
            I.dtab = static_cast<uint32_t>(ml);
            I.key0 = static_cast<uint32_t>(d.seed);
            I.key1 = static_cast<uint32_t>(d.seed >> 32);
            require(std::isfinite(d.beta_lo) && std::isfinite(d.beta_hi) && d.beta_lo > 0 &&
                        d.beta_hi >= d.beta_lo && std::log(d.beta_hi / d.beta_lo) <= 40.0,
                    "instance " + std::to_string(i) + ": need 0 < beta_lo <= beta_hi");
            I.beta_hi = d.beta_hi;
            I.c = d.sweeps > 1 ? (std::log(d.beta_hi) - std::log(d.beta_lo)) / (d.sweeps - 1) : 0.0;
            I.offset = d.offset;
            // position-space CSR with absolute edge indices, local neighbour positions
            const uint32_t e_base = static_cast<uint32_t>(ci.size());
            rp.push_back(e_base);
            for (uint32_t p = 0; p < d.n; ++p) {
                const uint32_t site = co.order[p];
                lin.push_back(ir.linear[site]);
                order.push_back(site);
                for (uint32_t e = ir.row_ptr[site]; e < ir.row_ptr[site + 1]; ++e) {
                    ci.push_back(co.pos_of[ir.col_idx[e]]);
                    wt.push_back(ir.weights[e]);
                }
                rp.push_back(static_cast<uint32_t>(ci.size()));
            }
            for (uint32_t c = 0; c <= co.num_colors; ++c) cs.push_back(co.class_start[c]);
            b->host_insts.push_back(I);
            b->max_ncol = std::max(b->max_ncol, I.ncol);
            b->max_dtab = std::max(b->max_dtab, I.dtab);
            b->total_sites += d.n;
            b->total_attempts += static_cast<uint64_t>(d.n) * 64ull * d.sweeps;
            work[i] = static_cast<uint64_t>(d.n + ir.nnz()) * d.sweeps;
        }
        b->host_cs = cs;
        b->queue.resize(count);
        std::iota(b->queue.begin(), b->queue.end(), 0u);
        std::stable_sort(b->queue.begin(), b->queue.end(),
                         [&](uint32_t a, uint32_t c) { return work[a] > work[c]; });
        b->insts.alloc(count);
        b->insts.upload(b->host_insts.data(), count, b->stream);
        b->q.alloc(count);
        b->q.upload(b->queue.data(), count, b->stream);
        auto put = [&](auto& buf, const auto& v) {
            buf.alloc(v.size());
            buf.upload(v.data(), v.size(), b->stream);
        };
        put(b->rp, rp);
        put(b->ci, ci);
        put(b->cs, cs);
        put(b->order, order);
        put(b->wt, wt);
        put(b->lin, lin);
        b->spins.alloc(lin.size());
        b->energy.alloc(static_cast<size_t>(count) * 64);
        b->accepted.alloc(count);
        b->next.alloc(1);
        b->thr.alloc(static_cast<size_t>(b->max_dtab) + 1);
        std::vector<uint32_t> zt(64, 0);
        put(b->zero_temp, zt);
        check(hipStreamSynchronize(b->stream), "batch upload");
        b->orders.clear();
        batch_reset(b.get());
        *out = b.release();
    });
}

int32_t ige_batch_destroy(ige_batch* b) {
    if (b) (void)hipSetDevice(b->device);
    delete b;
    return IGE_OK;
}

int32_t ige_batch_reset(ige_batch* b) {
    return guarded([&] {
        require(b != nullptr, "NULL batch");
        check(hipSetDevice(b->device), "hipSetDevice");
        batch_reset(b);
    });
}

int32_t ige_batch_run(ige_batch* b, uint32_t mode, uint32_t waves, uint32_t k0, uint32_t k1) {
    return guarded([&] {
        require(b != nullptr, "NULL batch");
        require(k0 <= k1, "need k0 <= k1");
        check(hipSetDevice(b->device), "hipSetDevice");
        if (mode == IGE_BATCH_PLAIN) {
            run_plain(b, k0, k1);
        } else {
            require(mode == IGE_BATCH_PERSISTENT, "unknown batch mode");
            const uint32_t w = waves ? waves : (b->tuned_waves ? b->tuned_waves : 4);
            run_persistent(b, w, k0, k1);
        }
    });
}

int32_t ige_batch_autotune(ige_batch* b, uint32_t probe_sweeps, uint32_t* waves_out,
                           double* ms_out) {
    return guarded([&] {
        require(b != nullptr && probe_sweeps >= 1, "bad arguments");
        check(hipSetDevice(b->device), "hipSetDevice");
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "event");
        check(hipEventCreate(&e1), "event");
        const uint32_t cand[4] = {1, 2, 4, 8};
        float best = 0.f;
        uint32_t best_w = 4;
        for (int i = 0; i < 4; ++i) {
            run_persistent(b, cand[i], 0, probe_sweeps);  // warm
            check(hipEventRecord(e0, b->stream), "event");
            run_persistent(b, cand[i], 0, probe_sweeps);
            check(hipEventRecord(e1, b->stream), "event");
            check(hipEventSynchronize(e1), "event");
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, e0, e1), "event");
            if (ms_out) ms_out[i] = ms;
            if (i == 0 || ms < best) {
                best = ms;
                best_w = cand[i];
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->tuned_waves = best_w;
        if (waves_out) *waves_out = best_w;
        batch_reset(b);  // the probes advanced the state; start over
    });
}

int32_t ige_batch_synchronize(ige_batch* b) {
    return guarded([&] {
        require(b != nullptr, "NULL batch");
        check(hipStreamSynchronize(b->stream), "batch synchronize");
    });
}

int32_t ige_batch_info(const ige_batch* b, uint64_t* total_sites, uint64_t* total_attempts) {
    return guarded([&] {
        require(b != nullptr, "NULL batch");
        if (total_sites) *total_sites = b->total_sites;
        if (total_attempts) *total_attempts = b->total_attempts;
    });
}

int32_t ige_batch_read(ige_batch* b, uint32_t index, uint64_t* words, int64_t* energies,
                       uint64_t* accepted) {
    return guarded([&] {
        require(b != nullptr && index < b->count, "bad batch index");
        check(hipSetDevice(b->device), "hipSetDevice");
        const ige::BatchInst& I = b->host_insts[index];
        check(hipStreamSynchronize(b->stream), "batch read");
        if (words) {
            std::vector<uint64_t> w(I.n);
            std::vector<uint32_t> ord(I.n);
            check(hipMemcpy(w.data(), b->spins.p + I.pos_base, I.n * sizeof(uint64_t),
                            hipMemcpyDeviceToHost),
                  "read words");
            check(hipMemcpy(ord.data(), b->order.p + I.pos_base, I.n * sizeof(uint32_t),
                            hipMemcpyDeviceToHost),
                  "read order");
            for (uint32_t p = 0; p < I.n; ++p) words[ord[p]] = w[p];
        }
        if (energies)
            check(hipMemcpy(energies, b->energy.p + static_cast<size_t>(index) * 64,
                            64 * sizeof(int64_t), hipMemcpyDeviceToHost),
                  "read energies");
        if (accepted) {
            unsigned long long a = 0;
            check(hipMemcpy(&a, b->accepted.p + index, sizeof(a), hipMemcpyDeviceToHost),
                  "read accepted");
            *accepted = a;
        }
    });
}

}  // extern "C"
