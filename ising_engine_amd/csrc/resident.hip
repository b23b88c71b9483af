// resident.hip -- whole multispin runs of small regular lattices in one
// workgroup per slot.
//
// A call of many sweeps on a small lattice is bound by launch latency when
// every colour class of every sweep is its own launch (a class of 512 sites
// is one wave's work).  Here the slot's words stay in LDS for the whole call:
// each sweep builds the slot's threshold records from that sweep's row of the
// level table, the colour classes follow each other behind workgroup
// barriers, and the words go back to HBM at the end.  The site update is
// regular_site (multispin_ops.h), with the same uniforms, thresholds and tie
// resolution as the colour-class launches, so a state ends with the same
// bits either way.  Slots are independent (no exchange), so one workgroup per
// slot needs no device-wide barrier.
#include <hip/hip_runtime.h>

#include "ige_math.h"
#include "kernels.hpp"
#include "multispin_ops.h"

namespace ige {

namespace {

using namespace msc;

constexpr uint32_t kResWaves = 8;
constexpr uint32_t kResThreads = kResWaves * kWave;
constexpr uint32_t kResRecCap = 64;  // tie records per wave and class pass
constexpr uint32_t kFlushSites = 8;  // sites per lane between counter flushes

template <int DEG>
__global__ __launch_bounds__(kResThreads) void resident_sweeps_kernel(
    const uint32_t* __restrict__ ci, const uint32_t* __restrict__ sgn,
    const uint32_t* __restrict__ class_start, uint32_t ncol, uint64_t* spins, uint32_t n,
    int32_t w0, const uint32_t* __restrict__ lev_rows, uint32_t row_stride,
    const uint32_t* __restrict__ temp_of, int64_t* energy, unsigned long long* accepted,
    uint32_t sweep0, uint32_t sweeps, uint32_t key0, uint32_t key1) {
    static_assert(3 * kFlushSites < (1u << (kSW - 1)), "signed counters too narrow");
    constexpr int LMAX = DEG - DEG / 2;
    extern __shared__ uint64_t words[];  // this slot's n words
    __shared__ unsigned long long recs[(DEG + 1) * kQStride];
    __shared__ unsigned long long rec_und[kResWaves][kResRecCap];
    __shared__ unsigned long long rec_l1[kResWaves][kResRecCap];
    __shared__ unsigned long long rec_l2[kResWaves][DEG == 6 ? kResRecCap : 1];
    __shared__ uint32_t rec_pos[kResWaves][kResRecCap];
    __shared__ uint32_t qcount[kResWaves];
    __shared__ unsigned long long edef[kResWaves][kWave];
    __shared__ unsigned long long blk_e[kWave];
    __shared__ uint32_t blk_acc;
    const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1);
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t slot = blockIdx.x;
    uint64_t* g = spins + (uint64_t)slot * n;
    for (uint32_t i = tid; i < n; i += kResThreads) words[i] = g[i];
    edef[wv][lane] = 0;
    if (wv == 0) blk_e[lane] = 0;
    if (tid == 0) blk_acc = 0;
    const uint32_t tr = temp_of[slot * kWave + lane];
    const TieRecords tq{rec_und[wv], rec_l1[wv], rec_l2[wv], rec_pos[wv], &qcount[wv], kResRecCap};
    uint64_t sacc[kSW];
#pragma unroll
    for (int b = 0; b < kSW; ++b) sacc[b] = 0;
    int64_t e_lane = 0;  // this wave's dE of replica `lane` from flushed counters
    uint32_t nacc = 0, pending = 0;
    __syncthreads();
    for (uint32_t k = 0; k < sweeps; ++k) {
        const uint32_t sweep = sweep0 + k;
        // This sweep's threshold records: wave j < LMAX builds level 2 + 2j.
        if (wv < (uint32_t)LMAX)
            build_records(recs, lev_rows + (uint64_t)k * row_stride, tr, DEG, 2 + 2 * wv, 2 * DEG,
                          lane);
        __syncthreads();
        for (uint32_t c = 0; c < ncol; ++c) {
            const uint32_t cb = class_start[c], ce = class_start[c + 1];
            if (lane == 0) qcount[wv] = 0;
            __builtin_amdgcn_wave_barrier();
            for (uint32_t base = cb + wv * kWave; base < ce; base += kResThreads) {
                const uint32_t p = base + lane;
                if (p < ce)
                    regular_site<DEG>(words, words, p, load_row<DEG>(words, p, ci, sgn), recs,
                                      slot, sweep, key0, key1, w0, sacc, nacc, tq, edef[wv]);
                if (++pending == kFlushSites) {  // wave-uniform
                    e_lane += (int64_t)w0 * wave_column_sum_signed<kSW>(sacc, lane);
#pragma unroll
                    for (int b = 0; b < kSW; ++b) sacc[b] = 0;
                    pending = 0;
                }
            }
            __threadfence_block();
            __builtin_amdgcn_wave_barrier();
            drain_ties<DEG>(tq, words, recs, slot, sweep, key0, key1, w0, edef[wv], lane, nacc);
            __syncthreads();  // the next class reads this one's words
        }
    }
    e_lane += (int64_t)w0 * wave_column_sum_signed<kSW>(sacc, lane);
    for (uint32_t i = tid; i < n; i += kResThreads) g[i] = words[i];
    e_lane += (int64_t)edef[wv][lane];
    for (uint32_t off = 32; off > 0; off >>= 1) nacc += __shfl_xor(nacc, (int)off, 64);
    atomicAdd(&blk_e[lane], (unsigned long long)e_lane);
    if (lane == 0) atomicAdd(&blk_acc, nacc);
    __syncthreads();
    if (wv == 0) {
        energy[slot * kWave + lane] += (int64_t)blk_e[lane];
        if (lane == 0 && blk_acc) atomicAdd(accepted, (unsigned long long)blk_acc);
    }
}

constexpr uint32_t kClsWaves = kFuseMaxClassSites / kWave;
constexpr uint32_t kClsThreads = kClsWaves * kWave;
constexpr uint32_t kClsQueueCap = 256;  // tied replicas per wave and class (any degree)
constexpr uint32_t kClsRecCap = 32;     // tie records per wave and class (regular lattices)

// Colour classes [c0, c1) of one sweep in one workgroup per slot, one after
// the other behind workgroup barriers: classes too small to fill the GPU,
// whose separate launches would cost more than their work.  Words stay in
// HBM (the workgroup's own barriers order them); the update is the same
// regular_site / irregular_site as the colour-class kernel.
template <int B, int W, int DEG>
__global__ __launch_bounds__(kClsThreads) void classes_sweep_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const uint32_t* __restrict__ sgn, const uint32_t* __restrict__ class_start, uint32_t c0,
    uint32_t c1, uint64_t* spins, uint32_t n, uint32_t dmax, int32_t w0,
    const uint32_t* __restrict__ lev_thr, const uint32_t* __restrict__ temp_of, int64_t* energy,
    unsigned long long* accepted, uint32_t sweep_rel, uint32_t key0, uint32_t key1,
    const uint32_t* __restrict__ clock) {
    static_assert(3 * kFlushSites < (1u << (kSW - 1)) && W >= B + 3, "counters too narrow");
    extern __shared__ unsigned long long recs[];  // (dmax + 1) * kQStride
    __shared__ uint32_t queue[kClsWaves][DEG > 0 ? 1 : kClsQueueCap];
    __shared__ unsigned long long rec_und[kClsWaves][DEG > 0 ? kClsRecCap : 1];
    __shared__ unsigned long long rec_l1[kClsWaves][DEG > 0 ? kClsRecCap : 1];
    __shared__ unsigned long long rec_l2[kClsWaves][DEG == 6 ? kClsRecCap : 1];
    __shared__ uint32_t rec_pos[kClsWaves][DEG > 0 ? kClsRecCap : 1];
    __shared__ uint32_t qcount[kClsWaves];
    __shared__ unsigned long long edef[kClsWaves][kWave];
    __shared__ unsigned long long blk_e[kWave];
    __shared__ uint32_t blk_acc;
    const uint32_t sweep = clocked(clock, kClockSweep, sweep_rel);
    const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1);
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t slot = blockIdx.x;
    uint64_t* S = spins + (uint64_t)slot * n;
    {
        // The slot's threshold records (levels 2, 4, ... on a regular lattice).
        const uint32_t tr = temp_of[slot * kWave + lane];
        constexpr uint32_t m_step = DEG > 0 ? 2 : 1;
        build_records(recs, lev_thr, tr, dmax, (DEG > 0 ? 2 : 0) + m_step * wv,
                      m_step * kClsWaves, lane);
        edef[wv][lane] = 0;
        if (wv == 0) blk_e[lane] = 0;
        if (tid == 0) blk_acc = 0;
        __syncthreads();
    }
    const TieRecords tq{rec_und[wv], rec_l1[wv], rec_l2[wv], rec_pos[wv], &qcount[wv], kClsRecCap};
    uint64_t sacc[kSW], kacc[W], dacc[W];
#pragma unroll
    for (int b = 0; b < kSW; ++b) sacc[b] = 0;
#pragma unroll
    for (int b = 0; b < W; ++b) kacc[b] = dacc[b] = 0;
    int64_t e_lane = 0;
    uint32_t nacc = 0, pending = 0;
    auto flush = [&]() {
        if constexpr (DEG > 0) {
            e_lane += (int64_t)w0 * wave_column_sum_signed<kSW>(sacc, lane);
#pragma unroll
            for (int b = 0; b < kSW; ++b) sacc[b] = 0;
        } else {
            const int64_t ksum = wave_column_sum<W>(kacc, lane);
            const int64_t dsum = wave_column_sum<W>(dacc, lane);
            e_lane += ((int64_t)w0 * (2 * ksum - dsum)) / 2;
#pragma unroll
            for (int b = 0; b < W; ++b) kacc[b] = dacc[b] = 0;
        }
        pending = 0;
    };
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t cb = class_start[c], ce = class_start[c + 1];
        if (lane == 0) qcount[wv] = 0;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t base = cb + wv * kWave; base < ce; base += kClsThreads) {
            const uint32_t p = base + lane;
            if (p < ce) {
                if constexpr (DEG > 0)
                    regular_site<DEG>(S, S, p, load_row<DEG>(S, p, ci, sgn), recs, slot, sweep,
                                      key0, key1, w0, sacc, nacc, tq, edef[wv]);
                else
                    irregular_site<B, W>(S, S, p, p - cb, rp, ci, sgn, recs, slot, sweep, key0,
                                         key1, w0, kacc, dacc, nacc, queue[wv], &qcount[wv],
                                         kClsQueueCap, edef[wv]);
            }
            if (++pending == kFlushSites) flush();  // wave-uniform
        }
        __threadfence_block();
        __builtin_amdgcn_wave_barrier();
        if constexpr (DEG > 0)
            drain_ties<DEG>(tq, S, recs, slot, sweep, key0, key1, w0, edef[wv], lane, nacc);
        else
            drain_queue(queue[wv], &qcount[wv], kClsQueueCap, cb, S, recs, slot, sweep, key0,
                        key1, w0, edef[wv], lane, nacc);
        __syncthreads();  // the next class reads this one's words
    }
    flush();
    e_lane += (int64_t)edef[wv][lane];
    for (uint32_t off = 32; off > 0; off >>= 1) nacc += __shfl_xor(nacc, (int)off, 64);
    atomicAdd(&blk_e[lane], (unsigned long long)e_lane);
    if (lane == 0) atomicAdd(&blk_acc, nacc);
    __syncthreads();
    if (wv == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(energy + slot * kWave + lane), blk_e[lane]);
        if (lane == 0 && blk_acc) atomicAdd(accepted, (unsigned long long)blk_acc);
    }
}

// Rows of the level table: out[(k * cells + t) * (dmax + 1) + m] =
// metropolis_threshold(beta_rows[k * cells + t], m * w0 / 2), 0 for m = 0.
__global__ void level_rows_kernel(const double* __restrict__ beta_rows, uint32_t* __restrict__ out,
                                  uint64_t total, uint32_t cells, uint32_t dmax, int32_t w0) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint64_t kt = i / (dmax + 1);
    const uint32_t m = (uint32_t)(i % (dmax + 1));
    out[i] = m == 0 ? 0u : metropolis_threshold(beta_rows[kt], ((int64_t)m * w0) / 2);
}

}  // namespace

hipError_t launch_level_rows(const DevModel& m, uint32_t cells, const double* beta_rows,
                             uint32_t rows, uint32_t* out, hipStream_t stream) {
    const uint64_t total = (uint64_t)rows * cells * (m.dmax + 1);
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(level_rows_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0,
                       stream, beta_rows, out, total, cells, m.dmax, m.w0);
    return hipGetLastError();
}

hipError_t launch_classes_sweep(const DevModel& m, const DevState& s,
                                const uint32_t* class_start, uint32_t c0, uint32_t c1,
                                uint32_t sweep, uint64_t seed, const uint32_t* lev_thr,
                                hipStream_t stream) {
    if (c1 <= c0) return hipSuccess;
    const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);
    const uint32_t lds = (m.dmax + 1) * kQStride * sizeof(unsigned long long);
#define IGE_CLS_LAUNCH(B, W, DEG)                                                                  \
    hipLaunchKernelGGL((classes_sweep_kernel<B, W, DEG>), dim3(s.cells), dim3(kClsThreads), lds,   \
                       stream, m.rp, m.ci, m.sgn, class_start, c0, c1, s.spins, m.n, m.dmax,       \
                       m.w0, lev_thr, s.temp_of, s.energy, s.accepted, sweep, key0, key1, s.clock)
    if (m.regular_degree == 6) {
        IGE_CLS_LAUNCH(3, 6, 6);
    } else if (m.regular_degree == 4) {
        IGE_CLS_LAUNCH(3, 6, 4);
    } else if (m.dmax < 8) {
        IGE_CLS_LAUNCH(3, 6, 0);
    } else if (m.dmax < 16) {
        IGE_CLS_LAUNCH(4, 7, 0);
    } else if (m.dmax < 32) {
        IGE_CLS_LAUNCH(5, 8, 0);
    } else {
        return hipErrorInvalidValue;
    }
#undef IGE_CLS_LAUNCH
    return hipGetLastError();
}

hipError_t launch_resident_sweeps(const DevModel& m, const DevState& s,
                                  const uint32_t* class_start, uint32_t ncol,
                                  const uint32_t* lev_rows, uint32_t row_stride, uint32_t sweep0,
                                  uint32_t sweeps, uint64_t seed, hipStream_t stream) {
    if (m.n > kResidentMaxSites || (m.regular_degree != 4 && m.regular_degree != 6))
        return hipErrorInvalidValue;
    if (sweeps == 0) return hipSuccess;
    const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);
    const size_t lds = (size_t)m.n * sizeof(uint64_t);
    if (m.regular_degree == 6)
        hipLaunchKernelGGL(resident_sweeps_kernel<6>, dim3(s.cells), dim3(kResThreads), lds,
                           stream, m.ci, m.sgn, class_start, ncol, s.spins, m.n, m.w0, lev_rows,
                           row_stride, s.temp_of, s.energy, s.accepted, sweep0, sweeps, key0,
                           key1);
    else
        hipLaunchKernelGGL(resident_sweeps_kernel<4>, dim3(s.cells), dim3(kResThreads), lds,
                           stream, m.ci, m.sgn, class_start, ncol, s.spins, m.n, m.w0, lev_rows,
                           row_stride, s.temp_of, s.energy, s.accepted, sweep0, sweeps, key0,
                           key1);
    return hipGetLastError();
}

}  // namespace ige
