// descent.hip -- deterministic steepest descent over every replica of a state,
// and the best-seen (incumbent) replica.
//
// One descent step: every replica flips the site of its most negative dE
// (lowest site index on a tie; dE = 0 never moves), as SteepestDescent::apply
// does per sweep (src/engine_v2/operators/steepest_descent.rs:80-124); run to
// the end it is steepest_descent_1opt of src/solver/local_search.rs:19-60 for
// each replica.  A step is three launches on the state's stream:
//   scan    lane = replica.  dE of each (slot, site, lane) -- from the CSR rows
//           with the neighbour words as lane masks, or from the matrix-core
//           fields of a dense model -- folded into a per-lane minimum of the
//           key (dE + 2^31) << 32 | site, then one atomicMin per lane and wave
//           into keys[slot][lane];
//   apply   lane = replica.  Flips the chosen bit (64-bit atomic xor: lanes of
//           one slot may pick the same site), adds dE to the tracked energy,
//           counts the flip and clears the key for the next step;
//   finish  one thread.  A step without a flip marks the run done, after which
//           every later launch of the batch returns at once.
// dE is exact in int32: |field| <= max local field < 2^31.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "sweep_generic.h"

namespace ige {

namespace {

constexpr uint32_t kWave = 64;
constexpr uint32_t kBlock = 256;
constexpr uint32_t kWavesPerBlock = kBlock / kWave;
constexpr uint64_t kNoKey = ~0ull;

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint64_t descent_key(int32_t de, uint32_t site) {
    return de < 0 ? ((uint64_t)((uint32_t)de ^ 0x80000000u) << 32) | site : kNoKey;
}

__device__ __forceinline__ void fold_key(unsigned long long* keys, uint64_t kmin) {
    if (kmin != kNoKey) atomicMin(keys, (unsigned long long)kmin);
}

// Sparse scan: wave w covers positions [p0, p0 + chunk) of slot w / waves_per_slot.
__global__ __launch_bounds__(kBlock) void descent_scan_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin,
    const uint32_t* __restrict__ order, const uint8_t* __restrict__ clamped,
    const uint64_t* __restrict__ spins, uint32_t n, uint32_t cells, uint32_t chunk,
    uint32_t waves_per_slot, unsigned long long* __restrict__ keys,
    const unsigned long long* __restrict__ ctl) {
    if (ctl[kDescentDone]) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t w = uniform(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    if (w >= (uint64_t)waves_per_slot * cells) return;
    const uint32_t slot = w / waves_per_slot;
    const uint32_t p0 = (w % waves_per_slot) * chunk;
    const uint32_t p1 = min(n, p0 + chunk);
    const uint64_t* S = spins + (uint64_t)slot * n;
    uint64_t kmin = kNoKey;
    for (uint32_t p = p0; p < p1; ++p) {
        const uint32_t site = uniform(order[p]);
        if (clamped && clamped[site]) continue;
        int32_t h = lin[p];
        for (uint32_t e = rp[p]; e < rp[p + 1]; ++e) h += lane_bit(S[ci[e]]) ? wt[e] : 0;
        const int32_t de = lane_bit(S[p]) ? -h : h;
        const uint64_t k = descent_key(de, site);
        kmin = k < kmin ? k : kmin;
    }
    fold_key(keys + (uint64_t)slot * kWave + lane, kmin);
}

// Dense scan over fields[(slot*n + site)*64 + lane] (site order).
__global__ __launch_bounds__(kBlock) void descent_scan_fields_kernel(
    const int32_t* __restrict__ fields, const uint32_t* __restrict__ pos_of,
    const uint8_t* __restrict__ clamped, const uint64_t* __restrict__ spins, uint32_t n,
    uint32_t cells, uint32_t chunk, uint32_t waves_per_slot,
    unsigned long long* __restrict__ keys, const unsigned long long* __restrict__ ctl) {
    if (ctl[kDescentDone]) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t w = uniform(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    if (w >= (uint64_t)waves_per_slot * cells) return;
    const uint32_t slot = w / waves_per_slot;
    const uint32_t s0 = (w % waves_per_slot) * chunk;
    const uint32_t s1 = min(n, s0 + chunk);
    const uint64_t* S = spins + (uint64_t)slot * n;
    const int32_t* F = fields + (uint64_t)slot * n * kWave;
    uint64_t kmin = kNoKey;
    for (uint32_t s = s0; s < s1; ++s) {
        if (clamped && clamped[s]) continue;
        const int32_t h = F[(uint64_t)s * kWave + lane];
        const int32_t de = lane_bit(S[pos_of[s]]) ? -h : h;
        const uint64_t k = descent_key(de, s);
        kmin = k < kmin ? k : kmin;
    }
    fold_key(keys + (uint64_t)slot * kWave + lane, kmin);
}

// One wave per slot.
__global__ __launch_bounds__(kWave) void descent_apply_kernel(
    const uint32_t* __restrict__ pos_of, uint64_t* __restrict__ spins, uint32_t n,
    int64_t* __restrict__ energy, uint32_t* __restrict__ flips,
    unsigned long long* __restrict__ keys, unsigned long long* __restrict__ ctl) {
    if (ctl[kDescentDone]) return;
    const uint32_t lane = threadIdx.x;
    const uint64_t q = (uint64_t)blockIdx.x * kWave + lane;
    const uint64_t key = keys[q];
    keys[q] = kNoKey;
    const bool hit = key != kNoKey;
    if (hit) {
        const uint32_t site = (uint32_t)key;
        const int64_t de = (int64_t)(key >> 32) - 0x80000000ll;
        atomicXor(reinterpret_cast<unsigned long long*>(spins) + (uint64_t)blockIdx.x * n +
                      pos_of[site],
                  1ull << lane);
        energy[q] += de;
        flips[q] += 1;
    }
    const uint64_t m = wave_ballot(hit);
    if (lane == 0 && m) atomicAdd(ctl + kDescentStepFlips, (unsigned long long)__popcll(m));
}

__global__ void descent_finish_kernel(unsigned long long* ctl) {
    if (ctl[kDescentDone]) return;
    if (ctl[kDescentStepFlips] == 0) {
        ctl[kDescentDone] = 1;
        return;
    }
    ctl[kDescentSteps] += 1;
    ctl[kDescentFlips] += ctl[kDescentStepFlips];
    ctl[kDescentStepFlips] = 0;
}

__global__ void descent_reset_kernel(unsigned long long* keys, uint64_t count,
                                     unsigned long long* ctl) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) keys[i] = kNoKey;
    if (i < kDescentCtlWords) ctl[i] = 0;
}

// Incumbent update, one workgroup: argmin over (energy, temperature, lane) of
// the tracked energies; a strictly lower energy (or the first call) copies that
// replica's bits in site order and records where it was found.
constexpr uint32_t kTrackBlock = 1024;

__global__ __launch_bounds__(kTrackBlock) void track_best_kernel(
    const int64_t* __restrict__ energy, const uint32_t* __restrict__ temp_of,
    const uint64_t* __restrict__ spins, const uint32_t* __restrict__ order, uint32_t n,
    uint32_t cells, uint32_t sweep, BestRecord* __restrict__ best, uint8_t* __restrict__ best_x) {
    __shared__ int64_t se[kTrackBlock];
    __shared__ uint32_t sk[kTrackBlock], sq[kTrackBlock];
    __shared__ uint32_t take;
    const uint32_t tid = threadIdx.x;
    int64_t e = 0;
    uint32_t k = ~0u, qq = ~0u;
    for (uint32_t q = tid; q < cells * kWave; q += kTrackBlock) {
        const int64_t eq = energy[q];
        const uint32_t kq = temp_of[q] * kWave + (q & (kWave - 1));
        if (qq == ~0u || eq < e || (eq == e && kq < k)) {
            e = eq;
            k = kq;
            qq = q;
        }
    }
    se[tid] = e;
    sk[tid] = k;
    sq[tid] = qq;
    __syncthreads();
    for (uint32_t stride = kTrackBlock / 2; stride > 0; stride >>= 1) {
        if (tid < stride) {
            const uint32_t o = tid + stride;
            const bool other = sq[o] != ~0u &&
                               (sq[tid] == ~0u || se[o] < se[tid] || (se[o] == se[tid] && sk[o] < sk[tid]));
            if (other) {
                se[tid] = se[o];
                sk[tid] = sk[o];
                sq[tid] = sq[o];
            }
        }
        __syncthreads();
    }
    if (tid == 0) take = (sq[0] != ~0u) && (!best->valid || se[0] < best->energy);
    __syncthreads();
    if (!take) return;
    const uint32_t q = sq[0];
    const uint32_t slot = q / kWave, lane = q % kWave;
    const uint64_t* S = spins + (uint64_t)slot * n;
    for (uint32_t p = tid; p < n; p += kTrackBlock) best_x[order[p]] = (uint8_t)((S[p] >> lane) & 1u);
    if (tid == 0) {
        best->energy = se[0];
        best->valid = 1;
        best->sweep = sweep;
        best->temp = sk[0] / kWave;
        best->lane = lane;
    }
}

uint32_t ceil_div(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

}  // namespace

hipError_t launch_descent_reset(unsigned long long* keys, uint32_t cells, unsigned long long* ctl,
                                hipStream_t stream) {
    const uint64_t count = (uint64_t)cells * kWave;
    hipLaunchKernelGGL(descent_reset_kernel, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, stream,
                       keys, count, ctl);
    return hipGetLastError();
}

hipError_t launch_descent_step(const DevModel& m, const DevState& s, const int32_t* dense_fields,
                               const uint32_t* pos_of, const uint8_t* clamped,
                               unsigned long long* keys, uint32_t* flips, unsigned long long* ctl,
                               hipStream_t stream) {
    // About 16K waves per scan, at least 16 positions each.
    uint64_t chunk = ((uint64_t)m.n * s.cells + 16383) / 16384;
    if (chunk < 16) chunk = 16;
    if (chunk > 4096) chunk = 4096;
    const uint32_t waves_per_slot = ceil_div(m.n, chunk);
    const dim3 grid(ceil_div((uint64_t)waves_per_slot * s.cells, kWavesPerBlock));
    if (dense_fields)
        hipLaunchKernelGGL(descent_scan_fields_kernel, grid, dim3(kBlock), 0, stream, dense_fields,
                           pos_of, clamped, s.spins, m.n, s.cells, (uint32_t)chunk,
                           waves_per_slot, keys, ctl);
    else
        hipLaunchKernelGGL(descent_scan_kernel, grid, dim3(kBlock), 0, stream, m.rp, m.ci, m.wt,
                           m.lin, m.order, clamped, s.spins, m.n, s.cells, (uint32_t)chunk,
                           waves_per_slot, keys, ctl);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(descent_apply_kernel, dim3(s.cells), dim3(kWave), 0, stream, pos_of,
                       s.spins, m.n, s.energy, flips, keys, ctl);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(descent_finish_kernel, dim3(1), dim3(1), 0, stream, ctl);
    return hipGetLastError();
}

hipError_t launch_track_best(const DevModel& m, const DevState& s, uint32_t sweep,
                             BestRecord* best, uint8_t* best_x, hipStream_t stream) {
    hipLaunchKernelGGL(track_best_kernel, dim3(1), dim3(kTrackBlock), 0, stream, s.energy,
                       s.temp_of, s.spins, m.order, m.n, s.cells, sweep, best, best_x);
    return hipGetLastError();
}

}  // namespace ige
