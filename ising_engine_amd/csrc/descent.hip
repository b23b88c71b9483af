// descent.hip -- deterministic steepest descent over every replica of a state,
// and the best-seen (incumbent) replica.
//
// One descent step: every replica flips the site of its most negative dE
// (lowest site index on a tie; dE = 0 never moves), as SteepestDescent::apply
// does per sweep (src/engine_v2/operators/steepest_descent.rs:80-124); run to
// the end it is steepest_descent_1opt of src/solver/local_search.rs:19-60 for
// each replica.  A step is three launches on the state's stream:
//   scan    lane = replica.  dE of each (slot, site, lane) -- from the CSR rows
//           gathered eight positions at a time, or from the matrix-core
//           fields of a dense model -- folded into a per-lane minimum of the
//           key (dE + 2^31) << 32 | site, then one atomicMin per lane and
//           workgroup into keys[slot][lane];
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

constexpr uint32_t kScanBlock = 512;  // 8 waves, one slot per workgroup
constexpr uint32_t kScanWaves = kScanBlock / kWave;
constexpr int kScanGroup = 8;          // positions per gather group
constexpr int kScanInterleave = 1;     // groups gathered together

// Folds the per-lane minima of the workgroup's waves, then one atomicMin per lane.
__device__ __forceinline__ void fold_block(unsigned long long* keys, uint64_t kmin, uint32_t wave,
                                           uint32_t lane) {
    __shared__ uint64_t red[kScanWaves][kWave];
    red[wave][lane] = kmin;
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (uint32_t w = 1; w < kScanWaves; ++w) kmin = red[w][lane] < kmin ? red[w][lane] : kmin;
        fold_key(keys + lane, kmin);
    }
}

// Sparse scan.  Workgroup b covers slot b / blocks_per_slot; its wave w the
// positions [p0, p0 + chunk) with p0 = ((b % blocks_per_slot) * kScanWaves + w) * chunk.
// site_aux[p] = site of p, bit 31 set if the site is clamped.
__global__ __launch_bounds__(kScanBlock) void descent_scan_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin,
    const uint32_t* __restrict__ site_aux, const uint64_t* __restrict__ spins, uint32_t n,
    uint32_t chunk, uint32_t blocks_per_slot, unsigned long long* __restrict__ keys,
    const unsigned long long* __restrict__ ctl) {
    if (ctl[kDescentDone]) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = uniform(threadIdx.x >> 6);
    const uint32_t slot = blockIdx.x / blocks_per_slot;
    const uint64_t p0_64 = ((uint64_t)(blockIdx.x % blocks_per_slot) * kScanWaves + wave) * chunk;
    const uint32_t p0 = (uint32_t)min<uint64_t>(p0_64, n);
    const uint32_t p1 = min(n, p0 + chunk);
    const uint64_t* S = spins + (uint64_t)slot * n;
    uint64_t kmin = kNoKey;
    fields_groups<kScanGroup, kScanInterleave>(
        rp, ci, wt, lin, site_aux, S, p0, p1, lane,
        [&](uint32_t, uint64_t x, int32_t h, uint32_t aux) {
            const int32_t de = lane_bit(x) ? -h : h;
            const uint64_t k = (aux >> 31) ? kNoKey : descent_key(de, aux);
            kmin = k < kmin ? k : kmin;
        });
    fold_block(keys + (uint64_t)slot * kWave, kmin, wave, lane);
}

// Dense scan over fields[(slot*n + site)*64 + lane] (site order); workgroups
// as descent_scan_kernel, with sites in place of positions.  Each group of up
// to 64 sites gathers its words (through pos_of) and clamps one per lane,
// then reads the fields eight sites at a time.
__global__ __launch_bounds__(kScanBlock) void descent_scan_fields_kernel(
    const int32_t* __restrict__ fields, const uint32_t* __restrict__ pos_of,
    const uint8_t* __restrict__ clamped, const uint64_t* __restrict__ spins, uint32_t n,
    uint32_t chunk, uint32_t blocks_per_slot, unsigned long long* __restrict__ keys,
    const unsigned long long* __restrict__ ctl) {
    if (ctl[kDescentDone]) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = uniform(threadIdx.x >> 6);
    const uint32_t slot = blockIdx.x / blocks_per_slot;
    const uint64_t s0_64 = ((uint64_t)(blockIdx.x % blocks_per_slot) * kScanWaves + wave) * chunk;
    const uint32_t s0 = (uint32_t)min<uint64_t>(s0_64, n);
    const uint32_t s1 = min(n, s0 + chunk);
    const uint64_t* S = spins + (uint64_t)slot * n;
    const int32_t* F = fields + (uint64_t)slot * n * kWave;
    uint64_t kmin = kNoKey;
    for (uint32_t g = s0; g < s1; g += kWave) {
        const uint32_t cnt = min(kWave, s1 - g);
        uint64_t xw = 0;
        uint32_t cl = 0;
        if (lane < cnt) {
            xw = S[pos_of[g + lane]];
            cl = clamped ? clamped[g + lane] : 0u;
        }
        for (uint32_t q0 = 0; q0 < cnt; q0 += 8) {
            int32_t hv[8];
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                hv[j] = q0 + j < cnt ? F[(uint64_t)(g + q0 + j) * kWave + lane] : 0;
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j) {
                const uint32_t q = q0 + j;
                if (q >= cnt || __builtin_amdgcn_readlane(cl, q)) continue;
                const int32_t de = lane_bit(readlane64(xw, q)) ? -hv[j] : hv[j];
                const uint64_t k = descent_key(de, g + q);
                kmin = k < kmin ? k : kmin;
            }
        }
    }
    fold_block(keys + (uint64_t)slot * kWave, kmin, wave, lane);
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

// site_aux[p] = order[p] | (clamped[order[p]] != 0) << 31: the sparse scan's
// position -> site map with the clamp flag.
__global__ void descent_aux_kernel(const uint32_t* __restrict__ order,
                                   const uint8_t* __restrict__ clamped, uint32_t n,
                                   uint32_t* __restrict__ site_aux) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) site_aux[p] = order[p] | (clamped[order[p]] ? 0x80000000u : 0u);
}

__global__ void descent_reset_kernel(unsigned long long* keys, uint64_t count,
                                     unsigned long long* ctl) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) keys[i] = kNoKey;
    if (i < kDescentCtlWords) ctl[i] = 0;
}

// Incumbent update, one workgroup: argmin over (energy, temperature, lane) of
// the tracked energies; a strictly lower energy (or the first call) updates
// the record and sets take[0] = 1, take[1] = the winning slot * 64 + lane, for
// track_copy_kernel to copy that replica's bits.
constexpr uint32_t kTrackBlock = 1024;

__global__ __launch_bounds__(kTrackBlock) void track_best_kernel(
    const int64_t* __restrict__ energy, const uint32_t* __restrict__ temp_of, uint32_t cells,
    uint32_t replica_view, uint32_t sweep_rel, BestRecord* __restrict__ best,
    uint32_t* __restrict__ take, const uint32_t* __restrict__ clock) {
    const uint32_t sweep = clocked(clock, kClockSweep, sweep_rel);
    __shared__ int64_t se[kTrackBlock];
    __shared__ uint32_t sk[kTrackBlock], sq[kTrackBlock];
    const uint32_t tid = threadIdx.x;
    int64_t e = 0;
    uint32_t k = ~0u, qq = ~0u;
    for (uint32_t q = tid; q < cells * kWave; q += kTrackBlock) {
        const int64_t eq = energy[q];
        // the replica index: (temperature, lane), or temp_of itself in the replica view
        const uint32_t kq = replica_view ? temp_of[q] : temp_of[q] * kWave + (q & (kWave - 1));
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
    if (tid != 0) return;
    const bool t = (sq[0] != ~0u) && (!best->valid || se[0] < best->energy);
    take[0] = t ? 1u : 0u;
    take[1] = sq[0];
    if (t) {
        best->energy = se[0];
        best->valid = 1;
        best->sweep = sweep;
        best->temp = replica_view ? sk[0] : sk[0] / kWave;
        best->lane = sq[0] % kWave;
    }
}

// Copies the replica take[1] names into best_x (site order) when take[0] is set.
__global__ __launch_bounds__(kBlock) void track_copy_kernel(
    const uint64_t* __restrict__ spins, const uint32_t* __restrict__ order, uint32_t n,
    const uint32_t* __restrict__ take, uint8_t* __restrict__ best_x) {
    if (!take[0]) return;
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint32_t q = take[1];
    const uint32_t slot = q / kWave, lane = q % kWave;
    best_x[order[p]] = (uint8_t)((spins[(uint64_t)slot * n + p] >> lane) & 1u);
}

uint32_t ceil_div(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

}  // namespace

hipError_t launch_descent_aux(const uint32_t* order, const uint8_t* clamped, uint32_t n,
                              uint32_t* site_aux, hipStream_t stream) {
    hipLaunchKernelGGL(descent_aux_kernel, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, stream,
                       order, clamped, n, site_aux);
    return hipGetLastError();
}

hipError_t launch_descent_reset(unsigned long long* keys, uint32_t cells, unsigned long long* ctl,
                                hipStream_t stream) {
    const uint64_t count = (uint64_t)cells * kWave;
    hipLaunchKernelGGL(descent_reset_kernel, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, stream,
                       keys, count, ctl);
    return hipGetLastError();
}

hipError_t launch_descent_step(const DevModel& m, const DevState& s, const int32_t* dense_fields,
                               const uint32_t* pos_of, const uint8_t* clamped,
                               const uint32_t* site_aux, unsigned long long* keys, uint32_t* flips,
                               unsigned long long* ctl, hipStream_t stream) {
    // About 32K waves per scan; positions per wave a multiple of one
    // interleaved gather (sparse) or of eight sites (dense).
    uint64_t chunk = ((uint64_t)m.n * s.cells + 32767) / 32768;
    const uint64_t unit = dense_fields ? 8 : kScanGroup * kScanInterleave;
    chunk = (chunk + unit - 1) / unit * unit;
    if (chunk < unit) chunk = unit;
    if (chunk > 4096) chunk = 4096;
    const uint32_t blocks_per_slot = ceil_div(m.n, chunk * kScanWaves);
    const dim3 grid((uint32_t)((uint64_t)blocks_per_slot * s.cells));
    if (dense_fields)
        hipLaunchKernelGGL(descent_scan_fields_kernel, grid, dim3(kScanBlock), 0, stream,
                           dense_fields, pos_of, clamped, s.spins, m.n, (uint32_t)chunk,
                           blocks_per_slot, keys, ctl);
    else
        hipLaunchKernelGGL(descent_scan_kernel, grid, dim3(kScanBlock), 0, stream, m.rp, m.ci,
                           m.wt, m.lin, site_aux, s.spins, m.n, (uint32_t)chunk, blocks_per_slot,
                           keys, ctl);
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
                             BestRecord* best, uint32_t* take, uint8_t* best_x,
                             hipStream_t stream) {
    hipLaunchKernelGGL(track_best_kernel, dim3(1), dim3(kTrackBlock), 0, stream, s.energy,
                       s.temp_of, s.cells, s.replica_view, sweep, best, take, s.clock);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(track_copy_kernel, dim3(ceil_div(m.n, kBlock)), dim3(kBlock), 0, stream,
                       s.spins, m.order, m.n, take, best_x);
    return hipGetLastError();
}

}  // namespace ige
