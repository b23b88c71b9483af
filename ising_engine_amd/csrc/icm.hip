// icm.hip -- isoenergetic cluster moves (Houdayer / ICM, src/solver/icm.rs:91-185)
// on the MSC words.
//
// One workgroup per move: temperature t, r < 32, replicas (t, r) and (t, r + 32)
// found through slot_of (a configuration stays in its slot; exchange moves its
// temperature).  The workgroup
//   1. counts the sites where the two replicas disagree;
//   2. picks the floor(u |D| / 2^32)-th disagreeing site in site order as the
//      seed (u = Philox(t, r, round, tag 6)[0]), with a block-wide scan over
//      sites;
//   3. grows the seed's connected component of disagreeing sites by a
//      level-synchronous BFS: a visited bitmap and a queue of positions in the
//      move's scratch, claimed with atomicOr, appended with an LDS counter;
//   4. sums dE of flipping the cluster in replica A (int64, exact);
//   5. flips the cluster in both replicas (64-bit atomic XOR: other moves own
//      other bits of the same words) and moves dE from B to A.
// Every position enters the queue at most once, so the BFS ends after at most
// n levels; all threads follow the same LDS head / tail, so the loop is
// uniform across the workgroup.
#include <hip/hip_runtime.h>

#include "ige_math.h"
#include "kernels.hpp"
#include "sweep_generic.h"

namespace ige {

namespace {

constexpr uint32_t kIcmBlock = 512;
constexpr uint32_t kIcmWaves = kIcmBlock / 64;

// The bitmap is claimed with atomics and the queue filled by other waves, so
// both are read with agent-scope loads (from L2, never a stale L1 line).
__device__ __forceinline__ uint32_t load_l2(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_l2(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t disagree(const uint64_t* A, const uint64_t* B, uint32_t p,
                                             uint32_t r) {
    return (uint32_t)(((A[p] >> r) ^ (B[p] >> (r + 32))) & 1u);
}

// Block-wide exclusive scan of v (one value per thread); returns the exclusive
// prefix and writes the block total to *total.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if ((int)lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t w = 0; w < kIcmWaves; ++w) {
        if (w < wave) base += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return base + x - v;
}

__global__ __launch_bounds__(kIcmBlock) void icm_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin,
    const uint32_t* __restrict__ pos_of, uint64_t* spins, uint32_t n,
    const uint32_t* __restrict__ slot_of, int64_t* energy, uint32_t pair0, uint32_t round,
    uint32_t key0, uint32_t key1, uint32_t* scratch, uint64_t scratch_stride,
    unsigned long long* flipped) {
    __shared__ uint32_t wsum[kIcmWaves];
    __shared__ uint32_t s_seed, s_head, s_tail, s_found;
    __shared__ int64_t s_de[kIcmWaves];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t pair = pair0 + blockIdx.x;
    const uint32_t t = pair / 32, r = pair % 32;
    const uint32_t sa = slot_of[t * 64 + r], sb = slot_of[t * 64 + r + 32];
    uint64_t* A = spins + (uint64_t)sa * n;
    uint64_t* B = spins + (uint64_t)sb * n;
    uint32_t* visited = scratch + (uint64_t)blockIdx.x * scratch_stride;
    uint32_t* queue = visited + (n + 31) / 32;

    // 1. |D| and a cleared bitmap
    uint32_t cnt = 0;
    for (uint32_t p = tid; p < n; p += kIcmBlock) cnt += disagree(A, B, p, r);
    for (uint32_t k = tid; k < (n + 31) / 32; k += kIcmBlock) store_l2(visited + k, 0u);
    uint32_t count = 0;
    (void)block_scan(cnt, wsum, &count);
    if (count == 0) return;  // uniform: every thread saw the same total

    // 2. the seed: the k-th disagreeing site in site order
    const U32x4 u = philox4x32_10(t, r, round, kTagIcm << 24, key0, key1);
    const uint32_t k = (uint32_t)(((uint64_t)u.v[0] * count) >> 32);
    if (tid == 0) s_found = 0;
    uint32_t before = 0;
    for (uint32_t s0 = 0; s0 < n; s0 += kIcmBlock) {
        const uint32_t s = s0 + tid;
        const uint32_t p = s < n ? pos_of[s] : 0u;
        const uint32_t d = s < n ? disagree(A, B, p, r) : 0u;
        uint32_t tot = 0;
        const uint32_t ex = block_scan(d, wsum, &tot);
        if (d && before + ex == k) {
            s_seed = p;
            s_found = 1;
        }
        __syncthreads();
        if (s_found) break;
        before += tot;
    }

    // 3. BFS over disagreeing neighbours
    if (tid == 0) {
        store_l2(queue, s_seed);
        atomicOr(visited + (s_seed >> 5), 1u << (s_seed & 31));
        s_head = 0;
        s_tail = 1;
    }
    __syncthreads();
    for (;;) {
        // Every thread reads this level's bounds before any wave appends to
        // the queue: without the barrier a late wave could see a tail that
        // already counts this level's appends and load slots not yet stored.
        const uint32_t head = s_head, end = s_tail;
        __syncthreads();
        if (head == end) break;
        for (uint32_t q = head + tid; q < end; q += kIcmBlock) {
            const uint32_t i = load_l2(queue + q);
            for (uint32_t e = rp[i]; e < rp[i + 1]; ++e) {
                const uint32_t j = ci[e];
                if (!disagree(A, B, j, r)) continue;
                const uint32_t m = 1u << (j & 31);
                if (atomicOr(visited + (j >> 5), m) & m) continue;
                store_l2(queue + atomicAdd(&s_tail, 1u), j);
            }
        }
        __syncthreads();
        if (tid == 0) s_head = end;
        __syncthreads();
    }
    const uint32_t size = s_tail;

    // 4. dE of the flip in replica A (icm.rs:55-82), internal edges once (i < j)
    int64_t de = 0;
    for (uint32_t q = tid; q < size; q += kIcmBlock) {
        const uint32_t i = load_l2(queue + q);
        const int64_t xi = (int64_t)((A[i] >> r) & 1u);
        de += (int64_t)lin[i] * (1 - 2 * xi);
        for (uint32_t e = rp[i]; e < rp[i + 1]; ++e) {
            const uint32_t j = ci[e];
            const int64_t xj = (int64_t)((A[j] >> r) & 1u);
            const bool in_c = (load_l2(visited + (j >> 5)) >> (j & 31)) & 1u;
            if (in_c) {
                if (i < j) de += (int64_t)wt[e] * ((1 - xi) * (1 - xj) - xi * xj);
            } else {
                de += (int64_t)wt[e] * xj * (1 - 2 * xi);
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)de, off), hi = __shfl_xor((uint32_t)(de >> 32), off);
        de += (int64_t)(((uint64_t)hi << 32) | lo);
    }
    if (lane == 0) s_de[wave] = de;
    __syncthreads();

    // 5. flip in both replicas, move the energy
    const unsigned long long ma = 1ull << r, mb = 1ull << (r + 32);
    for (uint32_t q = tid; q < size; q += kIcmBlock) {
        const uint32_t i = load_l2(queue + q);
        if (sa == sb) {
            atomicXor(reinterpret_cast<unsigned long long*>(A + i), ma | mb);
        } else {
            atomicXor(reinterpret_cast<unsigned long long*>(A + i), ma);
            atomicXor(reinterpret_cast<unsigned long long*>(B + i), mb);
        }
    }
    if (tid == 0) {
        int64_t tot = 0;
        for (uint32_t w = 0; w < kIcmWaves; ++w) tot += s_de[w];
        energy[sa * 64 + r] += tot;
        energy[sb * 64 + r + 32] -= tot;
        if (flipped) atomicAdd(flipped, (unsigned long long)size);
    }
}

}  // namespace

uint64_t icm_scratch_words(uint32_t n) { return (uint64_t)(n + 31) / 32 + n; }

hipError_t launch_icm(const DevModel& m, const DevState& s, const uint32_t* pos_of,
                      uint32_t pair0, uint32_t pairs, uint32_t round, uint64_t seed,
                      uint32_t* scratch, unsigned long long* flipped, hipStream_t stream) {
    if (pairs == 0) return hipSuccess;
    hipLaunchKernelGGL(icm_kernel, dim3(pairs), dim3(kIcmBlock), 0, stream, m.rp, m.ci, m.wt,
                       m.lin, pos_of, s.spins, m.n, s.slot_of, s.energy, pair0, round,
                       (uint32_t)seed, (uint32_t)(seed >> 32), scratch, icm_scratch_words(m.n),
                       flipped);
    return hipGetLastError();
}

}  // namespace ige
