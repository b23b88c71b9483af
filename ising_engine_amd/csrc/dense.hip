// dense.hip -- Metropolis sweep for dense couplings (fully connected SK,
// dense QUBO) with the local fields recomputed on int8 matrix cores.
//
// Model: q_ij = scale * a_ij with a_ij int8, stored as an n_pad x n_pad int8
// matrix in position space.  One workgroup (4 waves) per slot walks the
// positions in blocks of 64:
//   1. every wave draws the block's uniforms (word view, the same Philox
//      words as the lane-per-replica kernel) into LDS, and the block's own
//      couplings scale * a[p0+p][p0+k] are staged in LDS;
//   2. the fields of the block, h[i][r] = lin + scale * sum_k a[p0+i][k] x_k(r),
//      are recomputed from the current spins as a 64 x n_pad x 64 int8 GEMM
//      with v_mfma_i32_32x32x32_i8 (exact: |sum| <= 127 n < 2^31).  The
//      B operand (the spins, 0/1 bytes) is expanded from replica-major bit
//      rows kept in LDS;
//   3. wave 0 (lane = replica) runs the block's positions in order: accept if
//      dE <= 0 or u < metropolis_threshold(beta, dE), then pushes the flip
//      into the fields of the block's later positions (rank-1 update with the
//      staged couplings), exactly the sequential order of the sweep;
//   4. the block's new words go back to HBM and into the LDS bit rows.
// Positions are visited one by one in colour-major order, which gives the
// same result as the colour-parallel kernels: positions of one colour class
// do not interact.
//
// The acceptance test uses a single-precision estimate of the threshold
// (v_exp_f32) with a guard band wide enough to contain the deterministic
// threshold; only proposals inside the band evaluate it exactly, so the
// decisions are those of metropolis_threshold bit for bit.
#include <hip/hip_runtime.h>

#include "ige_math.h"
#include "kernels.hpp"
#include "sweep_generic.h"

namespace ige {

namespace {

typedef int32_t v4i __attribute__((ext_vector_type(4)));
typedef int32_t v16i __attribute__((ext_vector_type(16)));

constexpr uint32_t kDenseThreads = 256;  // 4 waves
constexpr uint32_t kHStride = 65;        // padded row of the block-field tile

struct DenseLds {
    uint64_t* xt;      // [64][xs] replica-major spin bits, xs = n_pad / 64 + 1
    int32_t* hblk;     // [64][kHStride] fields of the block, [position][replica]
    uint32_t* ublk;    // [64][64] uniforms, [position][lane]
    int32_t* wblk;     // [64][64] scale * a[p0 + p][p0 + k], [p][k]
};

__host__ __device__ inline size_t dense_lds_bytes(uint32_t n_pad) {
    const size_t xs = n_pad / 64 + 1;
    return 64 * xs * sizeof(uint64_t) + 64 * kHStride * sizeof(int32_t) +
           2 * 64 * 64 * sizeof(uint32_t);
}

__device__ __forceinline__ DenseLds carve(unsigned char* base, uint32_t xs) {
    DenseLds l;
    l.xt = reinterpret_cast<uint64_t*>(base);
    l.hblk = reinterpret_cast<int32_t*>(l.xt + 64 * xs);
    l.ublk = reinterpret_cast<uint32_t*>(l.hblk + 64 * kHStride);
    l.wblk = reinterpret_cast<int32_t*>(l.ublk + 64 * 64);
    return l;
}

// 16 bits -> 16 bytes, byte i = bit i (0 or 1).  n * 0x00204081 places the
// four bits of a nibble n at bits 0, 8, 16, 24 with no overlapping terms.
__device__ __forceinline__ v4i spread16(uint32_t b) {
    v4i r;
    r[0] = (int32_t)(((b & 0xFu) * 0x00204081u) & 0x01010101u);
    r[1] = (int32_t)((((b >> 4) & 0xFu) * 0x00204081u) & 0x01010101u);
    r[2] = (int32_t)((((b >> 8) & 0xFu) * 0x00204081u) & 0x01010101u);
    r[3] = (int32_t)((((b >> 12) & 0xFu) * 0x00204081u) & 0x01010101u);
    return r;
}

// Replica-major bit rows of one slot: xt[r * xs + c] bit k = bit r of word
// 64c + k (zero past n).  Every wave transposes chunks wave, wave + 4, ...
__device__ __forceinline__ void load_bit_rows(const uint64_t* words, uint32_t n, uint32_t nchunk,
                                              uint32_t xs, uint64_t* xt, uint32_t wave,
                                              uint32_t lane) {
    for (uint32_t c = wave; c < nchunk; c += 4) {
        const uint32_t p = c * 64 + lane;
        const uint64_t w = p < n ? words[p] : 0ull;
        xt[lane * xs + c] = wave_transpose(w, lane);
    }
}

// Fields of positions [p0, p0 + 64) for all 64 replicas into hblk, by the
// four waves: wave w owns rows 32 (w >> 1) .. +31 and replicas 32 (w & 1) .. +31.
// Operand bytes of lane l: row / replica l % 32, k = 16 (l / 32) + b within
// each 32-wide k step; A and B use the same k assignment, so the product
// does not depend on how the instruction orders k internally.  Result
// element v of lane l is row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32.
__device__ __forceinline__ void block_fields(const DevDense& d, const uint64_t* xt, uint32_t xs,
                                             uint32_t p0, int32_t* hblk, uint32_t wave,
                                             uint32_t lane) {
    const uint32_t nchunk = d.n_pad / 64;
    const uint32_t rbase = 32 * (wave >> 1), jbase = 32 * (wave & 1);
    const uint32_t kh = 16 * (lane >> 5);
    const int8_t* arow = d.a + (size_t)(p0 + rbase + (lane & 31)) * d.n_pad + kh;
    const uint64_t* xr = xt + (size_t)(jbase + (lane & 31)) * xs;
    v16i acc = {};
    for (uint32_t kc = 0; kc < nchunk; ++kc) {
        const uint64_t xb = xr[kc];
        const v4i a0 = *reinterpret_cast<const v4i*>(arow + (size_t)kc * 64);
        const v4i a1 = *reinterpret_cast<const v4i*>(arow + (size_t)kc * 64 + 32);
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, spread16((uint32_t)(xb >> kh)), acc, 0,
                                                    0, 0);
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(
            a1, spread16((uint32_t)(xb >> (32 + kh))), acc, 0, 0, 0);
    }
    const uint32_t j = jbase + (lane & 31);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const uint32_t i = rbase + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        hblk[i * kHStride + j] = d.lin[p0 + i] + d.scale * acc[v];
    }
}

// u < metropolis_threshold(beta, de) for 1 <= de <= dlim.  a estimates
// T = 2^32 exp(-beta de) within a relative 1e-5 (single-precision product
// and v_exp_f32); the deterministic threshold lies within T (1 +- 1e-14) + 1.
// Outside the band [a (1 - 1e-4) - 4, a (1 + 1e-4) + 4] the answer is
// therefore decided by a; inside it the threshold is evaluated.
__device__ __attribute__((noinline)) bool exact_accept(uint32_t u, int32_t de, double beta) {
    return u < metropolis_threshold(beta, (int64_t)de);
}

__device__ __forceinline__ bool dense_accept(uint32_t u, int32_t de, float nb, double beta) {
    const float a = __builtin_amdgcn_exp2f((float)de * nb) * 4294967296.0f;
    const float lo = a * 0.9999f - 4.0f;
    const float hi = a * 1.0001f + 4.0f;
    if (lo > 0.0f && u < (uint32_t)lo) return true;
    if (hi < 4294967040.0f && u > (uint32_t)hi) return false;
    return exact_accept(u, de, beta);
}

__global__ __launch_bounds__(kDenseThreads) void sweep_dense_kernel(DevDense d, DevState s,
                                                                    const double* __restrict__ beta,
                                                                    uint32_t sweep, uint32_t key0,
                                                                    uint32_t key1) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t nchunk = d.n_pad / 64, xs = nchunk + 1;
    const DenseLds L = carve(lds_raw, xs);
    const uint32_t slot = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t n = d.n;
    uint64_t* words = s.spins + (size_t)slot * n;
    load_bit_rows(words, n, nchunk, xs, L.xt, wave, lane);
    const double b = beta[s.temp_of[slot * 64 + lane]];
    const uint32_t dlim = threshold_limit(b, d.dtab);
    const float nb = (float)(-b * 1.4426950408889634);
    int64_t eacc = 0;
    uint32_t nacc = 0;
    __syncthreads();
    for (uint32_t blk = 0; blk < nchunk; ++blk) {
        const uint32_t p0 = blk * 64;
        const uint32_t cnt = min(64u, n - p0);
        // 1. uniforms of positions p0 + 16 wave .. +15, and the block couplings
#pragma unroll
        for (uint32_t g = 0; g < 4; ++g) {
            const uint32_t q = 16 * wave + 4 * g;
            const U32x4 u =
                philox4x32_10((p0 + q) >> 2, slot, sweep, (kTagWord << 24) | lane, key0, key1);
#pragma unroll
            for (uint32_t t = 0; t < 4; ++t) L.ublk[(q + t) * 64 + lane] = u.v[t];
        }
        {
            const uint32_t k = threadIdx.x >> 2, c0 = (threadIdx.x & 3u) * 16;
            const v4i a = *reinterpret_cast<const v4i*>(d.a + (size_t)(p0 + k) * d.n_pad + p0 + c0);
#pragma unroll
            for (uint32_t t = 0; t < 16; ++t) {
                const int32_t av = (int32_t)(int8_t)((uint32_t)a[t >> 2] >> (8 * (t & 3)));
                L.wblk[k * 64 + c0 + t] = d.scale * av;
            }
        }
        // 2. fields of the block on the matrix cores
        block_fields(d, L.xt, xs, p0, L.hblk, wave, lane);
        __syncthreads();
        // 3. the block's positions in order, lane = replica
        if (wave == 0) {
            int32_t hb[64];
#pragma unroll
            for (int k = 0; k < 64; ++k) hb[k] = L.hblk[k * kHStride + lane];
            uint64_t wv = lane < cnt ? words[p0 + lane] : 0ull;
#pragma unroll
            for (int p = 0; p < 64; ++p) {
                if ((uint32_t)p >= cnt) continue;  // past n in the last block
                const uint64_t x = readlane64(wv, p);
                const bool xb = lane_bit(x);
                const int32_t de = xb ? -hb[p] : hb[p];
                bool acc = de <= 0;
                if (!acc && (uint32_t)de <= dlim)
                    acc = dense_accept(L.ublk[p * 64 + lane], de, nb, b);
                const uint64_t m = wave_ballot(acc);
                if (m) {
                    const int32_t dx = acc ? (xb ? -1 : 1) : 0;
#pragma unroll
                    for (int k = p + 1; k < 64; ++k) hb[k] += dx * L.wblk[p * 64 + k];
                    eacc += acc ? de : 0;
                    nacc += (uint32_t)__builtin_popcountll(m);
                    wv = lane == (uint32_t)p ? (x ^ m) : wv;
                }
            }
            // 4. new words to HBM and to the bit rows
            if (lane < cnt) words[p0 + lane] = wv;
            L.xt[lane * xs + blk] = wave_transpose(wv, lane);
        }
        __syncthreads();
    }
    if (wave == 0) {
        s.energy[slot * 64 + lane] += eacc;
        if (lane == 0 && nacc) atomicAdd(s.accepted, (unsigned long long)nacc);
    }
}

// Local fields of every position on the matrix cores: one workgroup per
// (slot, block of 64 positions); out[(slot * n + site) * 64 + lane].
__global__ __launch_bounds__(kDenseThreads) void dense_fields_kernel(DevDense d, DevState s,
                                                                     int32_t* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t nchunk = d.n_pad / 64, xs = nchunk + 1;
    const DenseLds L = carve(lds_raw, xs);
    const uint32_t slot = blockIdx.x, blk = blockIdx.y;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t n = d.n, p0 = blk * 64;
    load_bit_rows(s.spins + (size_t)slot * n, n, nchunk, xs, L.xt, wave, lane);
    __syncthreads();
    block_fields(d, L.xt, xs, p0, L.hblk, wave, lane);
    __syncthreads();
    const uint32_t cnt = min(64u, n - p0);
    for (uint32_t i = wave; i < cnt; i += 4)
        out[((size_t)slot * n + d.order[p0 + i]) * 64 + lane] = L.hblk[i * kHStride + lane];
}

bool set_lds_limit(const void* fn, size_t bytes) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
           hipSuccess;
}

}  // namespace

size_t dense_lds_bytes_host(uint32_t n_pad) { return dense_lds_bytes(n_pad); }

hipError_t launch_sweep_dense(const DevDense& d, const DevState& s, uint32_t sweep, uint64_t seed,
                              const double* beta, hipStream_t stream) {
    const size_t lds = dense_lds_bytes(d.n_pad);
    if (!set_lds_limit(reinterpret_cast<const void*>(&sweep_dense_kernel), lds))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(sweep_dense_kernel, dim3(s.cells), dim3(kDenseThreads), lds, stream, d, s,
                       beta, sweep, (uint32_t)seed, (uint32_t)(seed >> 32));
    return hipGetLastError();
}

hipError_t launch_dense_fields(const DevDense& d, const DevState& s, int32_t* out,
                               hipStream_t stream) {
    const size_t lds = dense_lds_bytes(d.n_pad);
    if (!set_lds_limit(reinterpret_cast<const void*>(&dense_fields_kernel), lds))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(dense_fields_kernel, dim3(s.cells, d.n_pad / 64), dim3(kDenseThreads), lds,
                       stream, d, s, out);
    return hipGetLastError();
}

}  // namespace ige
