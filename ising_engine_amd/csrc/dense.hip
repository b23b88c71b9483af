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

constexpr int kPrefetch = 8;             // k chunks of the coupling rows in flight

// LDS: the replica-major spin bits xt[64][xs] (xs = n_pad / 64 + 1) are
// dynamic; the per-block tiles are static so that the unrolled sequential
// loop addresses them with immediate offsets.
struct BlockTiles {
    int32_t h[2][64 * kHStride];  // fields of this block and the next, [position][replica]
    uint32_t u[64 * 64];          // uniforms, [position][lane]
    int4 w[64 * 16];              // scale * a[p0 + p][p0 + k], row p = 16 int4
    uint64_t xnew[64];            // the block's new words as replica-major bit rows
};

__host__ __device__ inline size_t dense_lds_bytes(uint32_t n_pad) {
    return 64 * (size_t)(n_pad / 64 + 1) * sizeof(uint64_t);
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
// The coupling matrix is stored in exactly this operand order: tile (g, c)
// holds rows 32g.. and columns 32c.. as 64 lanes x 16 bytes, so every load
// of a wave reads 1 KB of contiguous memory (tiles of one row group follow
// each other along k).
__device__ __forceinline__ void block_fields(const DevDense& d, const uint64_t* xt, uint32_t xs,
                                             uint32_t p0, int32_t* hblk, uint32_t wave,
                                             uint32_t lane) {
    const uint32_t nchunk = d.n_pad / 64;
    const uint32_t rbase = 32 * (wave >> 1), jbase = 32 * (wave & 1);
    const uint32_t kh = 16 * (lane >> 5);
    const uint32_t ntile = d.n_pad / 32;  // tiles per row group
    const int8_t* arow = d.a + ((size_t)((p0 + rbase) / 32) * ntile) * 1024 + 16 * lane;
    const uint64_t* xr = xt + (size_t)(jbase + (lane & 31)) * xs;
    v16i acc = {};
    // Coupling rows stream from L2 / HBM with kPrefetch chunks in flight.
    // n_pad is a multiple of 64 * kPrefetch, so the loop needs no guards; the
    // last kPrefetch iterations re-load the final chunk instead of running
    // past the row.
    v4i pa0[kPrefetch], pa1[kPrefetch];
#pragma unroll
    for (int i = 0; i < kPrefetch; ++i) {
        pa0[i] = *reinterpret_cast<const v4i*>(arow + (size_t)i * 2048);
        pa1[i] = *reinterpret_cast<const v4i*>(arow + (size_t)i * 2048 + 1024);
    }
    for (uint32_t kb = 0; kb < nchunk; kb += kPrefetch) {
#pragma unroll
        for (int i = 0; i < kPrefetch; ++i) {
            const uint32_t kc = kb + i;
            const v4i a0 = pa0[i], a1 = pa1[i];
            const uint32_t kn = min(kc + (uint32_t)kPrefetch, nchunk - 1);
            pa0[i] = *reinterpret_cast<const v4i*>(arow + (size_t)kn * 2048);
            pa1[i] = *reinterpret_cast<const v4i*>(arow + (size_t)kn * 2048 + 1024);
            const uint64_t xb = xr[kc];
            acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, spread16((uint32_t)(xb >> kh)), acc,
                                                        0, 0, 0);
            acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                a1, spread16((uint32_t)(xb >> (32 + kh))), acc, 0, 0, 0);
        }
    }
    const uint32_t j = jbase + (lane & 31);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const uint32_t i = rbase + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        hblk[i * kHStride + j] = d.lin[p0 + i] + d.scale * acc[v];
    }
}

// Rows 32 rh .. +31 of the fields of positions [p0, p0 + 64), all 64 replicas
// (two 32 x 32 tiles sharing each coupling load), into hblk; one wave.
__device__ __forceinline__ void half_block_fields(const DevDense& d, const uint64_t* xt,
                                                  uint32_t xs, uint32_t p0, int32_t* hblk,
                                                  uint32_t rh, uint32_t lane) {
    const uint32_t nchunk = d.n_pad / 64;
    const uint32_t kh = 16 * (lane >> 5);
    const uint32_t ntile = d.n_pad / 32;
    const int8_t* arow = d.a + ((size_t)((p0 + 32 * rh) / 32) * ntile) * 1024 + 16 * lane;
    const uint64_t* xr0 = xt + (size_t)(lane & 31) * xs;
    const uint64_t* xr1 = xt + (size_t)(32 + (lane & 31)) * xs;
    v16i acc0 = {}, acc1 = {};
    v4i pa0[kPrefetch], pa1[kPrefetch];
#pragma unroll
    for (int i = 0; i < kPrefetch; ++i) {
        pa0[i] = *reinterpret_cast<const v4i*>(arow + (size_t)i * 2048);
        pa1[i] = *reinterpret_cast<const v4i*>(arow + (size_t)i * 2048 + 1024);
    }
    for (uint32_t kb = 0; kb < nchunk; kb += kPrefetch) {
#pragma unroll
        for (int i = 0; i < kPrefetch; ++i) {
            const uint32_t kc = kb + i;
            const v4i a0 = pa0[i], a1 = pa1[i];
            const uint32_t kn = min(kc + (uint32_t)kPrefetch, nchunk - 1);
            pa0[i] = *reinterpret_cast<const v4i*>(arow + (size_t)kn * 2048);
            pa1[i] = *reinterpret_cast<const v4i*>(arow + (size_t)kn * 2048 + 1024);
            const uint64_t x0 = xr0[kc], x1 = xr1[kc];
            acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, spread16((uint32_t)(x0 >> kh)), acc0,
                                                         0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, spread16((uint32_t)(x1 >> kh)), acc1,
                                                         0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                a1, spread16((uint32_t)(x0 >> (32 + kh))), acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                a1, spread16((uint32_t)(x1 >> (32 + kh))), acc1, 0, 0, 0);
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const uint32_t i = 32 * rh + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        hblk[i * kHStride + (lane & 31)] = d.lin[p0 + i] + d.scale * acc0[v];
        hblk[i * kHStride + 32 + (lane & 31)] = d.lin[p0 + i] + d.scale * acc1[v];
    }
}

// Adds to the fields of block blk + 1 (hnext) the change made by block blk's
// flips: scale * a[next rows][block blk columns] . (new - old), the spin
// changes as -1 / 0 / +1 bytes; wave w owns one 32 x 32 tile.
__device__ __forceinline__ void next_block_correction(const DevDense& d, const uint64_t* xt,
                                                      uint32_t xs, uint32_t blk,
                                                      const uint64_t* xnew, int32_t* hnext,
                                                      uint32_t wave, uint32_t lane) {
    const uint32_t ntile = d.n_pad / 32;
    const uint32_t rh = wave >> 1, jh = wave & 1;
    const uint32_t kh = 16 * (lane >> 5);
    const uint32_t j = 32 * jh + (lane & 31);
    const int8_t* t = d.a + ((size_t)(2 * (blk + 1) + rh) * ntile + 2 * blk) * 1024 + 16 * lane;
    const uint64_t o = xt[(size_t)j * xs + blk], nw = xnew[j];
    const uint64_t up = nw & ~o, dn = o & ~nw;
    v16i acc = {};
#pragma unroll
    for (uint32_t c = 0; c < 2; ++c) {
        const v4i a = *reinterpret_cast<const v4i*>(t + c * 1024);
        const uint32_t sh = 32 * c + kh;
        const v4i bp = spread16((uint32_t)(up >> sh));
        const v4i bn = spread16((uint32_t)(dn >> sh));
        const v4i bd = bp | (bn * 0xFF);  // bytes 0x01 (+1), 0xFF (-1) or 0
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bd, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const uint32_t i = 32 * rh + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        hnext[i * kHStride + j] += d.scale * acc[v];
    }
}

// Uniforms of block blk (wave w: positions 16 w .. +15) and its couplings.
__device__ __forceinline__ void stage_block(const DevDense& d, BlockTiles& T, uint32_t blk,
                                            uint32_t slot, uint32_t sweep, uint32_t key0,
                                            uint32_t key1, uint32_t wave, uint32_t lane) {
    const uint32_t p0 = blk * 64;
#pragma unroll
    for (uint32_t g = 0; g < 4; ++g) {
        const uint32_t q = 16 * wave + 4 * g;
        const U32x4 u =
            philox4x32_10((p0 + q) >> 2, slot, sweep, (kTagWord << 24) | lane, key0, key1);
#pragma unroll
        for (uint32_t t = 0; t < 4; ++t) T.u[(q + t) * 64 + lane] = u.v[t];
    }
    const int4* src = reinterpret_cast<const int4*>(d.wdiag + (size_t)blk * 4096);
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) T.w[t * 256 + threadIdx.x] = src[t * 256 + threadIdx.x];
}

// Acceptance u < metropolis_threshold(beta, de) with threshold_band: a
// proposal inside the band is "ambiguous" and needs the threshold itself,
// evaluated out of line to keep the unrolled block code small.
__device__ __attribute__((noinline)) bool exact_accept(uint32_t u, int32_t de, double beta) {
    return u < metropolis_threshold(beta, (int64_t)de);
}



// The 64 positions of one block in order (lane = replica).  hb: the block's
// fields at its start; wv: lane k holds the word of position p0 + k; wt: the
// block's couplings scale * a[p0 + p][p0 + k] in LDS (read as broadcasts).
// kFast: straight-line code for a full block; ambiguous proposals are only
// reported (the returned mask is nonzero) and the caller runs the block again
// with kFast = false, which evaluates their thresholds and honours cnt < 64.
template <bool kFast>
__device__ __forceinline__ uint64_t block_chain(int32_t (&hb)[64], uint64_t& wv,
                                                const uint32_t* ut, const int4* wt, uint32_t cnt,
                                                uint32_t lane, uint32_t dlim, double nbd,
                                                double beta, int64_t& eacc, uint32_t& nacc) {
    uint64_t amb_any = 0;
    // Row p of the block couplings (entries k > p) is read one position ahead.
    int4 wn[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) wn[j] = wt[j];
#pragma unroll
    for (int p = 0; p < 64; ++p) {
        int4 wc[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) wc[j] = wn[j];
        if (p + 1 < 64) {
#pragma unroll
            for (int j = (p + 2) / 4; j < 16; ++j) wn[j] = wt[(p + 1) * 16 + j];
        }
        if (!kFast && (uint32_t)p >= cnt) continue;  // past n in the last block
        const uint64_t x = readlane64(wv, p);
        const bool xb = lane_bit(x);
        const int32_t de = xb ? -hb[p] : hb[p];
        const uint32_t u = ut[p * 64 + lane];
        uint32_t lo_u, hi_u;
        threshold_band(de, nbd, lo_u, hi_u);
        const bool pos = de > 0 && (uint32_t)de <= dlim;
        bool acc = de <= 0 || (pos && u < lo_u);
        const bool amb = pos && u >= lo_u && u <= hi_u;
        if (kFast)
            amb_any |= wave_ballot(amb);
        else if (amb)
            acc = exact_accept(u, de, beta);
        const uint64_t m = wave_ballot(acc);
        const int32_t dx = acc ? (xb ? -1 : 1) : 0;
#pragma unroll
        for (int k = p + 1; k < 64; ++k) {
            const int4 c = wc[k >> 2];
            const int32_t w = (k & 3) == 0 ? c.x : (k & 3) == 1 ? c.y : (k & 3) == 2 ? c.z : c.w;
            hb[k] += __mul24(dx, w);
        }
        eacc += acc ? de : 0;
        nacc += (uint32_t)__builtin_popcountll(m);
        wv = lane == (uint32_t)p ? (x ^ m) : wv;
    }
    return amb_any;
}

__global__ __launch_bounds__(kDenseThreads) void sweep_dense_kernel(DevDense d, DevState s,
                                                                    const double* __restrict__ beta,
                                                                    uint32_t sweep, uint32_t key0,
                                                                    uint32_t key1) {
    extern __shared__ __align__(16) uint64_t xt[];
    __shared__ __align__(16) BlockTiles T;
    const uint32_t nchunk = d.n_pad / 64, xs = nchunk + 1, nblk = (d.n + 63) / 64;
    const uint32_t slot = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t n = d.n;
    uint64_t* words = s.spins + (size_t)slot * n;
    load_bit_rows(words, n, nchunk, xs, xt, wave, lane);
    const double b = beta[s.temp_of[slot * 64 + lane]];
    const uint32_t dlim = threshold_limit(b, d.dtab);
    const double nbd = -b * 1.4426950408889634;
    int64_t eacc = 0;
    uint32_t nacc = 0;
    __syncthreads();
    // Block 0: uniforms, couplings and fields.
    stage_block(d, T, 0, slot, sweep, key0, key1, wave, lane);
    if (wave == 1 || wave == 2) half_block_fields(d, xt, xs, 0, T.h[0], wave - 1, lane);
    __syncthreads();
    // Per block: wave 0 runs the block while waves 1 and 2 compute the next
    // block's fields from the spins before this block; after a barrier all
    // waves add this block's flips to those fields (matrix cores) and stage
    // the next block's uniforms and couplings.
    for (uint32_t blk = 0; blk < nblk; ++blk) {
        const uint32_t p0 = blk * 64;
        const uint32_t cnt = min(64u, n - p0);
        const uint32_t cur = blk & 1u;
        const bool more = blk + 1 < nblk;
        if (wave == 0) {
            const int32_t* hc = T.h[cur];
            int32_t hb[64];
#pragma unroll
            for (int k = 0; k < 64; ++k) hb[k] = hc[k * kHStride + lane];
            const uint64_t w0 = lane < cnt ? words[p0 + lane] : 0ull;
            const int64_t e0 = eacc;
            const uint32_t n0 = nacc;
            uint64_t wv = w0;
            bool redo = cnt < 64;
            if (!redo)
                redo = block_chain<true>(hb, wv, T.u, T.w, 64, lane, dlim, nbd, b, eacc, nacc) != 0;
            if (redo) {
#pragma unroll
                for (int k = 0; k < 64; ++k) hb[k] = hc[k * kHStride + lane];
                wv = w0;
                eacc = e0;
                nacc = n0;
                block_chain<false>(hb, wv, T.u, T.w, cnt, lane, dlim, nbd, b, eacc, nacc);
            }
            if (lane < cnt) words[p0 + lane] = wv;
            T.xnew[lane] = wave_transpose(wv, lane);
        } else if (more && (wave == 1 || wave == 2)) {
            half_block_fields(d, xt, xs, p0 + 64, T.h[cur ^ 1u], wave - 1, lane);
        }
        __syncthreads();
        if (more) {
            next_block_correction(d, xt, xs, blk, T.xnew, T.h[cur ^ 1u], wave, lane);
            stage_block(d, T, blk + 1, slot, sweep, key0, key1, wave, lane);
        }
        __syncthreads();
        if (wave == 0) xt[lane * xs + blk] = T.xnew[lane];
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
    extern __shared__ __align__(16) uint64_t xt[];
    __shared__ __align__(16) int32_t h[64 * kHStride];
    const uint32_t nchunk = d.n_pad / 64, xs = nchunk + 1;
    const uint32_t slot = blockIdx.x, blk = blockIdx.y;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t n = d.n, p0 = blk * 64;
    load_bit_rows(s.spins + (size_t)slot * n, n, nchunk, xs, xt, wave, lane);
    __syncthreads();
    block_fields(d, xt, xs, p0, h, wave, lane);
    __syncthreads();
    const uint32_t cnt = min(64u, n - p0);
    for (uint32_t i = wave; i < cnt; i += 4)
        out[((size_t)slot * n + d.order[p0 + i]) * 64 + lane] = h[i * kHStride + lane];
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
    // one workgroup per block that holds positions (n_pad may add empty blocks)
    hipLaunchKernelGGL(dense_fields_kernel, dim3(s.cells, (d.n + 63) / 64), dim3(kDenseThreads), lds,
                       stream, d, s, out);
    return hipGetLastError();
}

}  // namespace ige
