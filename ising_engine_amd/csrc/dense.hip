// dense.hip -- Metropolis sweep for dense couplings (fully connected SK,
// dense QUBO) with the local fields recomputed on int8 matrix cores.
//
// Model: q_ij = scale * a_ij with a_ij int8, stored as an n_pad x n_pad int8
// matrix in position space.  Two sweeps share the field recompute below:
// sweep_dense_bound_kernel (further down; decisions by accept bounds built
// before the sweep) whenever every beta is >= kDenseBoundMinBeta and the
// field range fits its 24-bit arithmetic, and sweep_dense_kernel otherwise.
// sweep_dense_kernel: one workgroup (4 waves) per slot walks the positions in
// blocks of 64:
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

#include "gang_sync.h"
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

// ---------------------------------------------------------------------------
// Accept-bound chain: sweep_dense_bound_kernel.
//
// The kernel above decides every proposal inside the sequential loop (a
// double-precision threshold band) and pushes every flip into all later
// fields of its block, so its one sequential wave issues about 120
// instructions per position.  Here that wave keeps only what is sequential:
//  * Accept bounds.  Before the sweep, dense_bounds_kernel turns the uniform
//    u of every (position, replica) into D = the largest dE in [0, dlim] with
//    dE == 0 or u < metropolis_threshold(beta, dE), on the whole GPU (the
//    sweep itself holds one CU per slot).  In the sweep a helper wave folds D
//    with the spin's sign into one key K = -(sgn + D + 1) (sgn = -x); the
//    chain accepts iff (h ^ sgn) + K < 0, i.e. dE = (1 - 2x) h <= D.
//    metropolis_threshold is non-increasing in dE (det_exp is within a few ulp
//    of exp, and beta >= kDenseBoundMinBeta separates neighbouring dE by far
//    more), so the accepted dE form the prefix [.., D] and the decision
//    equals the threshold's.
//  * Pulled fields.  The chain keeps the flips of the current half block (32
//    positions) as int8 bytes (+1, -1 or 0 in byte q of word q / 4) and forms
//    the field of position i as h0_i + scale * sum_q a_iq dx_q with v_dot4
//    over the words written so far; the coupling rows are read from LDS.
//  * Between the two halves of a block, two 32x32x32 int8 MFMAs add the first
//    half's flips to the second half's fields.
// Helper waves meanwhile compute the next block's fields from the spins on
// the matrix cores (as above) and stage its keys and coupling rows.  Waves:
// 0 the chain; 4 the keys and rows (it shares a SIMD with the chain, so it
// gets the light work); 1-3 and 5-7 the fields (all 64 rows, a sixth of the
// columns each, so that every expansion of spin bits into operand bytes feeds
// both row halves); 1, 2, 3 and 5 also combine the fields' tiles.

constexpr uint32_t kBoundThreads = 512;  // 8 waves
constexpr int kFieldPrefetch = 4;        // chunks of coupling rows in flight per field wave
constexpr uint32_t kRS = 68;             // dwords per replica row of the [replica][position] tiles

struct BoundTiles {
    int32_t h[64 * kRS];         // fields of the next block at its start, [replica][position]
    int32_t key[2][64 * kRS];    // accept keys of blocks blk and blk + 1, [replica][position]
    int32_t hacc[2][2][16][64];  // sums a.x of the next block: row half, replica half,
                                 // element, lane (added to by the six field waves,
                                 // cleared when read)
    uint4 rows[2][128];          // coupling rows (hdiag) of blocks blk and blk + 1
    v4i corr[2][2][64];          // coupling tiles (row half, column tile) of the next
                                 // block's rows against the current block's columns
    int32_t lin[64];             // linear terms of the next block
    v4i diag[2][64];             // coupling tile of rows 32.. against columns 0..31 of
                                 // blocks blk and blk + 1 (the chain's half correction)
    uint64_t xold[64];           // the finished block's spins per replica, before
    uint64_t xnew[64];           // and after the chain
    v4i corr2[2][2][64];         // tiles of the next block's rows against the columns
                                 // of the block before the current one
    uint64_t xold2[64];          // xold and xnew of the block before
    uint64_t xnew2[64];
};

struct BoundLane {
    double beta;  // this lane's inverse temperature
    double cd;    // ln 2 / beta
    double eps;   // error bound of y below (without its 1e-15 y part)
    uint32_t dlim;
};

// D = the largest d in [0, dlim] with d == 0 or u < metropolis_threshold(beta, d).
// u < thr(d) <=> u + 1 <= 2^32 det_exp(-beta d) <=> d <= y + theta with
// y = (32 - log2(u + 1)) ln 2 / beta and |theta| <= 1.5 eta / beta for a
// relative error eta of det_exp (a few ulp).  y is formed from u + 1 = 2^e m
// with log2(m) on v_log_f32 (absolute error below 2^-21 on [1, 2], checked
// over every float there; m itself is rounded to 24 bits, 8.6e-8 more), the
// rest in double: |y - y_true| <= 5.7e-7 ln2/beta + 1e-15 y.  When the
// fractional part of y is farther than eps = 6e-7 ln2/beta + 1e-12 + 1e-15 y
// from an integer, D = floor(y) (capped at dlim); otherwise D is one of two
// neighbours and one exact threshold decides (eps < 1/2 for
// beta >= kDenseBoundMinBeta).
__device__ __forceinline__ uint32_t accept_bound(uint32_t u, const BoundLane& bp) {
    if (u == 0xFFFFFFFFu) return 0u;
    const uint32_t v = u + 1u;
    const uint32_t z = __builtin_clz(v);
    const float mf = (float)(v << z) * 0x1p-31f;  // u + 1 = 2^(31 - z) mf, mf in [1, 2]
    const double y = ((double)(z + 1u) - (double)__builtin_amdgcn_logf(mf)) * bp.cd;
    if (y >= (double)bp.dlim + 2.0) return bp.dlim;
    const double d0 = floor(y);
    const double fr = y - d0;
    const double ep = bp.eps + 1e-15 * y;
    uint32_t D = (uint32_t)d0;
    if (fr < ep) {
        if (D > 0u && !exact_accept(u, (int32_t)D, bp.beta)) D -= 1u;
    } else if (fr > 1.0 - ep) {
        if (exact_accept(u, (int32_t)D + 1, bp.beta)) D += 1u;
    }
    return D < bp.dlim ? D : bp.dlim;
}

// Accept bounds of one sweep: D[((slot * nblk + blk) * 16 + g) * 64 + lane][q]
// for position 64 blk + 4 g + q of replica lane (one Philox call per thread:
// the word-view uniforms), -0x40000001 past n (a key that rejects).  One
// wave per (slot, block, group).  Workgroup 0 also zeroes the sweep's gang
// flags (zero: sync_count words at sync).
__global__ __launch_bounds__(256) void dense_bounds_kernel(DevDense d, DevState s,
                                                           const double* __restrict__ beta,
                                                           uint32_t sweep, uint32_t key0,
                                                           uint32_t key1, int4* __restrict__ D,
                                                           uint32_t waves, uint32_t* sync,
                                                           uint32_t sync_count) {
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < sync_count; i += blockDim.x) sync[i] = 0u;
    const uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= waves) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nblk = (d.n + 63) / 64;
    const uint32_t g = w % 16, blk = (w / 16) % nblk, slot = w / (16 * nblk);
    BoundLane bp;
    bp.beta = beta[s.temp_of[slot * 64 + lane]];
    bp.cd = 0.6931471805599453 / bp.beta;
    bp.eps = 6e-7 * bp.cd + 1e-12;
    bp.dlim = threshold_limit(bp.beta, d.dtab);
    const uint32_t p = blk * 64 + 4 * g;
    const U32x4 u = philox4x32_10(p >> 2, slot, sweep, (kTagWord << 24) | lane, key0, key1);
    int32_t v[4];
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q)
        v[q] = p + q < d.n ? (int32_t)accept_bound(u.v[q], bp) : -0x40000001;
    D[(size_t)w * 64 + lane] = int4{v[0], v[1], v[2], v[3]};
}

// Keys of block nb into key[replica][position] from its accept bounds Db
// (this lane = replica: K = -(sgn + D + 1), sgn = -x at the block's start),
// the block's coupling rows (hdiag tiles 2 nb, 2 nb + 1) into rows, its tile
// of rows 32.. against columns 0..31 into diag, and what combine_fields
// reads for block nb: the tiles of its rows against block nb - 1's columns
// (and, when `stale`, block nb - 2's) and its linear terms.  One wave.
__device__ __forceinline__ void stage_keys(const DevDense& d, uint32_t nb, const int4* Db,
                                           const uint64_t* xt, uint32_t xs, int32_t* key,
                                           uint4* rows, v4i* diag, v4i (*corr)[2][64],
                                           v4i (*corr2)[2][64], bool stale, int32_t* lin,
                                           uint32_t lane) {
    const uint32_t ntile = d.n_pad / 32;
    const uint4* src = reinterpret_cast<const uint4*>(d.hdiag + (size_t)nb * 2048);
    rows[lane] = src[lane];
    rows[64 + lane] = src[64 + lane];
    lin[lane] = d.lin[nb * 64 + lane];
    diag[lane] = *reinterpret_cast<const v4i*>(
        d.a + ((size_t)(2 * nb + 1) * ntile + 2 * nb) * 1024 + 16 * lane);
    if (nb > 0) {
#pragma unroll
        for (uint32_t r = 0; r < 2; ++r)
#pragma unroll
            for (uint32_t c = 0; c < 2; ++c)
                corr[r][c][lane] = *reinterpret_cast<const v4i*>(
                    d.a + ((size_t)(2 * nb + r) * ntile + 2 * (nb - 1) + c) * 1024 + 16 * lane);
    }
    if (stale) {
#pragma unroll
        for (uint32_t r = 0; r < 2; ++r)
#pragma unroll
            for (uint32_t c = 0; c < 2; ++c)
                corr2[r][c][lane] = *reinterpret_cast<const v4i*>(
                    d.a + ((size_t)(2 * nb + r) * ntile + 2 * (nb - 2) + c) * 1024 + 16 * lane);
    }
    const uint64_t xb = xt[lane * xs + nb];
    int4 dv[16];
#pragma unroll
    for (int g = 0; g < 16; ++g) dv[g] = Db[g * 64 + lane];
#pragma unroll
    for (int g = 0; g < 16; ++g) {
        const int32_t k[4] = {dv[g].x, dv[g].y, dv[g].z, dv[g].w};
        int32_t kv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t sgn = -(int32_t)((xb >> (4 * g + q)) & 1u);
            kv[q] = -(sgn + k[q] + 1);
        }
        *reinterpret_cast<int4*>(key + lane * kRS + 4 * g) = int4{kv[0], kv[1], kv[2], kv[3]};
    }
}

// The coupling rows of a wave's first kFieldPrefetch chunks of [c0, c1)
// (clamped to c1 - 1), loaded ahead of field_sums: they do not depend on the
// spins, so a caller can issue them before it waits for spins.
struct FieldRows {
    v4i pa[kFieldPrefetch][4];  // [chunk][row half * 2 + tile]
};

__device__ __forceinline__ void field_rows(const DevDense& d, uint32_t p0, uint32_t c0,
                                           uint32_t c1, FieldRows& r, uint32_t lane) {
    if (c1 <= c0) return;
    const uint32_t ntile = d.n_pad / 32;
    const int8_t* ar0 = d.a + ((size_t)(p0 / 32) * ntile) * 1024 + 16 * lane;
    const int8_t* ar1 = ar0 + (size_t)ntile * 1024;
#pragma unroll
    for (int i = 0; i < kFieldPrefetch; ++i) {
        const size_t o = (size_t)min(c0 + i, c1 - 1) * 2048;
        r.pa[i][0] = *reinterpret_cast<const v4i*>(ar0 + o);
        r.pa[i][1] = *reinterpret_cast<const v4i*>(ar0 + o + 1024);
        r.pa[i][2] = *reinterpret_cast<const v4i*>(ar1 + o);
        r.pa[i][3] = *reinterpret_cast<const v4i*>(ar1 + o + 1024);
    }
}

// a.x of the 64 rows of the block at p0 over the 64-column chunks [c0, c1)
// (a multiple of kFieldPrefetch chunks, possibly none), all 64 replicas, added
// into hacc; one wave, with the first chunks' rows in r (field_rows).  Each
// chunk's spin bytes (two replica halves x two 32-column tiles) feed the MFMAs
// of both 32-row halves.
__device__ __forceinline__ void field_sums(const DevDense& d, const uint64_t* xt, uint32_t xs,
                                           uint32_t p0, uint32_t c0, uint32_t c1, FieldRows& r,
                                           int32_t (*hacc)[2][16][64], uint32_t lane) {
    // Chunks of coupling rows in flight: the rows stream from the memory
    // behind L2, so a wave keeps P * 4 KB in flight.  With whole groups of P
    // chunks the loop has no early exit, and the loads of the next group stay
    // in flight across the loop's back edge (the last group reloads its last
    // chunk: a branch around those loads measured slower).
    constexpr int P = kFieldPrefetch;
    if (c1 <= c0) return;
    const uint32_t kh = 16 * (lane >> 5);
    const uint32_t ntile = d.n_pad / 32;
    const int8_t* ar0 = d.a + ((size_t)(p0 / 32) * ntile) * 1024 + 16 * lane;
    const int8_t* ar1 = ar0 + (size_t)ntile * 1024;
    const uint64_t* xr0 = xt + (size_t)(lane & 31) * xs;
    const uint64_t* xr1 = xt + (size_t)(32 + (lane & 31)) * xs;
    v16i acc[2][2] = {};  // [row half][replica half]
    for (uint32_t kb = c0; kb < c1; kb += P) {
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const uint32_t kc = kb + i;
            v4i a[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) a[t] = r.pa[i][t];
            {
                const size_t o = (size_t)min(kc + (uint32_t)P, c1 - 1) * 2048;
                r.pa[i][0] = *reinterpret_cast<const v4i*>(ar0 + o);
                r.pa[i][1] = *reinterpret_cast<const v4i*>(ar0 + o + 1024);
                r.pa[i][2] = *reinterpret_cast<const v4i*>(ar1 + o);
                r.pa[i][3] = *reinterpret_cast<const v4i*>(ar1 + o + 1024);
            }
            const uint64_t x0 = xr0[kc], x1 = xr1[kc];
#pragma unroll
            for (int t = 0; t < 2; ++t) {  // 32-column tile of the chunk
                const v4i b0 = spread16((uint32_t)(x0 >> (32 * t + kh)));
                const v4i b1 = spread16((uint32_t)(x1 >> (32 * t + kh)));
#pragma unroll
                for (int rr = 0; rr < 2; ++rr) {
                    acc[rr][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[2 * rr + t], b0,
                                                                       acc[rr][0], 0, 0, 0);
                    acc[rr][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[2 * rr + t], b1,
                                                                       acc[rr][1], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) atomicAdd(&hacc[rr][j][v][lane], acc[rr][j][v]);
}

__device__ __forceinline__ void block_partial(const DevDense& d, const uint64_t* xt, uint32_t xs,
                                              uint32_t p0, uint32_t c0, uint32_t c1,
                                              int32_t (*hacc)[2][16][64], uint32_t lane) {
    FieldRows r;
    field_rows(d, p0, c0, c1, r, lane);
    field_sums(d, xt, xs, p0, c0, c1, r, hacc, lane);
}

constexpr uint32_t kPartWords = 2 * 2 * 16 * 64;  // one partial sum, BoundTiles::hacc layout
// Buffers of partial sums per helper: block b's in buffer b % kPartBufs.  A
// helper sums blocks 0, 1 and 2 without waiting and block b >= 3 once block
// b - 3 is published, which the main workgroup does after it has read block
// b - 3's sums, so three buffers are never overwritten before they are read.
constexpr uint32_t kPartBufs = 3;

// Fields of block nb at its start, one 32 x 32 tile (rows 32 rh.., replicas
// 32 jh..) per wave: lin + scale * (the sums in hacc + block nb - 1's flips
// applied to rows of block nb, and block nb - 2's when `stale`: a helper
// summed its columns from the spins before them), into h[replica][position];
// the sums are cleared for the next block.  nb == 0: no flips yet.  The
// coupling tiles and linear terms were staged in LDS (stage_keys).
template <bool kGang>
__device__ __forceinline__ void combine_fields(const DevDense& d, BoundTiles& T, uint32_t nb,
                                               uint32_t rh, uint32_t jh, uint32_t lane,
                                               bool stale) {
    const uint32_t j = 32 * jh + (lane & 31);
    v16i acc = {};
    if (nb > 0) {
        const uint32_t kh = 16 * (lane >> 5);
        const uint64_t o = T.xold[j], nw = T.xnew[j];
        const uint64_t up = nw & ~o, dn = o & ~nw;
#pragma unroll
        for (uint32_t c = 0; c < 2; ++c) {
            const uint32_t sh = 32 * c + kh;
            const v4i bd = spread16((uint32_t)(up >> sh)) | (spread16((uint32_t)(dn >> sh)) * 0xFF);
            acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(T.corr[rh][c][lane], bd, acc, 0, 0, 0);
        }
    }
    if constexpr (kGang) {
        if (stale) {
            const uint32_t kh = 16 * (lane >> 5);
            const uint64_t o = T.xold2[j], nw = T.xnew2[j];
            const uint64_t up = nw & ~o, dn = o & ~nw;
#pragma unroll
            for (uint32_t c = 0; c < 2; ++c) {
                const uint32_t sh = 32 * c + kh;
                const v4i bd =
                    spread16((uint32_t)(up >> sh)) | (spread16((uint32_t)(dn >> sh)) * 0xFF);
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(T.corr2[rh][c][lane], bd, acc, 0, 0,
                                                            0);
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const uint32_t i = 32 * rh + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int32_t sum = T.hacc[rh][jh][v][lane] + acc[v];
        T.hacc[rh][jh][v][lane] = 0;
        T.h[j * kRS + i] = T.lin[i] + d.scale * sum;
    }
}

// Flags of a slot's field gang (one XCD; flags and data meet in its L2 and
// are read device-coherently): word 0 counts the blocks whose spins the main
// workgroup has published, word 1 + k the partial sums helper k has
// published, word 5 the XCDs of the members, word 6 their first barrier.
constexpr uint32_t kGangPublished = 0, kGangPartial = 1, kGangXcds = 5, kGangBarrier = 6;
constexpr uint32_t kGangXcdStride = 8;  // XCDs: workgroups launched per member

__device__ __forceinline__ uint32_t load_flag(const uint32_t* f) {
    return __hip_atomic_load(const_cast<uint32_t*>(f), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Polls until *f >= v, at most spin_limit polls; on expiry raises *err (and
// every later wait then returns at once).  One lane calls it.
__device__ __forceinline__ void wait_flag(const uint32_t* f, uint32_t v, uint32_t spin_limit,
                                          uint32_t* err) {
    uint32_t polls = 0;
    while (load_flag(f) < v) {
        if (load_flag(err) != 0u) break;
        if (++polls > spin_limit) {
            atomicOr(err, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Adds the H helpers' sums for tile (rh, jh) (hp + k * hstride, helper k,
// read device-coherently once its flag reaches nb + 1) into hacc; one wave.
// Every helper's sums load before any is added (one round trip to the L2).
__device__ __forceinline__ void add_helper_sums(int32_t (*hacc)[2][16][64], const uint32_t* sy,
                                                uint32_t nb, const int32_t* hp, size_t hstride,
                                                uint32_t H, uint32_t rh, uint32_t jh,
                                                uint32_t lane, uint32_t spin_limit,
                                                uint32_t* err) {
    if (lane == 0)
        for (uint32_t k = 0; k < H; ++k) wait_flag(&sy[kGangPartial + k], nb + 1, spin_limit, err);
    __builtin_amdgcn_wave_barrier();
    int32_t part[kMaxDenseHelpers][16];
#pragma unroll
    for (uint32_t k = 0; k < kMaxDenseHelpers; ++k) {
        if (k < H) {
            const int32_t* src = hp + k * hstride + (rh * 2 + jh) * 16 * 64 + lane;
#pragma unroll
            for (int v = 0; v < 16; ++v)
                part[k][v] = __hip_atomic_load(const_cast<int32_t*>(src + v * 64),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        int32_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < kMaxDenseHelpers; ++k)
            if (k < H) sum += part[k][v];
        atomicAdd(&hacc[rh][jh][v][lane], sum);
    }
}

// Writes the low byte of e into byte k of w, keeping the other bytes (SDWA);
// k is a constant once the caller's loop is unrolled.
__device__ __forceinline__ void put_byte(uint32_t& w, int32_t e, int k) {
    switch (k) {
    case 0:
        asm("v_mov_b32_sdwa %0, %1 dst_sel:BYTE_0 dst_unused:UNUSED_PRESERVE src0_sel:DWORD"
            : "+v"(w) : "v"(e));
        break;
    case 1:
        asm("v_mov_b32_sdwa %0, %1 dst_sel:BYTE_1 dst_unused:UNUSED_PRESERVE src0_sel:DWORD"
            : "+v"(w) : "v"(e));
        break;
    case 2:
        asm("v_mov_b32_sdwa %0, %1 dst_sel:BYTE_2 dst_unused:UNUSED_PRESERVE src0_sel:DWORD"
            : "+v"(w) : "v"(e));
        break;
    default:
        asm("v_mov_b32_sdwa %0, %1 dst_sel:BYTE_3 dst_unused:UNUSED_PRESERVE src0_sel:DWORD"
            : "+v"(w) : "v"(e));
        break;
    }
}

// Bit q of the result = (byte q of the flip words is nonzero), q < 32: the
// bytes are 0x00, 0x01 or 0xFF, so bit 0 of each byte tells; the multiply
// moves bit 8k of a word to bit 24 + k with no overlapping terms.
__device__ __forceinline__ uint32_t flip_bits(const uint32_t (&dxp)[8]) {
    uint32_t f = 0;
#pragma unroll
    for (int g = 0; g < 8; ++g) f |= (((dxp[g] & 0x01010101u) * 0x01020408u) >> 24) << (4 * g);
    return f;
}

// Positions OFF .. OFF + 31 of the block in order (lane = replica).  xh: this
// replica's spins there at the block start (bit i); rows: row i's 32
// couplings a (int8, columns OFF .. OFF + 31) at rows[2 i], rows[2 i + 1]
// (LDS, read as broadcasts).  dxp receives the flips as bytes; eacc the
// accepted dE.
// kScale1: scale == 1, so the dot products accumulate onto h0 directly.
// The rows are read from LDS kRowAhead positions ahead of their use.
constexpr int kRowAhead = 3;

template <int OFF, bool kScale1>
__device__ __forceinline__ void chain_half(int32_t (&hr)[64], const int32_t (&kr)[64],
                                           uint32_t xh, const uint4* rows, int32_t scale,
                                           uint32_t (&dxp)[8], int32_t& eacc) {
#pragma unroll
    for (int g = 0; g < 8; ++g) dxp[g] = 0u;
    // groups holding flips of positions before i: 0 .. (i - 1) / 4
    auto groups = [](int i) { return (i + 3) >> 2; };
    // ring slot k % (kRowAhead + 1) holds row k; row i + kRowAhead is loaded at
    // position i into the slot of row i - 1, which is no longer needed
    uint4 ring[kRowAhead + 1][2];
#pragma unroll
    for (int i = 1; i < kRowAhead; ++i) {
        ring[i][0] = rows[2 * i];
        if (groups(i) > 4) ring[i][1] = rows[2 * i + 1];
    }
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        const int ng = groups(i);
        const int nx = i + kRowAhead;
        if (nx < 32) {
            ring[nx % (kRowAhead + 1)][0] = rows[2 * nx];
            if (groups(nx) > 4) ring[nx % (kRowAhead + 1)][1] = rows[2 * nx + 1];
        }
        const uint4 r0 = ring[i % (kRowAhead + 1)][0], r1 = ring[i % (kRowAhead + 1)][1];
        const uint32_t w[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
        int32_t s = kScale1 ? hr[OFF + i] : 0;
#pragma unroll
        for (int g = 0; g < ng; ++g)
            s = __builtin_amdgcn_sdot4((int32_t)dxp[g], (int32_t)w[g], s, false);
        const int32_t h = kScale1 ? s : hr[OFF + i] + __mul24(s, scale);
        const int32_t sgn = (int32_t)(xh << (31 - i)) >> 31;
        const uint32_t t = ((uint32_t)h ^ (uint32_t)sgn) + (uint32_t)kr[OFF + i];
        const int32_t m = (int32_t)t >> 31;
        const int32_t e = m & (sgn | 1);
        put_byte(dxp[i >> 2], e, i & 3);
        eacc = __mul24(e, h) + eacc;
    }
}

// Member m of a slot's gang takes these groups of kFieldPrefetch column chunks
// of every block's fields.  With helpers the main workgroup keeps two (for
// its two field waves that do not combine: the four that do wait for and add
// the helpers' sums while the chain runs) and the helpers share the rest.
__device__ __forceinline__ void gang_groups(uint32_t ngrp, uint32_t H, uint32_t m, uint32_t& gb,
                                            uint32_t& ge) {
    const uint32_t g0 = H ? 2u : ngrp;
    if (m == 0) {
        gb = 0;
        ge = g0;
    } else {
        gb = g0 + (m - 1) * (ngrp - g0) / H;
        ge = g0 + m * (ngrp - g0) / H;
    }
}

// One slot's sweep.  Without helpers (g.helpers == 0) workgroup blockIdx.x
// is slot blockIdx.x.  With helpers the grid is (kGangXcdStride * (1 + H),
// cells): the workgroups with blockIdx.x % kGangXcdStride == 0 are the
// slot's gang (workgroups are dispatched to the XCDs in turn, so they share
// one), member 0 the main workgroup and 1..H the helpers.  They check the XCD
// at a first barrier; if they do not share one, the helpers leave and the
// main workgroup computes every field itself.  A helper computes its columns'
// share of block b's fields from the spins at the start of block b - 2 (the
// main workgroup publishes every finished block's spins), so it has two
// blocks' time, and publishes the sums; the main workgroup's combine waves
// add them at the end of block b - 1's chain and apply block b - 2's flips in
// the helpers' columns with the other flip corrections.  Every wait is
// bounded by g.spin_limit polls.
// kGang: the launch has helpers (the grid above); without, no helper code is
// compiled in.
template <bool kScale1, bool kGang>
__global__ __launch_bounds__(kBoundThreads) void sweep_dense_bound_kernel(
    DevDense d, DevState s, const int4* __restrict__ bounds, DevDenseGang g) {
    extern __shared__ __align__(16) uint64_t xt[];
    __shared__ __align__(16) BoundTiles T;
    const uint32_t nchunk = d.n_pad / 64, xs = nchunk + 1, nblk = (d.n + 63) / 64;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t n = d.n;
    uint32_t slot, m = 0, H = 0;
    uint32_t* sy = nullptr;
    if constexpr (kGang) {
        if (blockIdx.x % kGangXcdStride != 0) return;  // not a member: on another XCD
        m = blockIdx.x / kGangXcdStride;
        slot = blockIdx.y;
        sy = g.sync + slot * kDenseGangSyncWords;
        if (threadIdx.x == 0) atomicOr(&sy[kGangXcds], 1u << xcc_id());
        uint32_t target = 0;
        gang_barrier(&sy[kGangBarrier], target, 1 + g.helpers, g.spin_limit, g.err);
        const bool one_xcd = __builtin_popcount(load_flag(&sy[kGangXcds])) == 1;
        if (!one_xcd && m > 0) return;  // the main workgroup computes every field
        H = one_xcd ? g.helpers : 0u;
    } else {
        slot = blockIdx.x;
    }
    uint64_t* words = s.spins + (size_t)slot * n;
    for (uint32_t c = wave; c < nchunk; c += kBoundThreads / 64) {
        const uint32_t p = c * 64 + lane;
        xt[lane * xs + c] = wave_transpose(p < n ? words[p] : 0ull, lane);
    }
    for (uint32_t i = threadIdx.x; i < kPartWords; i += kBoundThreads)
        (&T.hacc[0][0][0][0])[i] = 0;
    __syncthreads();
    const uint32_t ngrp = nchunk / kFieldPrefetch;  // n_pad is a multiple of 512
    uint32_t gb, ge;
    gang_groups(ngrp, H, m, gb, ge);
    uint64_t* xpub = g.xpub + (size_t)slot * nblk * 64;
    int32_t* part = g.part + (size_t)slot * g.helpers * kPartBufs * kPartWords;
    if (kGang && m > 0) {
        // A helper: all eight waves take columns; block b's sums go to buffer b % kPartBufs.
        const uint32_t c0 = (gb + wave * (ge - gb) / 8) * kFieldPrefetch;
        const uint32_t c1 = (gb + (wave + 1) * (ge - gb) / 8) * kFieldPrefetch;
        int32_t* mine = part + (size_t)(m - 1) * kPartBufs * kPartWords;
        // Block b's sums from the spins at the start of block b - 2 (two blocks of
        // slack; the main workgroup corrects block b - 2's flips in these columns).
        for (uint32_t b = 0; b < nblk; ++b) {
            // the coupling rows load while the wave waits for the spins
            FieldRows r;
            field_rows(d, b * 64, c0, c1, r, lane);
            if (b >= 3) {  // the spins of block b - 3, the last before block b - 2
                if (wave == 0) {
                    if (lane == 0) wait_flag(&sy[kGangPublished], b - 2, g.spin_limit, g.err);
                    __builtin_amdgcn_wave_barrier();
                    xt[lane * xs + (b - 3)] = __hip_atomic_load(xpub + (b - 3) * 64 + lane,
                                                                __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_AGENT);
                }
                __syncthreads();
            }
            field_sums(d, xt, xs, b * 64, c0, c1, r, T.hacc, lane);
            __syncthreads();
            int32_t* dst = mine + (b % kPartBufs) * kPartWords;
            for (uint32_t i = threadIdx.x; i < kPartWords; i += kBoundThreads) {
                dst[i] = (&T.hacc[0][0][0][0])[i];
                (&T.hacc[0][0][0][0])[i] = 0;
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_store(&sy[kGangPartial + m - 1], b + 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }
    const int4* Dslot = bounds + (size_t)slot * nblk * 1024;
    // Field waves 1, 2, 3, 5, 6, 7 are numbered fw = 0..5 and take the main
    // workgroup's groups [gb, ge) in whole prefetch groups (with helpers only
    // fw 4 and 5 do); the first four also combine tile (rh, jh).
    const bool fieldw = wave != 0 && wave != 4;
    const uint32_t fw = wave < 4 ? wave - 1 : wave - 2;
    const bool combw = fieldw && fw < 4;
    // with helpers only fw 4 and 5 take the main workgroup's columns
    const uint32_t nfw = H ? 2u : 6u, ifw = H ? fw - 4 : fw;
    const uint32_t ch0 = (!H || !combw) ? (gb + ifw * (ge - gb) / nfw) * kFieldPrefetch : 0u;
    const uint32_t ch1 = (!H || !combw) ? (gb + (ifw + 1) * (ge - gb) / nfw) * kFieldPrefetch : 0u;
    const uint32_t rh = fw >> 1, jh = fw & 1u;
    // block nb's fields need block nb - 2's flips in its columns when a helper
    // owns that chunk (the main workgroup owns chunks [0, ge * kFieldPrefetch))
    auto stale = [&](uint32_t nb) {
        return kGang && H && nb >= 2 && nb - 2 >= ge * kFieldPrefetch;
    };
    // a field wave's share of block nb's sums; the combine waves then add the
    // helpers' sums for their tile (still while the chain runs)
    auto fields = [&](uint32_t nb) {
        block_partial(d, xt, xs, nb * 64, ch0, ch1, T.hacc, lane);
        if (kGang && combw && H)
            add_helper_sums(T.hacc, sy, nb, part + (nb % kPartBufs) * kPartWords,
                            kPartBufs * kPartWords, H,
                            rh, jh, lane, g.spin_limit, g.err);
    };
    // Block 0: fields, keys and rows.
    if (fieldw) {
        fields(0);
    } else if (wave == 4) {
        stage_keys(d, 0, Dslot, xt, xs, T.key[0], T.rows[0], T.diag[0], T.corr, T.corr2, false,
                   T.lin, lane);
    }
    __syncthreads();
    if (combw) combine_fields<kGang>(d, T, 0, rh, jh, lane, false);
    __syncthreads();
    int64_t eacc = 0;
    uint32_t nacc = 0;
    for (uint32_t blk = 0; blk < nblk; ++blk) {
        const bool more = blk + 1 < nblk;
        if (wave == 0) {
            int32_t hr[64], kr[64];
            const int32_t* hs = T.h + lane * kRS;
            const int32_t* ks = T.key[blk & 1u] + lane * kRS;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int4 a = *reinterpret_cast<const int4*>(hs + 4 * q);
                const int4 b = *reinterpret_cast<const int4*>(ks + 4 * q);
                hr[4 * q] = a.x; hr[4 * q + 1] = a.y; hr[4 * q + 2] = a.z; hr[4 * q + 3] = a.w;
                kr[4 * q] = b.x; kr[4 * q + 1] = b.y; kr[4 * q + 2] = b.z; kr[4 * q + 3] = b.w;
            }
            const uint64_t myx = xt[lane * xs + blk];
            const v4i atile = T.diag[blk & 1u][lane];
            const uint4* rows = T.rows[blk & 1u];
            uint32_t dxp[8];
            int32_t eb = 0;
            chain_half<0, kScale1>(hr, kr, (uint32_t)myx, rows, d.scale, dxp, eb);
            const uint32_t f0 = flip_bits(dxp);
            // The first half's flips into the second half's fields.
            {
                v4i b0, b1;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const auto p = __builtin_amdgcn_permlane32_swap(dxp[q], dxp[4 + q], false, false);
                    b0[q] = (int32_t)p[0];
                    b1[q] = (int32_t)p[1];
                }
                const v16i c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(atile, b0, v16i{}, 0, 0, 0);
                const v16i c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(atile, b1, v16i{}, 0, 0, 0);
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const auto p = __builtin_amdgcn_permlane32_swap((uint32_t)c0[v], (uint32_t)c1[v],
                                                                    false, false);
                    const int r = 32 + 8 * (v >> 2) + (v & 3);
                    hr[r] += kScale1 ? (int32_t)p[0] : __mul24((int32_t)p[0], d.scale);
                    hr[r + 4] += kScale1 ? (int32_t)p[1] : __mul24((int32_t)p[1], d.scale);
                }
            }
            chain_half<32, kScale1>(hr, kr, (uint32_t)(myx >> 32), rows + 64, d.scale, dxp, eb);
            const uint32_t f1 = flip_bits(dxp);
            const uint64_t xn = myx ^ ((uint64_t)f1 << 32 | f0);
            nacc += (uint32_t)__builtin_popcount(f0) + (uint32_t)__builtin_popcount(f1);
            eacc += eb;
            if constexpr (kGang) {
                T.xold2[lane] = T.xold[lane];
                T.xnew2[lane] = T.xnew[lane];
            }
            T.xold[lane] = myx;
            T.xnew[lane] = xn;
        } else if (more && fieldw) {
            fields(blk + 1);
        } else if (more) {
            stage_keys(d, blk + 1, Dslot + (size_t)(blk + 1) * 1024, xt, xs,
                       T.key[(blk + 1) & 1u], T.rows[(blk + 1) & 1u], T.diag[(blk + 1) & 1u],
                       T.corr, T.corr2, stale(blk + 1), T.lin, lane);
        }
        __syncthreads();
        if (wave == 0) {
            const uint64_t xn = T.xnew[lane];
            xt[lane * xs + blk] = xn;
            const uint64_t w = wave_transpose(xn, lane);
            const uint32_t p = blk * 64 + lane;
            if (p < n) words[p] = w;
            if (kGang && H && blk + 4 <= nblk) {  // the helpers read blocks up to nblk - 4
                xpub[blk * 64 + lane] = xn;
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                if (lane == 0)
                    __hip_atomic_store(&sy[kGangPublished], blk + 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if (more && combw) {
            combine_fields<kGang>(d, T, blk + 1, rh, jh, lane, stale(blk + 1));
        }
        __syncthreads();
    }
    if (wave == 0) {
        s.energy[slot * 64 + lane] += eacc;
        uint32_t tot = nacc;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0 && tot) atomicAdd(s.accepted, (unsigned long long)tot);
    }
}

bool set_lds_limit(const void* fn, size_t bytes) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
           hipSuccess;
}

}  // namespace

size_t dense_lds_bytes_host(uint32_t n_pad) { return dense_lds_bytes(n_pad); }

bool dense_bounds_eligible(const DevDense& d) {
    return d.hdiag != nullptr && d.dtab < (1u << 22) && d.scale > 0 && d.scale < (1 << 22) &&
           sizeof(BoundTiles) + dense_lds_bytes(d.n_pad) <= 160u * 1024u;
}

size_t dense_bounds_words(const DevDense& d, uint32_t cells) {
    return (size_t)cells * ((d.n + 63) / 64) * 4096;
}

// Helpers only while every slot's gang fits one XCD (one workgroup per CU:
// the kernel's LDS) and has columns to share; a slot's fields take a few
// groups of kFieldPrefetch chunks, so small models keep fewer helpers.
uint32_t dense_gang_helpers(const DevDense& d, uint32_t cells) {
    constexpr uint32_t kXcdCus = 32;
    const uint32_t ngrp = d.n_pad / 64 / kFieldPrefetch;
    if (cells == 0 || cells * (1 + kMaxDenseHelpers) > kXcdCus || ngrp < 8) return 0;
    return kMaxDenseHelpers;
}

size_t dense_gang_xpub_words(const DevDense& d, uint32_t cells) {
    return (size_t)cells * ((d.n + 63) / 64) * 64;
}

size_t dense_gang_part_words(uint32_t cells, uint32_t helpers) {
    return (size_t)cells * helpers * kPartBufs * kPartWords;
}

hipError_t launch_sweep_dense(const DevDense& d, const DevState& s, uint32_t sweep, uint64_t seed,
                              const double* beta, int32_t* bounds, const DevDenseGang* gang,
                              hipStream_t stream) {
    const size_t lds = dense_lds_bytes(d.n_pad);
    if (bounds) {
        const bool gk = gang && gang->helpers;
        const void* fn =
            d.scale == 1
                ? (gk ? reinterpret_cast<const void*>(&sweep_dense_bound_kernel<true, true>)
                      : reinterpret_cast<const void*>(&sweep_dense_bound_kernel<true, false>))
                : (gk ? reinterpret_cast<const void*>(&sweep_dense_bound_kernel<false, true>)
                      : reinterpret_cast<const void*>(&sweep_dense_bound_kernel<false, false>));
        if (!dense_bounds_eligible(d) || !set_lds_limit(fn, lds)) return hipErrorInvalidValue;
        DevDenseGang g;
        dim3 grid(s.cells);
        if (gang && gang->helpers) {
            if (gang->helpers > kMaxDenseHelpers || !gang->sync || !gang->err || !gang->xpub ||
                !gang->part)
                return hipErrorInvalidValue;
            g = *gang;
            grid = dim3(kGangXcdStride * (1 + g.helpers), s.cells);
        }
        const uint32_t waves = s.cells * ((d.n + 63) / 64) * 16;
        int4* D = reinterpret_cast<int4*>(bounds);
        hipLaunchKernelGGL(dense_bounds_kernel, dim3((waves + 3) / 4), dim3(256), 0, stream, d, s,
                           beta, sweep, (uint32_t)seed, (uint32_t)(seed >> 32), D, waves, g.sync,
                           g.helpers ? s.cells * kDenseGangSyncWords : 0u);
        if (d.scale == 1 && gk)
            hipLaunchKernelGGL((sweep_dense_bound_kernel<true, true>), grid, dim3(kBoundThreads),
                               lds, stream, d, s, D, g);
        else if (d.scale == 1)
            hipLaunchKernelGGL((sweep_dense_bound_kernel<true, false>), grid, dim3(kBoundThreads),
                               lds, stream, d, s, D, g);
        else if (gk)
            hipLaunchKernelGGL((sweep_dense_bound_kernel<false, true>), grid, dim3(kBoundThreads),
                               lds, stream, d, s, D, g);
        else
            hipLaunchKernelGGL((sweep_dense_bound_kernel<false, false>), grid, dim3(kBoundThreads),
                               lds, stream, d, s, D, g);
        return hipGetLastError();
    }
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
