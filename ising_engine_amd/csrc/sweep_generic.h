// sweep_generic.h -- the lane-per-replica Metropolis update of one aligned
// group of (up to) four positions, shared by the single-state sweep kernel
// and the persistent multi-instance kernel.
//
// One wavefront, lane r = replica r.  The wave first gathers the whole
// group's CSR entries in one coalesced pass (lane e holds entry e: its
// coupling and its neighbour's u64 word), so every neighbour word of the four
// positions is in flight at once; it then walks the entries with readlane,
// applying each neighbour word as a lane mask (v_cndmask on an SGPR pair).
// Uniforms (word view): Philox4x32-10 with counter (p >> 2, slot, sweep,
// 1 << 24 | lane), word p & 3: one call per lane per group.
#pragma once

#include <hip/hip_runtime.h>

#include "ige_math.h"

namespace ige {

__device__ __forceinline__ bool lane_bit(uint64_t word) {
    return __builtin_amdgcn_inverse_ballot_w64(word);
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Value of v in lane (lane ^ J), with VALU lane exchanges where gfx950 has
// them: v_permlane32_swap / v_permlane16_swap (swap one half of the wave with
// the other), DPP row_ror:8 and quad_perm within rows, ds_swizzle for xor 4.
template <int J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, uint32_t lane) {
    if constexpr (J == 32) {
        const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (lane & 32u) ? p[0] : p[1];
    } else if constexpr (J == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane & 16u) ? p[0] : p[1];
    } else if constexpr (J == 8) {
        return __builtin_amdgcn_update_dpp(0u, v, 0x128, 0xF, 0xF, false);  // row_ror:8
    } else if constexpr (J == 4) {
        return __builtin_amdgcn_ds_swizzle(v, (4 << 10) | 0x1F);  // and 31, xor 4
    } else if constexpr (J == 2) {
        return __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xF, 0xF, false);  // quad_perm 2301
    } else {
        static_assert(J == 1, "lane distance must be a power of two <= 32");
        return __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xF, 0xF, false);  // quad_perm 1032
    }
}

// One stage of the 64x64 transpose for lane distance J < 32: lanes l and
// l ^ J swap the off-diagonal J x J blocks of each 2J x 2J block.  The lane
// with bit J clear keeps the low J bits of each 2J-bit group and takes its
// partner's, shifted up; the other keeps the high bits and takes its
// partner's, shifted down.
template <int J>
__device__ __forceinline__ uint64_t transpose_stage(uint64_t x, uint64_t m, uint32_t lane) {
    const uint64_t y = (uint64_t)xor_lane<J>((uint32_t)x, lane) |
                       ((uint64_t)xor_lane<J>((uint32_t)(x >> 32), lane) << 32);
    const bool h = (lane & (uint32_t)J) != 0;
    const uint64_t ys = h ? (y >> J) : (y << J);
    const uint64_t keep = h ? ~m : m;
    return (x & keep) | (ys & ~keep);
}

// 64x64 bit-matrix transpose across the wave: lane l holds row l on entry,
// column l on exit (bit k of the result = bit l of row k).  Branch-free.
__device__ __forceinline__ uint64_t wave_transpose(uint64_t x, uint32_t lane) {
    // Distance 32: the lower half of the wave trades its high words for the
    // upper half's low words, one v_permlane32_swap on the two 32-bit halves.
    {
        const auto p = __builtin_amdgcn_permlane32_swap((uint32_t)x, (uint32_t)(x >> 32), false,
                                                        false);
        x = (uint64_t)p[0] | ((uint64_t)p[1] << 32);
    }
    x = transpose_stage<16>(x, 0x0000FFFF0000FFFFull, lane);
    x = transpose_stage<8>(x, 0x00FF00FF00FF00FFull, lane);
    x = transpose_stage<4>(x, 0x0F0F0F0F0F0F0F0Full, lane);
    x = transpose_stage<2>(x, 0x3333333333333333ull, lane);
    x = transpose_stage<1>(x, 0x5555555555555555ull, lane);
    return x;
}

// base[i] as a global-memory access with a 32-bit byte offset (i * sizeof(T)
// < 2^32): pointers that come from a descriptor in memory are generic, and a
// generic access is a flat instruction that needs a 64-bit address per lane
// and waits on the LDS counter too; a global one takes the uniform base in
// SGPRs and one offset VGPR.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ T gld(const T* base, uint32_t i) {
    // C-style casts: the address-space conversion has no C++ cast
    return *(gptr<const T>)((gptr<const char>)base + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void gst(T* base, uint32_t i, T v) {
    *(gptr<T>)((gptr<char>)base + i * (uint32_t)sizeof(T)) = v;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t l) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// Acceptance u < metropolis_threshold(beta, de) without a table.  The
// threshold is first bracketed: a = 2^32 2^x with x = -beta de log2(e) in
// double precision (clamped to [-200, 0]), 2^frac(x) on v_exp_f32 and
// 2^floor(x) exact, so |a / T - 1| <= 3e-7 for T = 2^32 exp(-beta de), while
// the deterministic threshold lies within T (1 +- 1e-14) + 1.  Outside the
// band [a (1 - 2e-6) - 4, a (1 + 2e-6) + 4] the comparison is decided by a;
// inside it (probability below 4e-6) the threshold itself is evaluated.
// nbd = -beta log2(e).
__device__ __forceinline__ void threshold_band(int32_t de, double nbd, uint32_t& lo_u,
                                               uint32_t& hi_u) {
    const double x = fmin(fmax(nbd * (double)de, -200.0), 0.0);
    const double xi = floor(x);
    const float a = __builtin_ldexpf(__builtin_amdgcn_exp2f((float)(x - xi)), (int)xi + 32);
    const float lo = a * 0.999998f - 4.0f;
    const float hi = a * 1.000002f + 4.0f;
    lo_u = (uint32_t)fmaxf(lo, 0.0f);
    hi_u = hi < 4294967040.0f ? (uint32_t)hi : 0xFFFFFFFFu;
}

__device__ __forceinline__ bool band_accept(uint32_t u, int32_t de, double nbd, double beta) {
    uint32_t lo_u, hi_u;
    threshold_band(de, nbd, lo_u, hi_u);
    if (u < lo_u) return true;
    if (u > hi_u) return false;
    return u < metropolis_threshold(beta, (int64_t)de);
}

// Updates positions [p0, p1) of one slot; p0..p1-1 lie in one colour class
// and in one aligned group of four (p0 >> 2 == (p1 - 1) >> 2).
// S/W: the slot's words for reading / writing (same memory); read with
// vector loads, so words written by other waves before a barrier are seen.
// Metropolis (kHeatBath false): accept(u, de) returns
// u < metropolis_threshold(beta of this lane, de) for de >= 1; de <= 0 is
// accepted outright.  Heat-bath (kHeatBath true): accept(u, de, x) decides the
// flip itself from the uniform, dE and the current spin x.
// Adds the accepted dE to eacc (per lane) and the accepted count to nacc
// (wave-uniform).
template <bool kHeatBath = false, class Acc>
__device__ __forceinline__ void generic_group(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin, const uint64_t* S,
    uint64_t* W, uint32_t p0, uint32_t p1, uint32_t slot, uint32_t sweep, uint32_t key0,
    uint32_t key1, uint32_t lane, Acc&& accept, int64_t& eacc, uint32_t& nacc) {
    const U32x4 u = philox4x32_10(p0 >> 2, slot, sweep, (kTagWord << 24) | lane, key0, key1);
    const uint32_t cnt = p1 - p0;  // 1..4
    // row bounds of the group's positions (wave-uniform: declared so, so that
    // the entry loops below run on scalar counters)
    uint32_t rb[5];
#pragma unroll
    for (int q = 0; q < 5; ++q)
        rb[q] = __builtin_amdgcn_readfirstlane(gld(rp, p0 + min((uint32_t)q, cnt)));
    const uint64_t xw = lane < cnt ? gld(S, p0 + lane) : 0ull;
    int32_t h[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) h[q] = (uint32_t)q < cnt ? gld(lin, p0 + q) : 0;
    for (uint32_t base = rb[0]; base < rb[4]; base += 64) {
        const uint32_t e = base + lane;
        uint64_t nw = 0;
        int32_t nwt = 0;
        if (e < rb[4]) {
            nwt = gld(wt, e);
            nw = gld(S, gld(ci, e));
        }
        const uint32_t top = min(base + 64, rb[4]);
        // h += (bit ? w : 0) per entry, as a select of h + w: an empty
        // statement keeps each sum opaque, since the folded form needs w
        // copied into a VGPR first (one more VALU instruction per entry).
        // Two entries a pass, into two sums: half the loop's scalar
        // bookkeeping, and the second entry's readlanes fill the wait the
        // first one's select has on its lane mask.
        auto add = [&](int32_t& acc, uint32_t k) {
            const uint64_t word = readlane64(nw, k - base);
            const int32_t w = __builtin_amdgcn_readlane(nwt, k - base);
            int32_t hw = acc + w;
            asm volatile("" : "+v"(hw));
            acc = lane_bit(word) ? hw : acc;
        };
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t a = max(rb[q], base), b = min(rb[q + 1], top);
            int32_t h2 = 0;
            uint32_t k = a;
            for (; k + 1 < b; k += 2) {
                add(h[q], k);
                add(h2, k + 1);
            }
            if (k < b) add(h[q], k);
            h[q] += h2;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if ((uint32_t)q >= cnt) break;
        const uint64_t x = readlane64(xw, q);
        const int32_t de = lane_bit(x) ? -h[q] : h[q];
        const uint32_t wsel = (p0 + q) & 3u;
        const uint32_t uu = wsel == 0 ? u.v[0] : wsel == 1 ? u.v[1] : wsel == 2 ? u.v[2] : u.v[3];
        bool acc;
        if constexpr (kHeatBath) {
            acc = accept(uu, de, lane_bit(x));
        } else {
            acc = de <= 0;
            if (!acc) acc = accept(uu, de);
        }
        const uint64_t m = wave_ballot(acc);
        if (lane == 0) gst(W, p0 + q, x ^ m);
        eacc += acc ? de : 0;
        nacc += (uint32_t)__builtin_popcountll(m);
    }
}

// Adds to h[q] the entries [max(rb[q], base), min(rb[q+1], top)) of row q,
// whose couplings and neighbour words lane (k - base) holds.
template <int G>
__device__ __forceinline__ void walk_entries(const uint32_t (&rb)[G + 1], uint32_t base,
                                             uint32_t top, uint64_t nw, int32_t nwt,
                                             int32_t (&h)[G]) {
    // as generic_group's walk: selects of opaque sums, two entries a pass
    auto add = [&](int32_t& acc, uint32_t k) {
        const uint64_t word = readlane64(nw, k - base);
        const int32_t w = __builtin_amdgcn_readlane(nwt, k - base);
        int32_t hw = acc + w;
        asm volatile("" : "+v"(hw));
        acc = lane_bit(word) ? hw : acc;
    };
#pragma unroll
    for (int q = 0; q < G; ++q) {
        const uint32_t a = max(rb[q], base), b = min(rb[q + 1], top);
        int32_t h2 = 0;
        uint32_t k = a;
        for (; k + 1 < b; k += 2) {
            add(h[q], k);
            add(h2, k + 1);
        }
        if (k < b) add(h[q], k);
        h[q] += h2;
    }
}

// Local fields of positions [p0, p1) of one slot, lane = replica, K groups of
// G consecutive positions at a time.  The K groups' CSR entries are gathered
// together (lane e: coupling and neighbour word of a group's entry e, the
// first 64 entries of each group; longer groups fetch the rest in place), so
// K independent load chains are in flight before any is walked.  For every
// position p, visit(p, x, h, aux) receives its word x (wave-uniform), its
// field h = lin_p + sum_j q_pj x_j (per lane) and aux[p] (0 if aux is NULL).
template <int G, int K, class Visit>
__device__ __forceinline__ void fields_groups(const uint32_t* __restrict__ rp,
                                              const uint32_t* __restrict__ ci,
                                              const int32_t* __restrict__ wt,
                                              const int32_t* __restrict__ lin,
                                              const uint32_t* __restrict__ aux,
                                              const uint64_t* S, uint32_t p0, uint32_t p1,
                                              uint32_t lane, Visit&& visit) {
    for (uint32_t g0 = p0; g0 < p1; g0 += G * K) {
        uint32_t rb[K][G + 1], cnt[K];
        uint64_t xw[K], nw[K];
        int32_t hl[K], nwt[K];
        uint32_t ax[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t g = min(p1, g0 + k * G);
            cnt[k] = min((uint32_t)G, p1 - g);
#pragma unroll
            for (int q = 0; q <= G; ++q) rb[k][q] = rp[g + min((uint32_t)q, cnt[k])];
            xw[k] = 0;
            hl[k] = 0;
            ax[k] = 0;
            if (lane < cnt[k]) {
                xw[k] = S[g + lane];
                hl[k] = lin[g + lane];
                if (aux) ax[k] = aux[g + lane];
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t e = rb[k][0] + lane;
            nw[k] = 0;
            nwt[k] = 0;
            if (e < rb[k][G]) {
                nwt[k] = wt[e];
                nw[k] = S[ci[e]];
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            int32_t h[G];
#pragma unroll
            for (int q = 0; q < G; ++q) h[q] = __builtin_amdgcn_readlane(hl[k], q);
            walk_entries<G>(rb[k], rb[k][0], min(rb[k][0] + 64, rb[k][G]), nw[k], nwt[k], h);
            for (uint32_t base = rb[k][0] + 64; base < rb[k][G]; base += 64) {
                const uint32_t e = base + lane;
                uint64_t w2 = 0;
                int32_t t2 = 0;
                if (e < rb[k][G]) {
                    t2 = wt[e];
                    w2 = S[ci[e]];
                }
                walk_entries<G>(rb[k], base, min(base + 64, rb[k][G]), w2, t2, h);
            }
            const uint32_t g = g0 + k * G;
#pragma unroll
            for (int q = 0; q < G; ++q) {
                if ((uint32_t)q >= cnt[k]) continue;
                visit(g + q, readlane64(xw[k], q), h[q], __builtin_amdgcn_readlane(ax[k], q));
            }
        }
    }
}

// Positions [p0, p1) of one colour class, group by group.
template <bool kHeatBath = false, class Acc>
__device__ __forceinline__ void generic_positions(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin, const uint64_t* S,
    uint64_t* W, uint32_t p0, uint32_t p1, uint32_t slot, uint32_t sweep, uint32_t key0,
    uint32_t key1, uint32_t lane, Acc&& accept, int64_t& eacc, uint32_t& nacc) {
    for (uint32_t p = p0; p < p1;) {
        const uint32_t q1 = min(p1, (p & ~3u) + 4u);
        generic_group<kHeatBath>(rp, ci, wt, lin, S, W, p, q1, slot, sweep, key0, key1, lane,
                                 accept, eacc, nacc);
        p = q1;
    }
}

}  // namespace ige
