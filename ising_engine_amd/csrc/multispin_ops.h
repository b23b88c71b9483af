// multispin_ops.h -- the bit-sliced pieces of the multispin sweep (lane =
// site, one u64 word = 64 replicas), shared by the colour-class launch
// (sweep_multispin_kernel, kernels.hip) and the resident anneal
// (resident.hip).
//
// Model restriction (checked on the host): |q_ij| = w0 for every bond and
// lin_i = -1/2 sum_j q_ij, so dE_i(replica r) = w0 (k_r - d_i / 2) with
// k_r = #{ j : bit r of (x_i ^ x_j ^ [q_ij < 0]) }.  Replicas with dE <= 0
// are accepted outright; the rest compare their uniform u_r with the
// threshold thr_r on its kEagerPlanes leading bits (bit planes from Philox,
// every lane in lockstep), and the replicas whose leading bits tie are
// resolved later from a tail word.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ige_math.h"
#include "kernels.hpp"
#include "sweep_generic.h"

namespace ige {
namespace msc {

constexpr uint32_t kWave = 64;
constexpr uint32_t kTagTail = 5u;
constexpr int kSW = 6;  // planes of the signed energy counter of regular lattices

// Bitwise functions of three 64-bit words, one v_bitop3_b32 per 32-bit half.
// The truth-table byte of f(a, b, c) is f(0xF0, 0xCC, 0xAA): bit i of the
// table is f of bits (i >> 2) & 1, (i >> 1) & 1 and i & 1 of the operands.
constexpr uint32_t kXor3 = (0xF0 ^ 0xCC ^ 0xAA) & 0xFF;                         // a ^ b ^ c
constexpr uint32_t kMaj = ((0xF0 & 0xCC) | (0xF0 & 0xAA) | (0xCC & 0xAA)) & 0xFF;  // majority
constexpr uint32_t kAndOr = ((0xF0 & 0xCC) | 0xAA) & 0xFF;                      // (a & b) | c
constexpr uint32_t kAndAndNot = (0xF0 & 0xCC & ~0xAA) & 0xFF;                   // a & b & ~c
constexpr uint32_t kAndXnor = (0xF0 & ~(0xCC ^ 0xAA)) & 0xFF;                   // a & ~(b ^ c)
constexpr uint32_t kNotOr = (~0xF0 | 0xCC) & 0xFF;                              // ~a | b
constexpr uint32_t kNotAnd = (~0xF0 & 0xCC) & 0xFF;                             // ~a & b
constexpr uint32_t kXorAnd = ((0xF0 ^ 0xCC) & 0xAA) & 0xFF;                     // (a ^ b) & c
constexpr uint32_t kXnorAnd = (~(0xF0 ^ 0xCC) & 0xAA) & 0xFF;                   // ~(a ^ b) & c
constexpr uint32_t kMux = ((0xF0 & 0xCC) | (~0xF0 & 0xAA)) & 0xFF;              // a ? b : c
constexpr uint32_t kAnd3 = (0xF0 & 0xCC & 0xAA) & 0xFF;                         // a & b & c
// One step of a least-significant-first comparison u < t: with a = t's bit,
// b = u's bit and c = the verdict of the lower bits, (a ^ b) ? a : c.
constexpr uint32_t kLtStep = ((0xF0 & ~0xCC) | (~(0xF0 ^ 0xCC) & 0xAA)) & 0xFF;
static_assert(kXor3 == 0x96 && kMaj == 0xE8 && kMux == 0xCA && kLtStep == 0xB2, "truth tables");

template <uint32_t TT>
__device__ __forceinline__ uint64_t bop3(uint64_t a, uint64_t b, uint64_t c) {
    const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, (uint32_t)c, TT);
    const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(b >> 32),
                                                    (uint32_t)(c >> 32), TT);
    return ((uint64_t)hi << 32) | lo;
}

// As bop3 with the same 32-bit c in both halves.
template <uint32_t TT>
__device__ __forceinline__ uint64_t bop3s(uint64_t a, uint64_t b, uint32_t c) {
    const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, c, TT);
    const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(b >> 32), c, TT);
    return ((uint64_t)hi << 32) | lo;
}

// base[i] with a 32-bit byte offset (the caller guarantees i * sizeof(T) <
// 2^32), so a global load takes the uniform base in SGPRs and one offset VGPR
// instead of a 64-bit address per lane.
template <typename T>
__device__ __forceinline__ T ld32(const T* base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                       i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st32(T* base, uint32_t i, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(T)) = v;
}

// All ones if bit e of sg is set, else zero.
__device__ __forceinline__ uint32_t bit_mask(uint32_t sg, uint32_t e) {
    return (uint32_t)((int32_t)(sg << (31u - e)) >> 31);
}

// Sum over the wave's lanes of the per-lane bit-sliced counter acc[0..W-1],
// returned in lane r for replica r.
template <int W>
__device__ __forceinline__ int64_t wave_column_sum(const uint64_t* acc, uint32_t lane) {
    int64_t tot = 0;
#pragma unroll
    for (int b = 0; b < W; ++b)
        tot += (int64_t)__builtin_popcountll(wave_transpose(acc[b], lane)) << b;
    return tot;
}

// As wave_column_sum for W-plane two's-complement counters (plane W-1 weighs
// -2^(W-1)).
template <int W>
__device__ __forceinline__ int64_t wave_column_sum_signed(const uint64_t* acc, uint32_t lane) {
    int64_t tot = 0;
#pragma unroll
    for (int b = 0; b < W; ++b) {
        const int64_t cnt = (int64_t)__builtin_popcountll(wave_transpose(acc[b], lane)) << b;
        tot += b == W - 1 ? -cnt : cnt;
    }
    return tot;
}

// Resolves one tied replica: u_r's kEagerPlanes leading bits tie with thr_r's,
// so u_r < thr_r iff tail >> kEagerPlanes < thr_r's low bits.
__device__ __forceinline__ bool resolve_tail(uint32_t p, uint32_t slot, uint32_t sweep,
                                             uint32_t key0, uint32_t key1, uint32_t r,
                                             uint32_t thr) {
    const U32x4 o = philox4x32_10(p, slot, sweep, (kTagTail << 24) | (r >> 2), key0, key1);
    const uint32_t q = r & 3u;
    const uint32_t w = q == 0 ? o.v[0] : q == 1 ? o.v[1] : q == 2 ? o.v[2] : o.v[3];
    return (w >> kEagerPlanes) < (thr & ((1u << (32 - kEagerPlanes)) - 1u));
}

// Threshold of replica r at level m from a slot's threshold records (kQStride
// u64 per level: the kEagerPlanes leading bit planes, the nonzero mask, then
// the 64 thresholds two per word).
__device__ __forceinline__ uint32_t record_thr(const unsigned long long* recs, uint32_t m,
                                               uint32_t r) {
    const unsigned long long w = recs[m * kQStride + kEagerPlanes + 1 + (r >> 1)];
    return (r & 1u) ? (uint32_t)(w >> 32) : (uint32_t)w;
}

// Builds a slot's threshold records in LDS for levels m = m0, m0 + step, ...,
// <= dmax: lane r reads thr(temperature of lane r, m) from lev_thr ([cells][dmax
// + 1]) and nine ballots make the planes.  Every lane of the wave takes part.
// The reads of eight levels are issued before their ballots.
__device__ __forceinline__ void build_records(unsigned long long* recs, const uint32_t* lev_thr,
                                              uint32_t tr, uint32_t dmax, uint32_t m0,
                                              uint32_t step, uint32_t lane) {
    constexpr uint32_t kBatch = 8;
    for (uint32_t m1 = m0; m1 <= dmax; m1 += kBatch * step) {
        uint32_t th[kBatch];
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) {
            const uint32_t m = m1 + i * step;
            th[i] = m <= dmax ? lev_thr[tr * (dmax + 1) + m] : 0u;
        }
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) {
            const uint32_t m = m1 + i * step;
            if (m > dmax) break;  // uniform
            unsigned long long* out = recs + m * kQStride;
            uint64_t mine = 0;
#pragma unroll
            for (uint32_t b = 0; b < kEagerPlanes; ++b) {
                const uint64_t plane = wave_ballot((th[i] >> (31u - b)) & 1u);
                if (lane == b) mine = plane;
            }
            const uint64_t nz = wave_ballot(th[i] != 0);
            if (lane < kEagerPlanes) out[lane] = mine;
            if (lane == kEagerPlanes) out[kEagerPlanes] = nz;
            reinterpret_cast<uint32_t*>(out + kEagerPlanes + 1)[lane] = th[i];
        }
    }
}

// One wave's tie records on a regular lattice: per site with tied replicas,
// the tied mask, its subsets at levels 1 and 2 (level j has m = 2 + 2j, dE =
// w0 m / 2) and the position.  All in LDS; *count is the wave's fill.
struct TieRecords {
    unsigned long long* und;
    unsigned long long* l1;
    unsigned long long* l2;  // degree 6 only
    uint32_t* pos;
    uint32_t* count;
    uint32_t cap;
};

// Resolves the tied replicas u of position p: flips the accepted ones in Wr
// (atomic XOR: other lanes own other words of the same wave) and adds their
// dE to edef[r].
template <int DEG>
__device__ __forceinline__ void resolve_site_ties(uint64_t u, uint64_t l1, uint64_t l2, uint32_t p,
                                                  uint64_t* Wr, const unsigned long long* recs,
                                                  uint32_t slot, uint32_t sweep, uint32_t key0,
                                                  uint32_t key1, int32_t w0,
                                                  unsigned long long* edef, uint32_t& nacc) {
    while (u) {
        const uint32_t r = (uint32_t)__builtin_ctzll(u);
        u &= u - 1;
        const uint32_t m = ((l1 >> r) & 1u) ? 4u : (DEG == 6 && ((l2 >> r) & 1u)) ? 6u : 2u;
        if (resolve_tail(p, slot, sweep, key0, key1, r, record_thr(recs, m, r))) {
            atomicXor(reinterpret_cast<unsigned long long*>(Wr + p), 1ull << r);
            atomicAdd(&edef[r], (unsigned long long)(((int64_t)m * w0) / 2));
            nacc += 1;
        }
    }
}

// Drains a wave's tie records with all 64 lanes.
template <int DEG>
__device__ __forceinline__ void drain_ties(const TieRecords& tq, uint64_t* Wr,
                                           const unsigned long long* recs, uint32_t slot,
                                           uint32_t sweep, uint32_t key0, uint32_t key1,
                                           int32_t w0, unsigned long long* edef, uint32_t lane,
                                           uint32_t& nacc) {
    const uint32_t queued = min(*tq.count, tq.cap);
    for (uint32_t q = lane; q < queued; q += kWave)
        resolve_site_ties<DEG>(tq.und[q], tq.l1[q], DEG == 6 ? tq.l2[q] : 0ull, tq.pos[q], Wr,
                               recs, slot, sweep, key0, key1, w0, edef, nacc);
}

// What a regular-lattice update reads before its neighbour words: the word
// at p, the coupling sign bits and the DEG neighbour positions.  A caller
// loads the next site's row while it updates the current one, so the index
// loads' latency is spent on the current site's arithmetic.
template <int DEG>
struct SiteRow {
    uint64_t x;
    uint32_t sg;
    uint32_t j[DEG];
};

template <int DEG>
__device__ __forceinline__ SiteRow<DEG> load_row(const uint64_t* S, uint32_t p,
                                                 const uint32_t* ci, const uint32_t* sgn) {
    // word and entry byte offsets fit 32 bits (multispin_shape: n < 2^29,
    // nnz < 2^30)
    SiteRow<DEG> r;
    r.x = ld32(S, p);
    r.sg = ld32(sgn, p);
    const uint32_t e0 = p * (uint32_t)DEG;
#pragma unroll
    for (int e = 0; e < DEG; ++e) r.j[e] = ld32(ci, e0 + (uint32_t)e);
    return r;
}

// The update of position p on a regular lattice (degree DEG in {4, 6}, row p
// at p * DEG, count planes B = 3): reads the neighbours of row from S,
// writes the new word to Wr, adds (k - DEG/2) * flip to the signed counter
// sacc and the flip count to nacc, and records the tied replicas in tq (or,
// when tq is full, resolves them in place).
template <int DEG>
__device__ __forceinline__ void regular_site(const uint64_t* S, uint64_t* Wr, uint32_t p,
                                             const SiteRow<DEG>& row,
                                             const unsigned long long* recs, uint32_t slot,
                                             uint32_t sweep, uint32_t key0, uint32_t key1,
                                             int32_t w0, uint64_t (&sacc)[kSW], uint32_t& nacc,
                                             const TieRecords& tq, unsigned long long* edef) {
    static_assert(DEG == 4 || DEG == 6, "regular lattices of degree 4 or 6");
    const uint64_t x = row.x;
    const uint32_t sg = row.sg;
    auto bond = [&](uint32_t e) { return bop3s<kXor3>(x, ld32(S, row.j[e]), bit_mask(sg, e)); };
    // carry-save adder tree: c = popcount of the bond words, bit-sliced
    auto fa = [](uint64_t a, uint64_t b, uint64_t cin, uint64_t& sum, uint64_t& cout) {
        sum = bop3<kXor3>(a, b, cin);
        cout = bop3<kMaj>(a, b, cin);
    };
    uint64_t c[3];
    uint64_t s1, c1;
    fa(bond(0), bond(1), bond(2), s1, c1);
    if constexpr (DEG == 6) {
        uint64_t s2, c2;
        fa(bond(3), bond(4), bond(5), s2, c2);
        c[0] = s1 ^ s2;
        fa(c1, c2, s1 & s2, c[1], c[2]);
    } else {
        const uint64_t k3 = bond(3);
        c[0] = s1 ^ k3;
        const uint64_t t = s1 & k3;
        c[1] = c1 ^ t;
        c[2] = c1 & t;
    }
    // dE > 0 for k > DEG/2, at level m = 2k - DEG.  Degree 6 (k <= 6): pos =
    // c2, and k = 4, 5, 6 (c = 100, 101, 110) is level 2, 4, 6 = c0 ? 4 : c1 ?
    // 6 : 2.  Degree 4 (k <= 4): pos = c2 | c1 c0, and k = 3, 4 (011, 100) is
    // level c2 ? 4 : 2.  Outside pos the selected record is arbitrary: those
    // replicas are accepted whatever the comparison says.
    const uint64_t pos = DEG == 6 ? c[2] : bop3<kAndOr>(c[1], c[0], c[2]);
    auto level_row = [&](uint32_t b) {
        const unsigned long long* r = recs + b;
        if constexpr (DEG == 6)
            return bop3<kMux>(c[0], r[4 * kQStride], bop3<kMux>(c[1], r[6 * kQStride],
                                                                r[2 * kQStride]));
        else
            return bop3<kMux>(c[2], r[4 * kQStride], r[2 * kQStride]);
    };
    // kEagerPlanes lockstep bit planes, two per Philox call, compared least
    // significant first: lt = u < thr on the leading bits (0 where they tie),
    // eq = they tie.
    uint64_t lt = 0, eq = 0;
#pragma unroll
    for (int j = kEagerPlanes / 2 - 1; j >= 0; --j) {
        const U32x4 o = philox4x32_10(p, slot, sweep, (kTagPlane << 24) | (uint32_t)j, key0,
                                      key1);
#pragma unroll
        for (int h = 1; h >= 0; --h) {
            const uint32_t b = 2 * (uint32_t)j + (uint32_t)h;
            const uint64_t P = h == 0 ? ((uint64_t)o.v[0] | ((uint64_t)o.v[1] << 32))
                                      : ((uint64_t)o.v[2] | ((uint64_t)o.v[3] << 32));
            const uint64_t T = level_row(b);
            if (b == kEagerPlanes - 1) {
                lt = T & ~P;
                eq = ~(T ^ P);
            } else {
                lt = bop3<kLtStep>(T, P, lt);
                eq = bop3<kAndXnor>(eq, T, P);
            }
        }
    }
    // ties of replicas with dE > 0 and a nonzero threshold go to the tail word
    const uint64_t und = bop3<kAnd3>(eq, pos, level_row(kEagerPlanes));
    const uint64_t flip = bop3<kNotOr>(pos, lt, 0ull);
    st32(Wr, p, x ^ flip);
    nacc += (uint32_t)__builtin_popcountll(flip);
    // (k - DEG/2) * flip as a three-plane two's-complement addend:
    // k + (8 - DEG/2) mod 8, i.e. + 101b for DEG 6 and + 110b for DEG 4.
    uint64_t a0, a1, a2;
    if constexpr (DEG == 6) {
        a0 = bop3<kNotAnd>(c[0], flip, 0ull);          // ~c0, carry c0
        a1 = bop3<kXorAnd>(c[1], c[0], flip);          // c1 ^ c0, carry c1 c0
        a2 = bop3<kXnorAnd>(c[2], c[1] & c[0], flip);  // c2 + 1 + carry
    } else {
        a0 = c[0] & flip;                              // c0 + 0
        a1 = bop3<kNotAnd>(c[1], flip, 0ull);          // ~c1, carry c1
        a2 = bop3<kXnorAnd>(c[2], c[1], flip);         // c2 + 1 + carry
    }
    // sacc += a, a sign-extended from plane 2 (ripple add)
    uint64_t cs = sacc[0] & a0;
    sacc[0] ^= a0;
#pragma unroll
    for (int b = 1; b < kSW; ++b) {
        const uint64_t ab = b == 1 ? a1 : a2;
        const uint64_t sb = bop3<kXor3>(sacc[b], ab, cs);
        if (b + 1 < kSW) cs = bop3<kMaj>(sacc[b], ab, cs);
        sacc[b] = sb;
    }
    // Record the replicas whose leading bits tie.
    if (und) {
        const uint32_t at = atomicAdd(tq.count, 1u);
        // level 4 is k = 5 (c0 set) on degree 6, k = 4 (c2 set) on degree 4;
        // level 6 is k = 6 (c1 set)
        const uint64_t l1 = und & (DEG == 6 ? c[0] : c[2]);
        const uint64_t l2 = DEG == 6 ? (und & c[1]) : 0ull;
        if (at < tq.cap) {
            tq.und[at] = und;
            tq.l1[at] = l1;
            if constexpr (DEG == 6) tq.l2[at] = l2;
            tq.pos[at] = p;
        } else {
            __threadfence_block();
            resolve_site_ties<DEG>(und, l1, l2, p, Wr, recs, slot, sweep, key0, key1, w0, edef,
                                   nacc);
        }
    }
}

// The update of position p for any degree d < 2^B (row p at rp[p]): count
// planes from the row in groups, a level per k > d/2, the same lockstep
// planes as regular_site, unsigned counters kacc += k flip and dacc += d flip,
// and one queue entry per tied replica, (qoff << 11) | (r << 5) | m with qoff
// = p - (the queue's base position) < 2^21, resolved in place when the queue
// (qcap entries) is full.
template <int B, int W>
__device__ __forceinline__ void irregular_site(
    const uint64_t* S, uint64_t* Wr, uint32_t p, uint32_t qoff, const uint32_t* rp,
    const uint32_t* ci, const uint32_t* sgn, const unsigned long long* recs, uint32_t slot,
    uint32_t sweep, uint32_t key0, uint32_t key1, int32_t w0, uint64_t (&kacc)[W],
    uint64_t (&dacc)[W], uint32_t& nacc, uint32_t* queue, uint32_t* qcount, uint32_t qcap,
    unsigned long long* edef) {
    constexpr int LMAX = 1 << (B - 1);
    const uint64_t x = ld32(S, p);
    const uint32_t e0 = ld32(rp, p);
    const uint32_t d = ld32(rp, p + 1) - e0;
    const uint32_t sg = ld32(sgn, p);
    uint64_t c[B];
#pragma unroll
    for (int b = 0; b < B; ++b) c[b] = 0;
    // Rows in groups of kRowGroup entries: a group's index loads issue
    // together, then its word loads, so a row costs two memory round trips per
    // group rather than two per neighbour (the sweep of a small colour class
    // is one wave's walk along its rows).  Entries past the row's end repeat
    // its last one and are masked out of the count.
    constexpr uint32_t kRowGroup = 8;
    for (uint32_t g = 0; g < d; g += kRowGroup) {
        uint32_t j[kRowGroup];
        uint64_t wj[kRowGroup];
#pragma unroll
        for (uint32_t q = 0; q < kRowGroup; ++q) j[q] = ld32(ci, e0 + min(g + q, d - 1));
#pragma unroll
        for (uint32_t q = 0; q < kRowGroup; ++q) wj[q] = ld32(S, j[q]);
#pragma unroll
        for (uint32_t q = 0; q < kRowGroup; ++q) {
            const uint32_t e = g + q;
            uint64_t carry =
                bop3s<kXor3>(x, wj[q], bit_mask(sg, e & 31u)) & (e < d ? ~0ull : 0ull);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const uint64_t s = c[b] ^ carry;
                carry &= c[b];
                c[b] = s;
            }
        }
    }
    // Level masks for k = floor(d/2) + 1 + j, k <= d; level m = 2k - d.
    uint64_t lev[LMAX];
    uint32_t mrow[LMAX];
    // Most significant plane first (a least-significant-first comparison,
    // as on regular lattices, needs more registers here and costs the
    // degree < 32 kernel an occupancy step).
    uint64_t pos = 0, und = 0;
#pragma unroll
    for (int j = 0; j < LMAX; ++j) {
        const uint32_t k = (d >> 1) + 1 + (uint32_t)j;
        uint64_t eq = k <= d ? ~0ull : 0ull;
#pragma unroll
        for (int b = 0; b < B; ++b) eq &= ((k >> b) & 1u) ? c[b] : ~c[b];
        lev[j] = eq;
        mrow[j] = (k <= d ? 2 * k - d : 0) * kQStride;
        pos |= eq;
        und = bop3<kAndOr>(eq, recs[mrow[j] + kEagerPlanes], und);
    }
    // kEagerPlanes lockstep bit planes, two per Philox call.
    uint64_t lt = 0;
#pragma unroll
    for (uint32_t j = 0; j < kEagerPlanes / 2; ++j) {
        const U32x4 o = philox4x32_10(p, slot, sweep, (kTagPlane << 24) | j, key0, key1);
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h) {
            const uint32_t b = 2 * j + h;
            const uint64_t P = h == 0 ? ((uint64_t)o.v[0] | ((uint64_t)o.v[1] << 32))
                                      : ((uint64_t)o.v[2] | ((uint64_t)o.v[3] << 32));
            uint64_t T = lev[0] & recs[mrow[0] + b];
#pragma unroll
            for (int l = 1; l < LMAX; ++l) T = bop3<kAndOr>(lev[l], recs[mrow[l] + b], T);
            lt |= bop3<kAndAndNot>(und, T, P);
            und = bop3<kAndXnor>(und, P, T);
        }
    }
    const uint64_t flip = bop3<kNotOr>(pos, lt, 0ull);
    st32(Wr, p, x ^ flip);
    nacc += (uint32_t)__builtin_popcountll(flip);
    // kacc += k * flip, dacc += d * flip (bit-sliced ripple adds).
    uint64_t ck = 0, cd = 0;
#pragma unroll
    for (int b = 0; b < W; ++b) {
        if (b == 0) {  // half adder: no carry in
            const uint64_t ak = c[0] & flip;
            ck = kacc[0] & ak;
            kacc[0] ^= ak;
        } else if (b < B) {  // full adder
            const uint64_t ak = c[b] & flip;
            const uint64_t sk = bop3<kXor3>(kacc[b], ak, ck);
            ck = bop3<kMaj>(kacc[b], ak, ck);
            kacc[b] = sk;
        } else {  // carry propagation only
            const uint64_t sk = kacc[b] ^ ck;
            ck &= kacc[b];
            kacc[b] = sk;
        }
        const uint64_t ad = (b < B && ((d >> b) & 1u)) ? flip : 0ull;
        const uint64_t sd = bop3<kXor3>(dacc[b], ad, cd);
        cd = bop3<kMaj>(dacc[b], ad, cd);
        dacc[b] = sd;
    }
    // Queue the replicas whose leading bits tie.
    while (und) {
        const uint32_t r = (uint32_t)__builtin_ctzll(und);
        und &= und - 1;
        uint32_t m = 0;
#pragma unroll
        for (int l = 0; l < LMAX; ++l)
            if ((lev[l] >> r) & 1u) m = mrow[l] / kQStride;
        const uint32_t at = atomicAdd(qcount, 1u);
        if (at < qcap) {
            queue[at] = (qoff << 11) | (r << 5) | m;
        } else if (resolve_tail(p, slot, sweep, key0, key1, r, record_thr(recs, m, r))) {
            // queue full: resolve in place
            __threadfence_block();
            atomicXor(reinterpret_cast<unsigned long long*>(Wr + p), 1ull << r);
            atomicAdd(&edef[r], (unsigned long long)(((int64_t)m * w0) / 2));
            nacc += 1;
        }
    }
}

// Largest value of v over the wave's lanes (lanes that left early may add
// their stale values, which can only make the result larger).
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    return v;
}

// Half h of word i of S (replicas 32 h .. 32 h + 31), read device-coherently:
// the load goes to the L2, not the CU's L1, so it sees words other
// workgroups wrote since this CU last read them, with no cache invalidation.
__device__ __forceinline__ uint32_t ld_half(const uint64_t* S, uint32_t i, uint32_t h) {
    return __hip_atomic_load(
        reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(const_cast<uint64_t*>(S)) + i * 8u +
                                    h * 4u),
        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// v of the other lane of this lane's pair (lanes 2 i and 2 i + 1): DPP
// quad_perm [1, 0, 3, 2].  Both lanes of the pair must be active.
__device__ __forceinline__ uint32_t pair_swap(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
}

// As wave_column_sum for the 32-replica counters of irregular_site_pair
// (lane l holds replicas 32 (l & 1) ..).
template <int W>
__device__ __forceinline__ int64_t wave_column_sum_pair(const uint32_t* acc, uint32_t lane) {
    int64_t tot = 0;
#pragma unroll
    for (int b = 0; b < W; ++b)
        tot += (int64_t)__builtin_popcountll(
                   wave_transpose((uint64_t)acc[b] << (32u * (lane & 1u)), lane))
               << b;
    return tot;
}

// What irregular_site_pair reads of a position's row before any spin word:
// entry start, degree, coupling sign bits and the first kPairPre neighbour
// positions (entries past the degree repeat the last one; zero for degree
// 0; load_pair_row<B> fills pair_pre<B>() of them).  A caller loads it
// ahead, while it waits at a barrier.
constexpr uint32_t kPairPre = 16;
struct PairRow {
    uint32_t e0, d, sg;
    uint32_t j[kPairPre];
};

// Entries of a row read ahead for count planes B: 8 when every degree is
// below 8 (B = 3), else kPairPre.
template <int B>
constexpr uint32_t pair_pre() {
    return B == 3 ? kPairPre / 2 : kPairPre;
}

template <int B>
__device__ __forceinline__ PairRow load_pair_row(uint32_t p, const uint32_t* rp,
                                                 const uint32_t* ci, const uint32_t* sgn) {
    PairRow r;
    r.e0 = ld32(rp, p);
    r.d = ld32(rp, p + 1) - r.e0;
    r.sg = ld32(sgn, p);
#pragma unroll
    for (uint32_t q = 0; q < kPairPre; ++q) r.j[q] = 0;
    if (r.d) {
#pragma unroll
        for (uint32_t q = 0; q < pair_pre<B>(); ++q) r.j[q] = ld32(ci, r.e0 + min(q, r.d - 1));
    }
    return r;
}

// Full adder on bit planes: s = a ^ b ^ c, k = majority(a, b, c).
__device__ __forceinline__ void full_add(uint32_t a, uint32_t b, uint32_t c, uint32_t& s,
                                         uint32_t& k) {
    s = __builtin_amdgcn_bitop3_b32(a, b, c, kXor3);
    k = __builtin_amdgcn_bitop3_b32(a, b, c, kMaj);
}

// Carry-save trees: bit r of o[i] is bit i of the number of inputs with bit
// r set (8 inputs: 4 full and 3 half adders; 16: 11 and 4), where a ripple
// add per input costs two operations per plane.
__device__ __forceinline__ void count8(const uint32_t* b, uint32_t (&o)[4]) {
    uint32_t s0, s1, s2, k0, k1, k2, u, q;
    full_add(b[0], b[1], b[2], s0, k0);
    full_add(b[3], b[4], b[5], s1, k1);
    full_add(s0, s1, b[6], s2, k2);
    o[0] = s2 ^ b[7];
    const uint32_t k3 = s2 & b[7];
    full_add(k0, k1, k2, u, q);
    o[1] = u ^ k3;
    const uint32_t q1 = u & k3;
    o[2] = q ^ q1;
    o[3] = q & q1;
}

__device__ __forceinline__ void count16(const uint32_t* b, uint32_t (&o)[5]) {
    uint32_t s[5], k[8], t0, t1, u0, u1, u2, q[4], v, r0;
#pragma unroll
    for (int i = 0; i < 5; ++i) full_add(b[3 * i], b[3 * i + 1], b[3 * i + 2], s[i], k[i]);
    full_add(s[0], s[1], s[2], t0, k[5]);
    full_add(s[3], s[4], b[15], t1, k[6]);
    o[0] = t0 ^ t1;
    k[7] = t0 & t1;
    full_add(k[0], k[1], k[2], u0, q[0]);
    full_add(k[3], k[4], k[5], u1, q[1]);
    full_add(k[6], k[7], u0, u2, q[2]);
    o[1] = u1 ^ u2;
    q[3] = u1 & u2;
    full_add(q[0], q[1], q[2], v, r0);
    o[2] = v ^ q[3];
    const uint32_t r1 = v & q[3];
    o[3] = r0 ^ r1;
    o[4] = r0 & r1;
}

// The count planes of position p for replicas 32 h .. 32 h + 31 (h = lane &
// 1; bit r of c[b] = bit b of k_r), from the row prefetched in `row`; returns
// this lane's half of p's word.  The words of the first pair_pre<B>()
// entries load together before any is used, so a site of such a degree
// costs one round trip to the L2.
template <int B>
__device__ __forceinline__ uint32_t pair_counts(const uint64_t* S, uint32_t p, uint32_t h,
                                                const PairRow& row, const uint32_t* ci,
                                                uint32_t (&c)[B]) {
    static_assert(B >= 3 && B <= 5, "count planes");
    const uint32_t d = row.d, sg = row.sg;
    const uint32_t x = ld_half(S, p, h);
    // ripple add of entry e (past kPre)
    auto add = [&](uint32_t e, uint32_t w) {
        uint32_t carry = __builtin_amdgcn_bitop3_b32(x, w, bit_mask(sg, e & 31u), kXor3) &
                         (e < d ? ~0u : 0u);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const uint32_t s = c[b] ^ carry;
            carry &= c[b];
            c[b] = s;
        }
    };
    // The first kPre entries' words, each loaded only by the lanes whose
    // degree reaches it (a masked lane sends no request: the loads bypass
    // the L1, and the L2's request rate bounds a class of random gathers).
    // No branch on the wave's degrees: the compiler would hoist the first
    // loads and count them before issuing the rest.
    constexpr uint32_t kPre = pair_pre<B>(), kG = kPairPre / 2;
    uint32_t wj[kPre];
#pragma unroll
    for (uint32_t q = 0; q < kPre; ++q) {
        wj[q] = 0;
        if (q < d) wj[q] = ld_half(S, row.j[q], h);
    }
    // bit r of dis[q]: replica r disagrees with entry q's bond (x ^ w ^ sign),
    // zero past the degree
    uint32_t dis[kPre];
#pragma unroll
    for (uint32_t q = 0; q < kPre; ++q)
        dis[q] = __builtin_amdgcn_bitop3_b32(wj[q] ^ bit_mask(sg, q), x, q < d ? ~0u : 0u, kXorAnd);
    if constexpr (kPre == 16) {
        uint32_t o[5];
        count16(dis, o);
#pragma unroll
        for (int b = 0; b < B; ++b) c[b] = o[b];  // o[4] = 0 when B = 4 (d < 16)
    } else {
        uint32_t o[4];
        count8(dis, o);
#pragma unroll
        for (int b = 0; b < B; ++b) c[b] = o[b];  // B = 3: d < 8
    }
    // entries past kPre (degrees above 16), in groups of kG
    for (uint32_t g = kPre; g < d; g += kG) {
        uint32_t j[kG], wj[kG];
#pragma unroll
        for (uint32_t q = 0; q < kG; ++q) j[q] = ld32(ci, row.e0 + min(g + q, d - 1));
#pragma unroll
        for (uint32_t q = 0; q < kG; ++q) wj[q] = ld_half(S, j[q], h);
#pragma unroll
        for (uint32_t q = 0; q < kG; ++q) add(g + q, wj[q]);
    }
    return x;
}

// The rest of the update after pair_counts, on the first BF count planes
// (the wave's degrees are below 2^BF, so planes BF .. B - 1 are zero):
// returns the lane's flip mask (bit rb: replica 32 h + rb flips) and adds the
// flips to the counters.  Every level up to 2^(BF-1) is compared: skipping
// the levels past the wave's largest degree with uniform branches measured
// 14 % slower (the branches split the batches of record reads).  recs holds
// a record every RS u64 (the planes and the nonzero mask first);
// tie_thr(m, r) is replica r's full threshold at level m.
template <int BF, int B, int W, uint32_t RS = kQStride, class TieThr>
__device__ __forceinline__ uint32_t pair_update(
    const uint32_t (&c)[B], uint32_t p, uint32_t h, uint32_t d, const unsigned long long* recs,
    uint32_t slot, uint32_t sweep, uint32_t key0, uint32_t key1, uint32_t (&kacc)[W],
    uint32_t (&dacc)[W], uint32_t& nacc, TieThr&& tie_thr) {
    static_assert(kEagerPlanes == 8, "a pair's lanes compute two Philox calls each");
    constexpr int LMAX = 1 << (BF - 1);
    // recs as 32-bit halves: half h of record word i at rh[2 i + h].  Level m
    // = 2 k - d of k = floor(d/2) + 1 + j is 2 j + 2 - (d & 1): its record
    // starts 2 j RS words after rl's.  Levels of k > d (up to 2^BF) are read
    // but masked out by lev[j] = 0.
    const uint32_t m0 = 2u - (d & 1u);
    const uint32_t* rl = reinterpret_cast<const uint32_t*>(recs) + h + 2 * m0 * RS;
    uint32_t lev[LMAX];
    uint32_t pos = 0, und = 0;
#pragma unroll
    for (int j = 0; j < LMAX; ++j) {
        const uint32_t k = (d >> 1) + 1 + (uint32_t)j;
        uint32_t eq = k <= d ? ~0u : 0u;
#pragma unroll
        for (int b = 0; b < BF; ++b) eq &= ((k >> b) & 1u) ? c[b] : ~c[b];
        lev[j] = eq;
        pos |= eq;
        und = __builtin_amdgcn_bitop3_b32(eq, rl[2 * (2 * j * RS + kEagerPlanes)], und, kAndOr);
    }
    // Plane 2 j + t of call j: replicas 0..31 in word 2 t, 32..63 in 2 t + 1.
    uint32_t keep[4], send[4];
#pragma unroll
    for (uint32_t jj = 0; jj < 2; ++jj) {
        const U32x4 o = philox4x32_10(p, slot, sweep, (kTagPlane << 24) | (2 * h + jj), key0, key1);
#pragma unroll
        for (uint32_t t = 0; t < 2; ++t) {
            keep[2 * jj + t] = h ? o.v[2 * t + 1] : o.v[2 * t];
            send[2 * jj + t] = h ? o.v[2 * t] : o.v[2 * t + 1];
        }
    }
    uint32_t P[kEagerPlanes];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t recv = pair_swap(send[k]);
        P[k] = h ? recv : keep[k];
        P[4 + k] = h ? keep[k] : recv;
    }
    uint32_t lt = 0;
#pragma unroll
    for (uint32_t b = 0; b < kEagerPlanes; ++b) {
        uint32_t T = lev[0] & rl[2 * b];
#pragma unroll
        for (int l = 1; l < LMAX; ++l)
            T = __builtin_amdgcn_bitop3_b32(lev[l], rl[2 * (2 * l * RS + b)], T, kAndOr);
        lt |= __builtin_amdgcn_bitop3_b32(und, T, P[b], kAndAndNot);
        und = __builtin_amdgcn_bitop3_b32(und, P[b], T, kAndXnor);
    }
    // Tied replicas are resolved here from their tail words, before the word
    // is written, so their flips and energy changes go through the same
    // store and counters as the others.  A compared replica ties with
    // probability 1/256, so a lane of 32 replicas has one with probability
    // below 1/8 and a wave usually makes a single pass of the loop; a queue
    // drained after the class added its atomics' round trip to the latency
    // of every class.
    while (und) {
        const uint32_t rb = (uint32_t)__builtin_ctz(und);
        und &= und - 1;
        uint32_t m = 0;
#pragma unroll
        for (int l = 0; l < LMAX; ++l)
            if ((lev[l] >> rb) & 1u) m = m0 + 2 * l;
        const uint32_t r = 32 * h + rb;
        if (resolve_tail(p, slot, sweep, key0, key1, r, tie_thr(m, r))) lt |= 1u << rb;
    }
    const uint32_t flip = ~pos | lt;
    nacc += (uint32_t)__builtin_popcount(flip);
    uint32_t ck = 0, cd = 0;
#pragma unroll
    for (int b = 0; b < W; ++b) {
        if (b == 0) {
            const uint32_t ak = c[0] & flip;
            ck = kacc[0] & ak;
            kacc[0] ^= ak;
        } else if (b < BF) {
            const uint32_t ak = c[b] & flip;
            const uint32_t sk = __builtin_amdgcn_bitop3_b32(kacc[b], ak, ck, kXor3);
            ck = __builtin_amdgcn_bitop3_b32(kacc[b], ak, ck, kMaj);
            kacc[b] = sk;
        } else {
            const uint32_t sk = kacc[b] ^ ck;
            ck &= kacc[b];
            kacc[b] = sk;
        }
        const uint32_t ad = (b < BF && ((d >> b) & 1u)) ? flip : 0u;
        const uint32_t sd = __builtin_amdgcn_bitop3_b32(dacc[b], ad, cd, kXor3);
        cd = __builtin_amdgcn_bitop3_b32(dacc[b], ad, cd, kMaj);
        dacc[b] = sd;
    }
    return flip;
}

// irregular_site with a position's 64 replicas split over a lane pair: lane
// 2 i + h (h = lane & 1) updates replicas 32 h .. 32 h + 31 on 32-bit planes,
// so each lane does half the bit operations, and the pair shares the Philox
// calls of the position: lane h computes calls 2 h and 2 h + 1 (bit planes
// 4 h .. 4 h + 3 of all 64 replicas) and trades the other half's words with
// its partner.  The uniforms, thresholds and decisions are those of
// irregular_site (tied replicas resolved in place from the same tail words);
// kacc and dacc count this lane's 32 replicas (wave_column_sum_pair), tied
// flips included.  Both lanes of a pair take the same p, whose row
// is `row` (load_pair_row); dw is the largest degree of the wave (uniform).
// A site's update depends on its degree, counts, uniforms and thresholds,
// not on the planes' width, so when dw < 2^(B-1) the levels are compared on
// B - 1 planes, half as many (an ER graph of mean degree 6 and N = 20000 has
// a largest degree near 20, so B = 5, but almost every wave of 32 sites
// stays below 16).  Words are read device-coherently.  recs must hold
// levels 0 .. 2^B (those past the graph's largest degree are read, not
// used).
template <int B, int W>
__device__ __forceinline__ void irregular_site_pair(
    uint32_t dw, const uint64_t* S, uint64_t* Wr, uint32_t p, uint32_t h, const PairRow& row,
    const uint32_t* ci, const unsigned long long* recs, uint32_t slot, uint32_t sweep,
    uint32_t key0, uint32_t key1, uint32_t (&kacc)[W], uint32_t (&dacc)[W], uint32_t& nacc) {
    uint32_t c[B];
    const uint32_t x = pair_counts<B>(S, p, h, row, ci, c);
    auto thr = [&](uint32_t m, uint32_t r) { return record_thr(recs, m, r); };
    uint32_t flip;
    if (B > 3 && dw < (1u << (B - 1)))
        flip = pair_update<(B > 3 ? B - 1 : B), B, W>(c, p, h, row.d, recs, slot, sweep, key0,
                                                      key1, kacc, dacc, nacc, thr);
    else
        flip = pair_update<B, B, W>(c, p, h, row.d, recs, slot, sweep, key0, key1, kacc, dacc,
                                    nacc, thr);
    uint32_t* Wh = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(Wr) + p * 8u + h * 4u);
    *Wh = x ^ flip;
}

// Drains a wave's per-replica tie queue (irregular_site entries relative to
// qbase) with all 64 lanes.
__device__ __forceinline__ void drain_queue(const uint32_t* queue, const uint32_t* qcount,
                                            uint32_t qcap, uint32_t qbase, uint64_t* Wr,
                                            const unsigned long long* recs, uint32_t slot,
                                            uint32_t sweep, uint32_t key0, uint32_t key1,
                                            int32_t w0, unsigned long long* edef, uint32_t lane,
                                            uint32_t& nacc) {
    const uint32_t queued = min(*qcount, qcap);
    for (uint32_t q = lane; q < queued; q += kWave) {
        const uint32_t code = queue[q];
        const uint32_t p = qbase + (code >> 11), r = (code >> 5) & 63u, m = code & 31u;
        if (resolve_tail(p, slot, sweep, key0, key1, r, record_thr(recs, m, r))) {
            atomicXor(reinterpret_cast<unsigned long long*>(Wr + p), 1ull << r);
            atomicAdd(&edef[r], (unsigned long long)(((int64_t)m * w0) / 2));
            nacc += 1;
        }
    }
}

}  // namespace msc
}  // namespace ige
