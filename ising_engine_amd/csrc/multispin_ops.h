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

// Word i of S: a plain load, or (kCoherent) a device-coherent one that reads
// the L2 and not the CU's L1, for words other workgroups may have written
// since this CU last read them with no cache invalidation in between.
template <bool kCoherent>
__device__ __forceinline__ uint64_t ld_word(const uint64_t* S, uint32_t i) {
    if constexpr (kCoherent)
        return __hip_atomic_load(const_cast<uint64_t*>(reinterpret_cast<const uint64_t*>(
                                     reinterpret_cast<const char*>(S) + i * 8u)),
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return ld32(S, i);
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
__device__ __forceinline__ void build_records(unsigned long long* recs, const uint32_t* lev_thr,
                                              uint32_t tr, uint32_t dmax, uint32_t m0,
                                              uint32_t step, uint32_t lane) {
    for (uint32_t m = m0; m <= dmax; m += step) {
        const uint32_t th = lev_thr[tr * (dmax + 1) + m];
        unsigned long long* out = recs + m * kQStride;
        uint64_t mine = 0;
#pragma unroll
        for (uint32_t b = 0; b < kEagerPlanes; ++b) {
            const uint64_t plane = wave_ballot((th >> (31u - b)) & 1u);
            if (lane == b) mine = plane;
        }
        const uint64_t nz = wave_ballot(th != 0);
        if (lane < kEagerPlanes) out[lane] = mine;
        if (lane == kEagerPlanes) out[kEagerPlanes] = nz;
        reinterpret_cast<uint32_t*>(out + kEagerPlanes + 1)[lane] = th;
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
// (qcap entries) is full.  kCoherent: words are read with ld_word<true>.
template <int B, int W, bool kCoherent = false>
__device__ __forceinline__ void irregular_site(
    const uint64_t* S, uint64_t* Wr, uint32_t p, uint32_t qoff, const uint32_t* rp,
    const uint32_t* ci, const uint32_t* sgn, const unsigned long long* recs, uint32_t slot,
    uint32_t sweep, uint32_t key0, uint32_t key1, int32_t w0, uint64_t (&kacc)[W],
    uint64_t (&dacc)[W], uint32_t& nacc, uint32_t* queue, uint32_t* qcount, uint32_t qcap,
    unsigned long long* edef) {
    constexpr int LMAX = 1 << (B - 1);
    const uint64_t x = ld_word<kCoherent>(S, p);
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
        for (uint32_t q = 0; q < kRowGroup; ++q) wj[q] = ld_word<kCoherent>(S, j[q]);
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

// irregular_site on the narrowest count planes the wave's sites allow: a
// site's update depends on its degree, its counts, uniforms and thresholds,
// not on the planes' width, so when every degree dw of the wave is below
// 2^(B-1) the wave counts on B - 1 planes, with half the levels to compare.
// (An ER graph of mean degree 6 and N = 20000 has a maximum degree near 20,
// so B = 5, but almost every wave of 64 sites stays below 16.)
template <int B, int W, bool kCoherent = false>
__device__ __forceinline__ void irregular_site_fit(
    uint32_t dw, const uint64_t* S, uint64_t* Wr, uint32_t p, uint32_t qoff, const uint32_t* rp,
    const uint32_t* ci, const uint32_t* sgn, const unsigned long long* recs, uint32_t slot,
    uint32_t sweep, uint32_t key0, uint32_t key1, int32_t w0, uint64_t (&kacc)[W],
    uint64_t (&dacc)[W], uint32_t& nacc, uint32_t* queue, uint32_t* qcount, uint32_t qcap,
    unsigned long long* edef) {
    if constexpr (B > 3) {
        if (dw < (1u << (B - 1))) {
            irregular_site<B - 1, W, kCoherent>(S, Wr, p, qoff, rp, ci, sgn, recs, slot, sweep,
                                                key0, key1, w0, kacc, dacc, nacc, queue, qcount,
                                                qcap, edef);
            return;
        }
    }
    irregular_site<B, W, kCoherent>(S, Wr, p, qoff, rp, ci, sgn, recs, slot, sweep, key0, key1,
                                    w0, kacc, dacc, nacc, queue, qcount, qcap, edef);
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
