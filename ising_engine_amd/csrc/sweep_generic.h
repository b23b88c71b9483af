// sweep_generic.h -- the lane-per-replica Metropolis update of a run of
// positions, shared by the single-state sweep kernel and the batched
// persistent kernel.
//
// One wavefront, lane r = replica r.  Every address is wave-uniform, so CSR
// rows, couplings and neighbour words arrive through scalar loads, and a
// neighbour's spin enters the field as v_cndmask(word-as-lane-mask).
// Uniforms (word view): Philox4x32-10 with counter (p >> 2, slot, sweep,
// 1 << 24 | lane), word p & 3, one call per lane per aligned group of four
// positions.
#pragma once

#include <hip/hip_runtime.h>

#include "ige_math.h"

namespace ige {

__device__ __forceinline__ bool lane_bit(uint64_t word) {
    return __builtin_amdgcn_inverse_ballot_w64(word);
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Updates positions [p0, p1) (all in one colour class) of one slot.
// S/W: the slot's words for reading / writing (same memory).  S is not
// restrict-qualified here: a caller whose kernel argument is restrict and
// never written within the launch gets scalar loads; a persistent caller,
// which reads words that other waves wrote after a barrier, must not (the
// scalar cache is not coherent with vector stores) and passes a plain
// pointer, which makes the loads vector loads.  thr(de)
// returns metropolis_threshold(beta of this lane, de) for de >= 1.
// Adds the accepted dE to eacc (per lane) and the accepted count to nacc.
template <class Thr>
__device__ __forceinline__ void generic_positions(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin,
    const uint64_t* S, uint64_t* W, uint32_t p0, uint32_t p1, uint32_t slot,
    uint32_t sweep, uint32_t key0, uint32_t key1, uint32_t lane, Thr&& thr, int64_t& eacc,
    uint32_t& nacc) {
    U32x4 u{};
    uint32_t group = 0xFFFFFFFFu;
    for (uint32_t p = p0; p < p1; ++p) {
        if ((p >> 2) != group) {
            group = p >> 2;
            u = philox4x32_10(group, slot, sweep, (kTagWord << 24) | lane, key0, key1);
        }
        const uint32_t q = p & 3u;
        const uint32_t uu = q == 0 ? u.v[0] : q == 1 ? u.v[1] : q == 2 ? u.v[2] : u.v[3];
        const uint64_t x = S[p];
        int32_t h = lin[p];
        const uint32_t e1 = rp[p + 1];
        for (uint32_t e = rp[p]; e < e1; ++e) {
            const int32_t w = wt[e];
            h += lane_bit(S[ci[e]]) ? w : 0;
        }
        const int32_t de = lane_bit(x) ? -h : h;
        bool acc = de <= 0;
        if (!acc) acc = uu < thr(de);
        const uint64_t m = wave_ballot(acc);
        if (lane == 0) W[p] = x ^ m;
        eacc += acc ? de : 0;
        nacc += (uint32_t)__builtin_popcountll(m);
    }
}

}  // namespace ige
