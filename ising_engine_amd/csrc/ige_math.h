// ige_math.h -- counter-based randoms and the deterministic acceptance
// thresholds shared by host and device code.
//
// Everything here is plain integer or IEEE-754 double arithmetic, compiled
// with -ffp-contract=off, so host and gfx950 evaluate it to identical bits.
// That is what makes a GPU trajectory reproducible on the CPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IGE_HD __host__ __device__ __forceinline__

namespace ige {

// Stream tags: top byte of Philox counter word 3.
constexpr uint32_t kTagWord = 1u;   // lane-per-replica sweep uniforms
constexpr uint32_t kTagPlane = 2u;  // bit-plane sweep uniforms (multispin)
constexpr uint32_t kTagExch = 3u;   // replica-exchange uniforms
constexpr uint32_t kTagInit = 4u;   // default initial spins
constexpr uint32_t kTagIcm = 6u;    // cluster-move seed sites
constexpr uint32_t kTagRepx = 8u;   // replica-index exchange uniforms (replica view)

struct U32x4 {
    uint32_t v[4];
};

// A counter of a sweep step: `rel` as launched, plus the step clock's word
// when the launch belongs to a replayed graph (clock != NULL).
__device__ __forceinline__ uint32_t clocked(const uint32_t* clock, uint32_t word, uint32_t rel) {
    return clock ? rel + clock[word] : rel;
}

// a ^ b ^ c; one v_bitop3_b32 on the device.
IGE_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

// Philox4x32-10 (Salmon et al., SC'11): ten rounds of two 32x32->64
// multiplies with the Weyl-bumped key.
IGE_HD U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                           uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    return U32x4{{c0, c1, c2, c3}};
}

// exp(x) for x <= 0 to within ~1 ulp: Cody-Waite reduction by ln2 (fdlibm
// split), degree-13 Taylor polynomial in Horner form, exact scaling by 2^k.
// Values below exp(-40) return 0 (they only ever feed a 2^-32 threshold).
IGE_HD double det_exp(double x) {
    if (!(x > -40.0)) return 0.0;
    if (x >= 0.0) return 1.0;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double kd = __builtin_rint(x * inv_ln2);
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return __builtin_ldexp(p, (int)kd);
}

IGE_HD uint32_t prob_to_u32(double e) {
    const double v = e * 4294967296.0;
    return v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
}

// Metropolis threshold: a proposal with dE >= 1 is accepted iff u < thr, u a
// uniform 32-bit integer, so P(accept) = thr / 2^32 ~ exp(-beta dE).
IGE_HD uint32_t metropolis_threshold(double beta, int64_t de) {
    return prob_to_u32(det_exp(-(beta * (double)de)));
}

// Largest dE whose threshold can be nonzero at beta: thr(beta, d) = 0 once
// beta * d > 40 (det_exp returns 0 there); two units of slack keep the bound
// safe against rounding of 40 / beta.
IGE_HD uint32_t threshold_limit(double beta, uint32_t dtab) {
    if (!(beta > 0.0)) return dtab;
    const double lim = 40.0 / beta;
    return lim >= (double)dtab ? dtab : ((uint32_t)lim + 2u < dtab ? (uint32_t)lim + 2u : dtab);
}

// Heat-bath threshold: the new spin of a site with local field h is 1 iff
// u < thr, thr = floor(2^32 P) in [0, 2^32] with P = 1 / (1 + exp(beta h)),
// exp evaluated only at -beta |h| <= 0 (P = e / (1 + e) for h >= 0, 1 / (1 + e)
// for h < 0, e = exp(-beta |h|)).  beta -> inf gives 0, 2^31 or 2^32.
IGE_HD uint64_t heatbath_threshold(double beta, int64_t h) {
    const double a = beta * (double)(h < 0 ? -h : h);
    const double e = det_exp(-a);
    const double p = h >= 0 ? e / (1.0 + e) : 1.0 / (1.0 + e);
    return (uint64_t)(p * 4294967296.0);
}

// Replica-exchange threshold for delta = (b_t - b_{t+1}) (E_t - E_{t+1}) < 0.
IGE_HD uint32_t exchange_threshold(double delta) { return prob_to_u32(det_exp(delta)); }

}  // namespace ige
