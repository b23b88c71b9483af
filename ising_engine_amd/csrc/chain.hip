// chain.hip -- many Metropolis sweeps of a one-slot multispin state of
// irregular degree in one launch: every colour class of every sweep, either
// with the launch's workgroups meeting at a device-wide barrier after each
// class (sweep_chain_ms_kernel) or with each site waiting only for its own
// neighbours (sweep_flow_ms_kernel, further below; the default: 39 against
// 45 us a sweep on ER N=20000).
//
// For graphs whose classes are too large for one workgroup but too small to
// fill the GPU in a launch each (ER N=20000: 7 classes of about 2900 sites),
// the per-class launches cost more than the class itself.  Here the launch's
// waves keep a fixed share of every class (wave w takes positions cb + 32 w +
// lane / 2, + 32 W, ..., a lane pair per position: irregular_site_pair), so a
// class costs one site update per lane plus one barrier.  The updates are
// those of launch_sweep_multispin (uniforms keyed by position, slot and
// sweep, the same thresholds and tie resolution), so the chain is bit-exact
// against the per-class launches.
//
// A class's time is the latency of its slowest site update, so the update
// is cut for latency: a position's 64 replicas on two lanes (half the bit
// operations and Philox calls per lane), its row (degree, signs, the first
// neighbour positions) loaded while the wave waits at the barrier before the
// class, all of its neighbour words loaded at once (only up to the degree: a
// class's gathers are bound by the L2's request rate), and the counts from a
// carry-save tree.
//
// A device-scope barrier on a GPU of several XCDs writes back and
// invalidates the L2 of every XCD taking part, so the following class reads
// its rows and words again from memory.  The launch therefore places a
// slot's workgroups on one XCD: it launches kXcdStride times as many
// workgroups, of which those with blockIdx.x % kXcdStride == 0 take part
// (workgroups are dispatched to the XCDs in turn), and the members check that
// they share an XCD (the XCC_ID register) at a first device-scope barrier.
// If they do, the class barriers neither write back nor invalidate any cache
// (xcd_barrier: the words meet in the XCD's L2, and the site updates read
// them with device-coherent loads, which go to that L2), so the model's rows
// stay cached from class to class; otherwise every barrier is the
// device-scope gang_barrier.
//
// Measured on MI355X (ER N=20000, degree 6, 7 classes): with device-scope
// barriers 69-75 us a sweep against 91 us for the captured per-class
// launches; on one XCD with coherent word loads 64.5 us, and with each
// wave's sites counted on the narrowest planes their degrees allow 59.9 us.
// A lane per position: 60.1 us; lane pairs, rows loaded at the barrier and
// the words in one batch of masked loads 50.0 us; carry-save counts 48.8 us;
// the next sweep's records built at the last barrier 49.4 against 50.4 us
// (same GPU, alternating runs).  With 16 slots (192 workgroups,
// device-scope barriers) 369 us against 107 us, and a lane-per-replica form
// with one group of four positions per wave (182 workgroups) 181 us.
#include <hip/hip_runtime.h>

#include <atomic>

#include "gang_sync.h"
#include "ige_math.h"
#include "kernels.hpp"
#include "multispin_ops.h"

namespace ige {
namespace {

constexpr uint32_t kChainBlock = 256;  // 4 waves
constexpr uint32_t kChainWaves = kChainBlock / 64;
constexpr uint32_t kXcdStride = 8;     // XCDs of the GPU: workgroups per member
constexpr uint32_t kPairSites = 32;    // positions per wave and pass: a lane pair each
}  // namespace

constexpr uint32_t kChainMaxBlocks = kChainMaxSites / (64 * 4);  // workgroups per slot
static_assert(kChainMaxBlocks * kChainWaves * kPairSites * 2 >= kChainMaxSites,
              "a wave takes at most two passes of a class");

namespace {
// Bit-sliced counts and plane-view uniforms as sweep_multispin_kernel:
// blockIdx.y = slot; of the gridDim.x workgroups per slot, the gang of
// gridDim.x / kXcdStride with blockIdx.x % kXcdStride == 0 takes part and
// meets at the slot's own barrier bar[slot] after each class (bar[cells +
// slot]: the XCDs its members run on).  Wave w of a slot takes positions
// cb + 32 w + lane / 2, + 32 W, ... of every class (W waves per slot; lane &
// 1 picks the replica half), resolving its tied replicas in place.  Sweep k
// reads its level thresholds from lev + k * lev_stride.

template <int B, int W>
__global__ __launch_bounds__(kChainBlock, 2) void sweep_chain_ms_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const uint32_t* __restrict__ sgn, const uint32_t* __restrict__ cstart, uint32_t ncol,
    uint64_t* spins, uint32_t n, uint32_t dmax, int32_t w0, const uint32_t* __restrict__ lev,
    uint32_t lev_stride, const uint32_t* __restrict__ temp_of, int64_t* energy,
    unsigned long long* accepted, uint32_t sweep0, uint32_t sweeps, uint32_t key0,
    uint32_t key1, uint32_t* bar, uint32_t spin_limit, uint32_t* err) {
    using namespace msc;
    static_assert(W >= B + 3, "counters too narrow");
    constexpr uint32_t kFlush = 8;  // sites per lane between counter flushes
    extern __shared__ unsigned long long recs[];  // 2 buffers of (2^B + 1) * kQStride
    __shared__ unsigned long long blk_e[64];
    __shared__ uint32_t blk_acc;
    if (blockIdx.x % kXcdStride != 0) return;  // not a member: on another XCD
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t slot = blockIdx.y;
    const uint32_t gang = gridDim.x / kXcdStride;
    const uint32_t gw = (blockIdx.x / kXcdStride) * kChainWaves + wv, tw = gang * kChainWaves;
    uint64_t* S = spins + (uint64_t)slot * n;
    uint32_t* sbar = bar + slot;
    uint32_t* sxcd = bar + gridDim.y + slot;
    const uint32_t tr = temp_of[slot * 64 + lane];
    if (wv == 0) blk_e[lane] = 0;
    if (tid == 0) blk_acc = 0;
    uint32_t kacc[W], dacc[W];
#pragma unroll
    for (int b = 0; b < W; ++b) kacc[b] = dacc[b] = 0;
    int64_t e_lane = 0;
    uint32_t nacc = 0, pending = 0, target = 0;
    bool one_xcd = false;
    if (gang > 1) {
        if (tid == 0) atomicOr(sxcd, 1u << xcc_id());
        gang_barrier(sbar, target, gang, spin_limit, err);
        one_xcd = __builtin_popcount(
                      __hip_atomic_load(sxcd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 1;
    }
    auto flush = [&]() {
        const int64_t ksum = wave_column_sum_pair<W>(kacc, lane);
        const int64_t dsum = wave_column_sum_pair<W>(dacc, lane);
        e_lane += ((int64_t)w0 * (2 * ksum - dsum)) / 2;
#pragma unroll
        for (int b = 0; b < W; ++b) kacc[b] = dacc[b] = 0;
        pending = 0;
    };
    // The row of this wave's next position and the wave's largest degree
    // there.  Rows are read-only, so the first pass of a class loads them
    // while the wave waits at the barrier before it.
    PairRow row;
    uint32_t dw = 0;
    auto load_row = [&](uint32_t base, uint32_t ce) {
        const uint32_t p = base + (lane >> 1);
        if (p < ce) row = load_pair_row<B>(p, rp, ci, sgn);
        dw = __builtin_amdgcn_readfirstlane(wave_max_u32(p < ce ? row.d : 0u));
    };
    auto first_row = [&](uint32_t c) {
        load_row(__builtin_amdgcn_readfirstlane(cstart[c]) + gw * kPairSites,
                 __builtin_amdgcn_readfirstlane(cstart[c + 1]));
    };
    first_row(0);
    // Two record buffers: sweep k reads buffer k & 1, and the next sweep's
    // records are built while the last class's barrier is pending.
    constexpr uint32_t kRecWords = ((1u << B) + 1) * kQStride;
    if (sweeps > 0)
        build_records(recs, lev, tr, dmax, wv, kChainWaves, lane);
    __syncthreads();
    for (uint32_t k = 0; k < sweeps; ++k) {
        const uint32_t sweep = sweep0 + k;
        const unsigned long long* R = recs + (k & 1u) * kRecWords;
        for (uint32_t c = 0; c < ncol; ++c) {
            const uint32_t cb = __builtin_amdgcn_readfirstlane(cstart[c]);
            const uint32_t ce = __builtin_amdgcn_readfirstlane(cstart[c + 1]);
            for (uint32_t base = cb + gw * kPairSites; base < ce; base += tw * kPairSites) {
                const uint32_t p = base + (lane >> 1);
                if (p < ce)
                    irregular_site_pair<B, W>(dw, S, S, p, lane & 1u, row, ci, R, slot, sweep,
                                              key0, key1, kacc, dacc, nacc);
                if (++pending == kFlush) flush();  // wave-uniform
                if (base + tw * kPairSites < ce) load_row(base + tw * kPairSites, ce);
            }
            // Arrive, load the next class's rows, build the next sweep's
            // records after the last class and flush counters, then wait.
            if (gang > 1 && one_xcd)
                xcd_arrive(sbar, target, gang);
            else if (gang > 1)
                gang_arrive(sbar, target, gang);
            first_row(c + 1 < ncol ? c + 1 : 0);
            if (c + 1 == ncol && k + 1 < sweeps)
                build_records(recs + ((k + 1) & 1u) * kRecWords,
                              lev + (uint64_t)(k + 1) * lev_stride, tr, dmax, wv, kChainWaves,
                              lane);
            if (2 * pending >= kFlush) flush();
            if (gang > 1)
                gang_wait(sbar, target, spin_limit, err, !one_xcd);
            else
                __syncthreads();
        }
    }
    flush();
    for (uint32_t off = 32; off > 0; off >>= 1) nacc += __shfl_xor(nacc, (int)off, 64);
    atomicAdd(&blk_e[lane], (unsigned long long)e_lane);
    if (lane == 0) atomicAdd(&blk_acc, nacc);
    __syncthreads();
    if (wv == 0) {
        if (blk_e[lane])
            atomicAdd(reinterpret_cast<unsigned long long*>(energy + slot * 64 + lane), blk_e[lane]);
        if (lane == 0 && blk_acc) atomicAdd(accepted, (unsigned long long)blk_acc);
    }
}

// ---- Chained sweeps without barriers ------------------------------------
//
// sweep_flow_ms_kernel runs the same updates as sweep_chain_ms_kernel, but a
// site waits only for its own neighbours instead of a gang barrier after
// every class.  The words live in 64-bit units for the launch, U[2 p + h] =
// (v << 32) | replicas 32 h .. 32 h + 31 of site p, where v counts the
// launch's sweeps that have updated p; one relaxed agent-scope atomic store
// writes a half word with its count, and one atomic load reads both, so a
// reader that sees the count it needs holds the words that go with it.
// Site p of class c (positions from cb) at the launch's sweep k reads its
// neighbour j once v(j) >= k + [j < cb]: neighbours of earlier classes must
// have made this sweep's update, those of later classes the previous one.
// That is the order the class-by-class sweep reads them in, so the result
// is the same bits; and no neighbour can move further before p has been
// updated, since it waits for p in turn (v(j) can only be k + [j < cb] when
// p reads it).  Each wave takes its items (sweep, class, pass) in order and
// every wait is for an item earlier in that order, so every wait ends once
// all members run (the members are resident: checked on the host).
// Waits are bounded (spin_limit polls; on expiry err is raised and the
// launch finishes with stale reads, which the host reports as an error).

using namespace msc;

// Half h of site i's unit.
__device__ __forceinline__ uint64_t ld_unit(const uint64_t* U, uint32_t i, uint32_t h) {
    return __hip_atomic_load(const_cast<uint64_t*>(U) + 2u * i + h, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until the units of entries q with bit q of pend set carry counts of
// at least k + [j[q] < cb] (the whole wave leaves together), leaving their
// half words in wj.  A pass loads every pending unit; while some are not
// ready, the wave watches just one of them (the first pending entry of its
// first waiting lane) with one uniform load per poll, then makes another
// pass.  Repeated passes over every pending unit instead kept the L2 and the
// CUs' memory pipelines so busy that a sweep took ten times as long.
template <uint32_t N>
__device__ __forceinline__ void poll_units(const uint64_t* U, const uint32_t (&j)[N],
                                           uint32_t cb, uint32_t k, uint32_t h, uint32_t pend,
                                           uint32_t (&wj)[N], uint32_t spin_limit, uint32_t* err) {
    for (uint32_t polls = 0;;) {
        uint64_t u[N];
#pragma unroll
        for (uint32_t q = 0; q < N; ++q) {
            u[q] = 0;
            if ((pend >> q) & 1u) u[q] = ld_unit(U, j[q], h);
        }
#pragma unroll
        for (uint32_t q = 0; q < N; ++q) {
            const uint32_t need = k + (j[q] < cb ? 1u : 0u);
            if (((pend >> q) & 1u) && (uint32_t)(u[q] >> 32) >= need) {
                wj[q] = (uint32_t)u[q];
                pend &= ~(1u << q);
            }
        }
        const uint64_t waiting = wave_ballot(pend != 0u);
        if (waiting == 0) return;
        // this lane's first pending entry, then the first waiting lane's
        uint32_t jf = 0;
#pragma unroll
        for (int q = (int)N - 1; q >= 0; --q)
            if ((pend >> q) & 1u) jf = j[q];
        const uint32_t L = (uint32_t)__builtin_ctzll(waiting);
        const uint32_t jw = __builtin_amdgcn_readlane(jf, L);
        const uint32_t hw = L & 1u, need = k + (jw < cb ? 1u : 0u);
        for (;;) {
            if (++polls > spin_limit) {
                if (pend) atomicOr(err, 1u);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
            const uint64_t v = ld_unit(U, jw, hw);
            const uint32_t failed =
                __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)__builtin_amdgcn_readfirstlane(failed) != 0u) return;
            if ((uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) >= need) break;
        }
    }
}

// pair_counts over the units: waits for the neighbours' counts, then counts.
// Inactive lanes (act false) wait for nothing and count zero.
template <int B>
__device__ __forceinline__ uint32_t flow_counts(const uint64_t* U, uint32_t p, uint32_t h,
                                                bool act, const PairRow& row,
                                                const uint32_t* ci, uint32_t cb, uint32_t k,
                                                uint32_t (&c)[B], uint32_t spin_limit,
                                                uint32_t* err) {
    const uint32_t d = act ? row.d : 0u, sg = row.sg;
    constexpr uint32_t kPre = pair_pre<B>(), kG = kPairPre / 2;
    const uint32_t x = act ? (uint32_t)ld_unit(U, p, h) : 0u;  // this lane's own last store
    uint32_t jj[kPre], wj[kPre];
#pragma unroll
    for (uint32_t q = 0; q < kPre; ++q) {
        jj[q] = row.j[q];
        wj[q] = 0;
    }
    const uint32_t pre = d < kPre ? d : kPre;
    poll_units<kPre>(U, jj, cb, k, h, pre ? (uint32_t)((1ull << pre) - 1u) : 0u, wj, spin_limit,
                     err);
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
    // entries past kPre (degrees above 16), in groups of kG; the loop runs
    // as long as some lane of the wave has entries left
    const uint32_t dmx = __builtin_amdgcn_readfirstlane(wave_max_u32(d));
    for (uint32_t g = kPre; g < dmx; g += kG) {
        uint32_t j2[kG], w2[kG];
#pragma unroll
        for (uint32_t q = 0; q < kG; ++q) {
            j2[q] = g + q < d ? ld32(ci, row.e0 + g + q) : 0u;
            w2[q] = 0;
        }
        const uint32_t left = g < d ? min(d - g, kG) : 0u;
        poll_units<kG>(U, j2, cb, k, h, (1u << left) - 1u, w2, spin_limit, err);
#pragma unroll
        for (uint32_t q = 0; q < kG; ++q) {
            uint32_t carry = __builtin_amdgcn_bitop3_b32(x, w2[q], bit_mask(sg, (g + q) & 31u),
                                                         kXor3) &
                             (g + q < d ? ~0u : 0u);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const uint32_t sb = c[b] ^ carry;
                carry &= c[b];
                c[b] = sb;
            }
        }
    }
    return x;
}

// A wave's threshold records for one sweep, kFlowRS u64 per level m <= dmax
// (the kEagerPlanes leading bit planes of the 64 replicas' thresholds, then
// the nonzero mask): lane r reads thr(temperature of lane r, m).
constexpr uint32_t kFlowRS = kEagerPlanes + 1;
__device__ __forceinline__ void build_planes(unsigned long long* recs, const uint32_t* lev_thr,
                                             uint32_t tr, uint32_t dmax, uint32_t lane) {
    constexpr uint32_t kBatch = 8;
    for (uint32_t m1 = 0; m1 <= dmax; m1 += kBatch) {
        uint32_t th[kBatch];
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i)
            th[i] = m1 + i <= dmax ? lev_thr[tr * (dmax + 1) + m1 + i] : 0u;
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) {
            if (m1 + i > dmax) break;  // uniform
            uint64_t mine = 0;
#pragma unroll
            for (uint32_t b = 0; b < kEagerPlanes; ++b) {
                const uint64_t plane = wave_ballot((th[i] >> (31u - b)) & 1u);
                if (lane == b) mine = plane;
            }
            const uint64_t nz = wave_ballot(th[i] != 0);
            unsigned long long* out = recs + (m1 + i) * kFlowRS;
            if (lane < kEagerPlanes) out[lane] = mine;
            if (lane == kEagerPlanes) out[kEagerPlanes] = nz;
        }
    }
    // the wave reads these records next: keep the compiler from moving the
    // reads above the writes (LDS serves one wave's accesses in order)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int B, int W>
__global__ __launch_bounds__(kChainBlock, 2) void sweep_flow_ms_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const uint32_t* __restrict__ sgn, const uint32_t* __restrict__ cstart, uint32_t ncol,
    uint64_t* U, uint32_t dmax, int32_t w0, const uint32_t* __restrict__ lev,
    uint32_t lev_stride, const uint32_t* __restrict__ temp_of, int64_t* energy,
    unsigned long long* accepted, uint32_t sweep0, uint32_t sweeps, uint32_t key0,
    uint32_t key1, uint32_t spin_limit, uint32_t* err) {
    using namespace msc;
    static_assert(W >= B + 3, "counters too narrow");
    constexpr uint32_t kFlush = 8;  // sites per lane between counter flushes
    constexpr uint32_t kRecW = ((1u << B) + 1) * kFlowRS;  // levels 0 .. 2^B are read
    __shared__ unsigned long long recs[kChainWaves][kRecW];
    __shared__ uint32_t trs[64];
    __shared__ unsigned long long blk_e[64];
    __shared__ uint32_t blk_acc;
    if (blockIdx.x % kXcdStride != 0) return;  // not a member: on another XCD
    const uint32_t tid = threadIdx.x, lane = tid & 63u, h = lane & 1u;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t gw = (blockIdx.x / kXcdStride) * kChainWaves + wv;
    const uint32_t tw = (gridDim.x / kXcdStride) * kChainWaves;
    const uint32_t tr = temp_of[lane];  // one slot
    if (wv == 0) {
        trs[lane] = tr;
        blk_e[lane] = 0;
    }
    if (tid == 0) blk_acc = 0;
    __syncthreads();
    unsigned long long* R = recs[wv];
    uint32_t kacc[W], dacc[W];
#pragma unroll
    for (int b = 0; b < W; ++b) kacc[b] = dacc[b] = 0;
    int64_t e_lane = 0;
    uint32_t nacc = 0, pending = 0;
    auto flush = [&]() {
        const int64_t ksum = wave_column_sum_pair<W>(kacc, lane);
        const int64_t dsum = wave_column_sum_pair<W>(dacc, lane);
        e_lane += ((int64_t)w0 * (2 * ksum - dsum)) / 2;
#pragma unroll
        for (int b = 0; b < W; ++b) kacc[b] = dacc[b] = 0;
        pending = 0;
    };
    // row of position p for an active lane, zeros otherwise
    auto row_of = [&](uint32_t p, bool act) {
        PairRow r;
        r.e0 = r.d = r.sg = 0;
#pragma unroll
        for (uint32_t q = 0; q < kPairPre; ++q) r.j[q] = 0;
        if (act) r = load_pair_row<B>(p, rp, ci, sgn);
        return r;
    };
    for (uint32_t k = 0; k < sweeps; ++k) {
        const uint32_t sweep = sweep0 + k;
        const uint32_t* levk = lev + (uint64_t)k * lev_stride;
        if (k == 0 || lev_stride) build_planes(R, levk, tr, dmax, lane);
        auto tie_thr = [&](uint32_t m, uint32_t r) { return levk[trs[r] * (dmax + 1) + m]; };
        for (uint32_t c = 0; c < ncol; ++c) {
            const uint32_t cb = __builtin_amdgcn_readfirstlane(cstart[c]);
            const uint32_t ce = __builtin_amdgcn_readfirstlane(cstart[c + 1]);
            for (uint32_t base = cb + gw * kPairSites; base < ce; base += tw * kPairSites) {
                const uint32_t p = base + (lane >> 1);
                const bool act = p < ce;
                const PairRow row = row_of(p, act);
                const uint32_t dw =
                    __builtin_amdgcn_readfirstlane(wave_max_u32(act ? row.d : 0u));
                uint32_t cnt[B];
                const uint32_t x = flow_counts<B>(U, p, h, act, row, ci, cb, k, cnt, spin_limit,
                                                  err);
                if (act) {
                    uint32_t flip;
                    if (B > 3 && dw < (1u << (B - 1)))
                        flip = pair_update<(B > 3 ? B - 1 : B), B, W, kFlowRS>(
                            cnt, p, h, row.d, R, 0u, sweep, key0, key1, kacc, dacc, nacc, tie_thr);
                    else
                        flip = pair_update<B, B, W, kFlowRS>(cnt, p, h, row.d, R, 0u, sweep, key0,
                                                             key1, kacc, dacc, nacc, tie_thr);
                    __hip_atomic_store(U + 2u * p + h,
                                       ((uint64_t)(k + 1u) << 32) | (uint64_t)(x ^ flip),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (++pending == kFlush) flush();  // wave-uniform
            }
        }
    }
    flush();
    for (uint32_t off = 32; off > 0; off >>= 1) nacc += __shfl_xor(nacc, (int)off, 64);
    atomicAdd(&blk_e[lane], (unsigned long long)e_lane);
    if (lane == 0) atomicAdd(&blk_acc, nacc);
    __syncthreads();
    if (wv == 0) {
        if (blk_e[lane]) atomicAdd(reinterpret_cast<unsigned long long*>(energy + lane), blk_e[lane]);
        if (lane == 0 && blk_acc) atomicAdd(accepted, (unsigned long long)blk_acc);
    }
}

// U[2 p + h] = half h of S[p] with count 0, and back.
__global__ void flow_pack_kernel(const uint64_t* __restrict__ S, uint64_t* __restrict__ U,
                                 uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2u * n) U[i] = (uint32_t)(S[i >> 1] >> (32u * (i & 1u)));
}

__global__ void flow_unpack_kernel(const uint64_t* __restrict__ U, uint64_t* __restrict__ S,
                                   uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) S[i] = (uint64_t)(uint32_t)U[2u * i] | ((uint64_t)(uint32_t)U[2u * i + 1u] << 32);
}

__global__ void chain_fill_kernel(uint32_t* p, uint32_t count) {
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0u;
}

}  // namespace

namespace {

// Two buffers of threshold records of levels 0 .. 2^B for the count planes
// B of degree dmax (irregular_site_pair reads levels up to 2^B).
uint32_t chain_lds(uint32_t dmax) {
    const uint32_t planes = dmax < 8 ? 3 : dmax < 16 ? 4 : 5;
    return 2 * ((1u << planes) + 1) * kQStride * sizeof(unsigned long long);
}

template <int B, int W>
int chain_ms_per_cu(uint32_t lds) {
    int pc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &pc, reinterpret_cast<const void*>(sweep_chain_ms_kernel<B, W>), kChainBlock, lds) !=
        hipSuccess)
        return 0;
    return pc;
}

}  // namespace

uint32_t chain_ms_capacity_blocks(uint32_t dmax) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    const uint32_t lds = chain_lds(dmax);
    const int pc = dmax < 8 ? chain_ms_per_cu<3, 6>(lds)
                 : dmax < 16 ? chain_ms_per_cu<4, 7>(lds)
                 : dmax < 32 ? chain_ms_per_cu<5, 8>(lds) : 0;
    return pc > 0 && ncu > 0 ? (uint32_t)(pc * ncu) : 0u;
}

hipError_t launch_sweep_chain_ms(const DevModel& m, const DevState& s, const uint32_t* cstart,
                                 uint32_t ncol, uint32_t max_class, const uint32_t* lev,
                                 uint32_t lev_stride, uint32_t sweep0, uint32_t sweeps,
                                 uint64_t seed, uint32_t* bar, uint32_t* err, uint32_t spin_limit,
                                 hipStream_t stream) {
    const uint32_t cap = chain_ms_capacity_blocks(m.dmax);
    if (cap < s.cells || m.dmax >= 32) return hipErrorInvalidValue;
    // one site per lane of the largest class, at most kChainMaxBlocks
    // workgroups per slot and what one XCD holds for every slot at once
    const uint32_t want = (max_class + kPairSites * kChainWaves - 1) / (kPairSites * kChainWaves);
    uint32_t per_slot = cap / kXcdStride / s.cells;
    if (per_slot > kChainMaxBlocks) per_slot = kChainMaxBlocks;
    if (want < per_slot) per_slot = want ? want : 1u;
    if (per_slot == 0) return hipErrorInvalidValue;
    const uint32_t lds = chain_lds(m.dmax);
    const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);
    const uint64_t per_launch = ((1ull << 31) / per_slot) / (ncol ? ncol : 1u);
    for (uint32_t done = 0; done < sweeps;) {
        const uint32_t k =
            (uint32_t)((uint64_t)(sweeps - done) < per_launch ? sweeps - done : per_launch);
        hipLaunchKernelGGL(chain_fill_kernel, dim3(1), dim3(256), 0, stream, bar, 2 * s.cells);
        const uint32_t* L = lev + (uint64_t)done * lev_stride;
#define IGE_CHAIN_MS(B, W)                                                                       \
    hipLaunchKernelGGL((sweep_chain_ms_kernel<B, W>), dim3(per_slot * kXcdStride, s.cells),     \
                       dim3(kChainBlock),                                                        \
                       lds, stream, m.rp, m.ci, m.sgn, cstart, ncol, s.spins, m.n, m.dmax, m.w0, \
                       L, lev_stride, s.temp_of, s.energy, s.accepted, sweep0 + done, k, key0,   \
                       key1, bar, spin_limit, err)
        if (m.dmax < 8)
            IGE_CHAIN_MS(3, 6);
        else if (m.dmax < 16)
            IGE_CHAIN_MS(4, 7);
        else
            IGE_CHAIN_MS(5, 8);
#undef IGE_CHAIN_MS
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        done += k;
    }
    return hipSuccess;
}

namespace {
template <int B, int W>
int flow_ms_per_cu() {
    int pc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &pc, reinterpret_cast<const void*>(sweep_flow_ms_kernel<B, W>), kChainBlock, 0) !=
        hipSuccess)
        return 0;
    return pc;
}
}  // namespace

uint32_t flow_ms_capacity_blocks(uint32_t dmax) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    const int pc = dmax < 8 ? flow_ms_per_cu<3, 6>()
                 : dmax < 16 ? flow_ms_per_cu<4, 7>()
                 : dmax < 32 ? flow_ms_per_cu<5, 8>() : 0;
    return pc > 0 && ncu > 0 ? (uint32_t)(pc * ncu) : 0u;
}

hipError_t launch_sweep_flow_ms(const DevModel& m, const DevState& s, const uint32_t* cstart,
                                uint32_t ncol, uint32_t max_class, const uint32_t* lev,
                                uint32_t lev_stride, uint32_t sweep0, uint32_t sweeps,
                                uint64_t seed, uint64_t* units, uint32_t* err,
                                uint32_t spin_limit, hipStream_t stream) {
    const uint32_t cap = flow_ms_capacity_blocks(m.dmax);
    if (s.cells != 1 || m.dmax >= 32 || cap == 0 || sweeps == 0xFFFFFFFFu)
        return hipErrorInvalidValue;
    // one site per lane pair of the largest class, at most kChainMaxBlocks
    // workgroups and what one XCD holds
    const uint32_t want = (max_class + kPairSites * kChainWaves - 1) / (kPairSites * kChainWaves);
    uint32_t members = cap / kXcdStride;
    if (members > kChainMaxBlocks) members = kChainMaxBlocks;
    if (want < members) members = want ? want : 1u;
    if (members == 0) return hipErrorInvalidValue;
    const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);
    hipLaunchKernelGGL(flow_pack_kernel, dim3((2 * m.n + 255) / 256), dim3(256), 0, stream,
                       s.spins, units, m.n);
#define IGE_FLOW_MS(B, W)                                                                        \
    hipLaunchKernelGGL((sweep_flow_ms_kernel<B, W>), dim3(members * kXcdStride), dim3(kChainBlock), \
                       0, stream, m.rp, m.ci, m.sgn, cstart, ncol, units, m.dmax, m.w0, lev,     \
                       lev_stride, s.temp_of, s.energy, s.accepted, sweep0, sweeps, key0, key1,  \
                       spin_limit, err)
    if (m.dmax < 8)
        IGE_FLOW_MS(3, 6);
    else if (m.dmax < 16)
        IGE_FLOW_MS(4, 7);
    else
        IGE_FLOW_MS(5, 8);
#undef IGE_FLOW_MS
    hipLaunchKernelGGL(flow_unpack_kernel, dim3((m.n + 255) / 256), dim3(256), 0, stream, units,
                       s.spins, m.n);
    return hipGetLastError();
}

}  // namespace ige
