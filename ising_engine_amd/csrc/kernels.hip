// kernels.hip -- gfx950 kernels of the sweep / exchange hot path.
//
// Wave64 throughout.  Two sweep kernels share one acceptance rule
// (dE <= 0, or u < metropolis_threshold(beta, dE)) and differ in how a wave
// maps onto the MSC words:
//   sweep_generic    lane = replica.  The wave walks the positions of one
//                    colour class for one slot; every word it touches is
//                    wave-uniform, so CSR rows and neighbour words arrive
//                    through scalar loads and a neighbour's spin is applied
//                    with the word itself as the lane mask (v_cndmask on an
//                    SGPR pair).  Any integer CSR model.
//   sweep_multispin  lane = site.  Each lane owns one u64 word and evaluates
//                    all 64 replicas with bitwise arithmetic (+-J models).
#include <hip/hip_runtime.h>

#include <atomic>

#include "ige_math.h"
#include "kernels.hpp"
#include "multispin_ops.h"
#include "sweep_generic.h"

namespace ige {

namespace {

using namespace msc;  // kWave, the bit operations and the site updates

constexpr uint32_t kBlock = 256;  // 4 waves
constexpr uint32_t kWavesPerBlock = kBlock / kWave;

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return wave_ballot(p); }

uint32_t ceil_div(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

// Positions per wave of the lane-per-replica sweeps over a colour class of
// `len` positions in each of `cells` slots: about 16K waves per launch
// (256 CUs x 4 SIMDs x 16) and at least one aligned group of 4 positions (one
// Philox call per lane) per wave; a class too small to give 128 waves at that
// is split further, down to one position per wave, since there the launch's
// duration is one wave's walk along its rows.
uint64_t lane_per_replica_chunk(uint32_t len, uint32_t cells) {
    uint64_t chunk = ((uint64_t)len * cells + 16383) / 16384;
    if (chunk < 4) {
        chunk = 4;
        while (chunk > 1 && (uint64_t)ceil_div(len, chunk) * cells < 128) chunk >>= 1;
    }
    if (chunk > 4096) chunk = 4096;
    return chunk;
}

// ---------------------------------------------------------------------------
// Threshold table: one thread per (t, d).
// ---------------------------------------------------------------------------
__global__ void thr_table_kernel(const double* __restrict__ beta, uint32_t* __restrict__ thr,
                                 uint32_t cells, uint32_t dtab, const uint32_t* __restrict__ clock) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    beta += (uint64_t)clocked(clock, kClockRow, 0) * cells;
    const uint64_t row = (uint64_t)dtab + 1;
    if (i >= row * cells) return;
    const uint32_t t = (uint32_t)(i / row);
    const uint32_t d = (uint32_t)(i % row);
    thr[i] = d == 0 ? 0xFFFFFFFFu : metropolis_threshold(beta[t], (int64_t)d);
}

// ---------------------------------------------------------------------------
// Generic sweep, lane = replica (generic_positions): each wave owns positions
// [p0, p1) of the colour class for one slot.
// ---------------------------------------------------------------------------
template <bool kLdsThr>
__global__ __launch_bounds__(kBlock) void sweep_generic_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin,
    const uint64_t* __restrict__ spins_rd, uint64_t* spins_wr, uint32_t n, uint32_t cells,
    uint32_t temps, const uint32_t* __restrict__ temp_of, const uint32_t* __restrict__ thr,
    uint32_t dtab, int64_t* energy, unsigned long long* accepted, uint32_t c_begin,
    uint32_t c_end, uint32_t chunk, uint32_t waves_per_slot, uint32_t sweep_rel, uint32_t key0,
    uint32_t key1, const uint32_t* __restrict__ clock) {
    extern __shared__ uint32_t lds_thr[];
    const uint32_t sweep = clocked(clock, kClockSweep, sweep_rel);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    if (kLdsThr) {
        const uint32_t total = temps * (dtab + 1);
        for (uint32_t i = threadIdx.x; i < total; i += kBlock) lds_thr[i] = thr[i];
        __syncthreads();
    }
    const uint32_t wave = uniform(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    const uint32_t slot = wave / waves_per_slot;
    const uint32_t p0 = c_begin + (wave % waves_per_slot) * chunk;
    if (slot >= cells || p0 >= c_end) return;
    const uint32_t p1 = min(p0 + chunk, c_end);

    const uint64_t* __restrict__ S = spins_rd + (uint64_t)slot * n;
    uint64_t* W = spins_wr + (uint64_t)slot * n;
    const uint32_t trow = temp_of[slot * kWave + lane] * (dtab + 1);

    int64_t eacc = 0;
    uint32_t nacc = 0;
    generic_positions(rp, ci, wt, lin, S, W, p0, p1, slot, sweep, key0, key1, lane,
                      [&](uint32_t uu, int32_t de) {
                          return uu < (kLdsThr ? lds_thr[trow + (uint32_t)de]
                                               : thr[trow + (uint32_t)de]);
                      },
                      eacc, nacc);
    atomicAdd(reinterpret_cast<unsigned long long*>(energy + slot * kWave + lane),
              (unsigned long long)eacc);
    if (lane == 0) atomicAdd(accepted, (unsigned long long)nacc);
}

// ---------------------------------------------------------------------------
// Heat-bath sweep (gibbs_color_sweep.rs:119-174), lane = replica, same
// mapping and uniforms as sweep_generic_kernel: the new spin is
// u < hb[temperature][h + dtab] with h the local field, and the lane flips
// when that differs from the current spin.  hb: [cells][2 dtab + 1] u64
// thresholds in [0, 2^32] (launch_heatbath_table).
// ---------------------------------------------------------------------------
__global__ void heatbath_table_kernel(const double* __restrict__ beta, uint64_t* __restrict__ hb,
                                      uint32_t cells, uint32_t dtab,
                                      const uint32_t* __restrict__ clock) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    beta += (uint64_t)clocked(clock, kClockRow, 0) * cells;
    const uint64_t row = 2ull * dtab + 1;
    if (i >= row * cells) return;
    const uint32_t t = (uint32_t)(i / row);
    const int64_t h = (int64_t)(i % row) - (int64_t)dtab;
    hb[i] = heatbath_threshold(beta[t], h);
}

template <bool kLdsThr>
__global__ __launch_bounds__(kBlock) void sweep_heatbath_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin,
    const uint64_t* __restrict__ spins_rd, uint64_t* spins_wr, uint32_t n, uint32_t cells,
    uint32_t temps, const uint32_t* __restrict__ temp_of, const uint64_t* __restrict__ hb,
    uint32_t dtab, int64_t* energy, unsigned long long* accepted, uint32_t c_begin,
    uint32_t c_end, uint32_t chunk, uint32_t waves_per_slot, uint32_t sweep_rel, uint32_t key0,
    uint32_t key1, const uint32_t* __restrict__ clock) {
    extern __shared__ uint64_t lds_hb[];
    const uint32_t sweep = clocked(clock, kClockSweep, sweep_rel);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t row = 2 * dtab + 1;
    if (kLdsThr) {
        const uint32_t total = temps * row;
        for (uint32_t i = threadIdx.x; i < total; i += kBlock) lds_hb[i] = hb[i];
        __syncthreads();
    }
    const uint32_t wave = uniform(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    const uint32_t slot = wave / waves_per_slot;
    const uint32_t p0 = c_begin + (wave % waves_per_slot) * chunk;
    if (slot >= cells || p0 >= c_end) return;
    const uint32_t p1 = min(p0 + chunk, c_end);

    const uint64_t* __restrict__ S = spins_rd + (uint64_t)slot * n;
    uint64_t* W = spins_wr + (uint64_t)slot * n;
    const uint32_t trow = temp_of[slot * kWave + lane] * row + dtab;

    int64_t eacc = 0;
    uint32_t nacc = 0;
    generic_positions<true>(rp, ci, wt, lin, S, W, p0, p1, slot, sweep, key0, key1, lane,
                            [&](uint32_t uu, int32_t de, bool x) {
                                const int32_t h = x ? -de : de;
                                const uint64_t t = kLdsThr ? lds_hb[trow + h] : hb[trow + h];
                                return ((uint64_t)uu < t) != x;
                            },
                            eacc, nacc);
    atomicAdd(reinterpret_cast<unsigned long long*>(energy + slot * kWave + lane),
              (unsigned long long)eacc);
    if (lane == 0) atomicAdd(accepted, (unsigned long long)nacc);
}

// ---------------------------------------------------------------------------
// DEO replica exchange number `round`: one wave per adjacent pair (t, t+1)
// with t = round mod 2; lane r decides for replica lane r and swaps
// temperature indices.  Under a step clock the phase is known only on the
// device: the grid then covers the larger phase and waves past the
// population's last pair return.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWave) void exchange_kernel(
    const double* __restrict__ beta, const int64_t* __restrict__ energy, uint32_t* temp_of,
    uint32_t* slot_of, unsigned long long* pair_stats, uint32_t cells, uint32_t pop_temps,
    uint32_t pairs_per_pop, uint32_t round_rel, uint32_t key0, uint32_t key1,
    const uint32_t* __restrict__ clock) {
    const uint32_t r = threadIdx.x;
    const uint32_t round = clocked(clock, kClockRound, round_rel);
    beta += (uint64_t)clocked(clock, kClockRow, 0) * cells;
    // populations of pop_temps levels each: pairs stay inside one population
    const uint32_t pop = blockIdx.x / pairs_per_pop, k = blockIdx.x % pairs_per_pop;
    const uint32_t local = (round & 1u) + 2u * k;
    if (local + 1 >= pop_temps) return;
    const uint32_t t = pop * pop_temps + local;
    const uint32_t s1 = slot_of[t * kWave + r];
    const uint32_t s2 = slot_of[(t + 1) * kWave + r];
    const double delta = (beta[t] - beta[t + 1]) *
                         (double)(energy[s1 * kWave + r] - energy[s2 * kWave + r]);
    const U32x4 u = philox4x32_10(t, r >> 2, round, kTagExch << 24, key0, key1);
    const uint32_t q = r & 3u;
    const uint32_t uu = q == 0 ? u.v[0] : q == 1 ? u.v[1] : q == 2 ? u.v[2] : u.v[3];
    const bool acc = delta >= 0.0 || uu < exchange_threshold(delta);
    if (acc) {
        slot_of[t * kWave + r] = s2;
        slot_of[(t + 1) * kWave + r] = s1;
        temp_of[s1 * kWave + r] = t + 1;
        temp_of[s2 * kWave + r] = t;
    }
    const uint64_t m = ballot(acc);
    if (r == 0) {
        pair_stats[2 * t] += kWave;
        pair_stats[2 * t + 1] += (unsigned long long)__builtin_popcountll(m);
    }
}

// ---------------------------------------------------------------------------
// Multispin sweep, lane = site.  Model restriction (checked on the host):
// |q_ij| = w0 for every bond and lin_i = -1/2 sum_j q_ij, so that
//   dE_i(replica r) = w0 (k_r - d_i / 2),
//   k_r = #{ j : bit r of (x_i ^ x_j ^ [q_ij < 0]) }.
// A lane counts k for its 64 replicas in B bit planes and accepts dE <= 0
// outright.  For the rest it compares the uniform u_r with the threshold
// thr_r on its kEagerPlanes (eight) leading bits: eight Philox bit planes
// against the threshold's eight leading bit planes, every lane in lockstep, no
// early exit.  A replica whose leading bits tie (probability 2^-8) goes to a
// per-wave queue in LDS; the queue is drained by all 64 lanes at once, each
// comparing the low 24 bits of its entry's tail word with the threshold's low
// bits, and accepted
// entries flip their word with an atomic XOR.  Per-replica energy changes of
// the lockstep flips are summed in bit-sliced counters and turned into
// integers by one 64x64 bit transpose per counter plane; queued flips add
// theirs through an LDS accumulator.
// ---------------------------------------------------------------------------
constexpr uint32_t kQueueCap = 1024; // deferred replicas per wave (any degree)
constexpr uint32_t kRecCap = 128;    // deferred sites per wave (regular lattices)
constexpr uint32_t kMaxPerLane = 8;  // sites per lane, any degree
// Regular lattices: a lane's signed counter (kSW planes) holds 10 terms in [-3, 3].
constexpr uint32_t kMaxPerLaneRegular = 10;

// B: bit planes of the neighbour count (degree < 2^B).  W: width of the
// energy counters.  DEG > 0: every site has exactly DEG neighbours and row p
// starts at p * DEG (regular lattices); DEG == 0: any degree < 2^B.
template <int B, int W, int DEG>
__global__ __launch_bounds__(kBlock, (DEG > 0 ? 5 : 1)) void sweep_multispin_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const uint32_t* __restrict__ sgn, const uint64_t* __restrict__ spins_rd, uint64_t* spins_wr,
    uint32_t n, uint32_t dmax, int32_t w0, const uint32_t* __restrict__ lev_thr,
    const uint32_t* __restrict__ temp_of,
    int64_t* energy, unsigned long long* accepted, uint32_t c_begin, uint32_t c_end,
    uint32_t per_lane, uint32_t span, uint32_t sweep_rel, uint32_t key0, uint32_t key1,
    const uint32_t* __restrict__ clock) {
    static_assert(kMaxPerLane <= 8 && W >= B + 3, "energy counters too narrow");
    const uint32_t sweep = clocked(clock, kClockSweep, sweep_rel);
    extern __shared__ unsigned long long lds_q[];
    // Ties wait in a per-wave LDS queue: one entry per replica for any degree;
    // on a regular lattice one record per site instead (its tied replicas, the
    // ones among them at levels 1 and 2, and where the site is).
    __shared__ uint32_t queue[kWavesPerBlock][DEG > 0 ? 1 : kQueueCap];
    __shared__ unsigned long long rec_und[kWavesPerBlock][DEG > 0 ? kRecCap : 1];
    __shared__ unsigned long long rec_l1[kWavesPerBlock][DEG > 0 ? kRecCap : 1];
    __shared__ unsigned long long rec_l2[kWavesPerBlock][DEG == 6 ? kRecCap : 1];
    __shared__ uint32_t rec_pos[kWavesPerBlock][DEG > 0 ? kRecCap : 1];
    __shared__ uint32_t qcount[kWavesPerBlock];
    __shared__ unsigned long long edef[kWavesPerBlock][kWave];
    __shared__ unsigned long long blk_e[kWave];
    __shared__ uint32_t blk_done, blk_acc;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wv = uniform(threadIdx.x >> 6);
    const uint32_t slot = blockIdx.y;
    // Workgroups go to the 8 XCDs round robin by dispatch order; when the row
    // of blocks divides by 8, XCD k takes the k-th eighth of the class, so a
    // block's neighbour words (the adjacent position ranges) sit in the same
    // L2 as its own.
    const uint32_t gx = gridDim.x;
    const uint32_t bx = (gx & 7u) ? blockIdx.x : (blockIdx.x & 7u) * (gx >> 3) + (blockIdx.x >> 3);
    const uint32_t wave = uniform(bx * kWavesPerBlock + wv);
    // Wave w takes positions [base, wave_end) of the class, span of them
    // (span <= kWave * per_lane); lane l takes base + l + kWave * it.
    const uint32_t base = c_begin + wave * span;
    const uint32_t wave_end = min(c_end, base + span);
    const uint64_t* __restrict__ S = spins_rd + (uint64_t)slot * n;
    // Regular lattices: the first site's row is in flight while the records
    // are built.
    SiteRow<(DEG > 0 ? DEG : 4)> row{};
    if constexpr (DEG > 0) {
        if (base + lane < wave_end) row = load_row<DEG>(S, base + lane, ci, sgn);
    }
    {
        // The slot's threshold records, one per level m <= dmax, kQStride u64
        // each: bit r of plane b (b < kEagerPlanes) is bit 31 - b of thr_r, the
        // threshold of the temperature that lane r of this slot holds; then
        // the mask of nonzero thr_r and the 64 thr_r themselves.
        // With a fixed even degree only the levels m = 2, 4, ..., DEG are read.
        const uint32_t tr = temp_of[slot * kWave + lane];
        constexpr uint32_t m_step = DEG > 0 ? 2 : 1;
        build_records(lds_q, lev_thr, tr, dmax, (DEG > 0 ? 2 : 0) + m_step * wv,
                      m_step * kWavesPerBlock, lane);
        edef[wv][lane] = 0;
        if (lane == 0) qcount[wv] = 0;
        if (wv == 0) blk_e[lane] = 0;
        if (threadIdx.x == 0) blk_done = blk_acc = 0;
        __syncthreads();
    }
    if (base >= c_end) return;
    // waves of this block that have sites (the rest returned above)
    const uint32_t first_wave = bx * kWavesPerBlock;
    const uint32_t active = min(kWavesPerBlock, (c_end - c_begin + span - 1) / span - first_wave);
    uint64_t* Wr = spins_wr + (uint64_t)slot * n;
    const TieRecords tq{rec_und[wv], rec_l1[wv], rec_l2[wv], rec_pos[wv], &qcount[wv], kRecCap};

    // Bit-sliced per-replica sums over this lane's sites.  Any degree:
    // kacc = sum k * flip and dacc = sum d * flip, so dE = w0 (2 kacc - dacc) / 2.
    // Regular lattices (d = DEG even): sacc = sum (k - DEG/2) * flip in kSW-plane
    // two's complement (each term in [-3, 3], at most kMaxPerLaneRegular of them), so
    // dE = w0 sacc.
    static_assert(3 * kMaxPerLaneRegular < (1u << (kSW - 1)), "signed counters too narrow");
    uint64_t kacc[W], dacc[W], sacc[kSW];
#pragma unroll
    for (int b = 0; b < W; ++b) kacc[b] = dacc[b] = 0;
#pragma unroll
    for (int b = 0; b < kSW; ++b) sacc[b] = 0;
    uint32_t nacc = 0;

    if constexpr (DEG > 0) {
        // The next site's row loads while this one updates (the class's sites
        // share no bonds, so no update writes a word another row reads).
        uint32_t p = base + lane;
        for (uint32_t it = 0; it < per_lane && p < wave_end; ++it) {
            const uint32_t pn = p + kWave;
            const bool more = it + 1 < per_lane && pn < wave_end;
            SiteRow<DEG> next = row;
            if (more) next = load_row<DEG>(S, pn, ci, sgn);
            regular_site<DEG>(S, Wr, p, row, lds_q, slot, sweep, key0, key1, w0, sacc, nacc, tq,
                              edef[wv]);
            row = next;
            p = pn;
        }
    } else {
        for (uint32_t it = 0; it < per_lane; ++it) {
            const uint32_t p = base + it * kWave + lane;
            if (p >= wave_end) break;
            irregular_site<B, W>(S, Wr, p, p - base, rp, ci, sgn, lds_q, slot, sweep, key0, key1,
                                 w0, kacc, dacc, nacc, queue[wv], &qcount[wv], kQueueCap,
                                 edef[wv]);
        }
    }
    // Drain the queue with every lane of the wave.
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
    if constexpr (DEG > 0) {
        drain_ties<DEG>(tq, Wr, lds_q, slot, sweep, key0, key1, w0, edef[wv], lane, nacc);
    } else {
        drain_queue(queue[wv], &qcount[wv], kQueueCap, base, Wr, lds_q, slot, sweep, key0, key1,
                    w0, edef[wv], lane, nacc);
    }
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
    int64_t de;
    if constexpr (DEG > 0) {
        de = (int64_t)w0 * wave_column_sum_signed<kSW>(sacc, lane) + (int64_t)edef[wv][lane];
    } else {
        const int64_t ksum = wave_column_sum<W>(kacc, lane);
        const int64_t dsum = wave_column_sum<W>(dacc, lane);
        de = ((int64_t)w0 * (2 * ksum - dsum)) / 2 + (int64_t)edef[wv][lane];
    }
    for (uint32_t off = 32; off > 0; off >>= 1) nacc += __shfl_xor(nacc, (int)off, 64);
    // Combine the block's waves in LDS; the last wave to finish publishes.
    atomicAdd(&blk_e[lane], (unsigned long long)de);
    if (lane == 0) atomicAdd(&blk_acc, nacc);
    __threadfence_block();
    uint32_t done = 0;
    if (lane == 0) done = atomicAdd(&blk_done, 1u);
    done = __shfl(done, 0, 64);
    if (done + 1 == active) {
        __threadfence_block();
        atomicAdd(reinterpret_cast<unsigned long long*>(energy + slot * kWave + lane),
                  __hip_atomic_load(&blk_e[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (lane == 0)
            atomicAdd(accepted, (unsigned long long)__hip_atomic_load(
                                    &blk_acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
    }
}

// One thread per (temperature t, level m), m in [0, dmax]: the threshold of
// dE = m w0 / 2 at beta[t] (0 for m = 0, whose moves are accepted outright).
__global__ void level_thr_kernel(const double* __restrict__ beta, uint32_t* __restrict__ lev_thr,
                                 uint32_t cells, uint32_t dmax, int32_t w0,
                                 const uint32_t* __restrict__ clock) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    beta += (uint64_t)clocked(clock, kClockRow, 0) * cells;
    if (i >= cells * (dmax + 1)) return;
    const uint32_t t = i / (dmax + 1), m = i % (dmax + 1);
    lev_thr[i] = m == 0 ? 0u : metropolis_threshold(beta[t], ((int64_t)m * w0) / 2);
}

__global__ __launch_bounds__(64) void clock_kernel(uint32_t* clock, uint32_t a, uint32_t b,
                                                  uint32_t c, uint32_t add) {
    if (threadIdx.x != 0) return;
    clock[kClockSweep] = (add ? clock[kClockSweep] : 0u) + a;
    clock[kClockRound] = (add ? clock[kClockRound] : 0u) + b;
    clock[kClockRow] = (add ? clock[kClockRow] : 0u) + c;
}

__global__ void identity_perm_kernel(uint32_t* temp_of, uint32_t* slot_of, uint32_t cells,
                                     uint32_t replica_view) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * kWave) return;
    temp_of[i] = replica_view ? i : i / kWave;
    slot_of[i] = replica_view ? i : i / kWave;
}

// Slot-ladder view -> replica view: the configuration at (temperature t,
// lane) becomes replica t * 64 + lane, so replica identities survive the
// switch.  One thread per position; rho is a bijection, so no two threads
// write the same entry.
__global__ void replica_index_perm_kernel(uint32_t* temp_of, uint32_t* slot_of, uint32_t cells) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cells * kWave) return;
    const uint32_t rho = temp_of[q] * kWave + (q & (kWave - 1));
    temp_of[q] = rho;
    slot_of[rho] = q;
}

// ReplicaExchange round in the replica view (replica_exchange.rs:61-109): one
// thread per pair (i, i + 1), i = phase + 2 k.  The pairs of a round are
// disjoint, so each thread owns its two replicas and its statistics.
__global__ void exchange_replicas_kernel(const double* __restrict__ beta,
                                         const int64_t* __restrict__ energy, uint32_t* temp_of,
                                         uint32_t* slot_of, unsigned long long* pair_stats,
                                         uint32_t replicas, uint32_t phase, uint32_t round,
                                         uint32_t key0, uint32_t key1) {
    const uint32_t i = phase + 2u * (blockIdx.x * blockDim.x + threadIdx.x);
    if (i + 1 >= replicas) return;
    const uint32_t qi = slot_of[i], qj = slot_of[i + 1];
    const double delta = (beta[i] - beta[i + 1]) * (double)(energy[qi] - energy[qj]);
    const U32x4 u = philox4x32_10(i >> 2, 0u, round, kTagRepx << 24, key0, key1);
    const uint32_t w = i & 3u;
    const uint32_t uu = w == 0 ? u.v[0] : w == 1 ? u.v[1] : w == 2 ? u.v[2] : u.v[3];
    const bool acc = delta >= 0.0 || uu < exchange_threshold(delta);
    pair_stats[2ull * i] += 1;
    if (acc) {
        pair_stats[2ull * i + 1] += 1;
        slot_of[i] = qj;
        slot_of[i + 1] = qi;
        temp_of[qi] = i + 1;
        temp_of[qj] = i;
    }
}

__global__ void init_spins_kernel(const uint32_t* __restrict__ order, uint64_t* spins, uint32_t n,
                                  uint32_t cells, uint32_t key0, uint32_t key1) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * cells) return;
    const uint32_t slot = (uint32_t)(i / n);
    const uint32_t p = (uint32_t)(i % n);
    const U32x4 o = philox4x32_10(order[p], slot, 0u, kTagInit << 24, key0, key1);
    spins[i] = (uint64_t)o.v[0] | ((uint64_t)o.v[1] << 32);
}

// Energy from scratch, lane = replica: A = sum_i x_i (lin_i + h_i) with
// h_i = lin_i + sum_j q_ij x_j, so that E = offset + A / 2 exactly.
// Workgroup b serves slot b / blocks_per_slot; its wave w the `chunk`
// positions from ((b % blocks_per_slot) * 4 + w) * chunk, in groups of eight
// whose rows are gathered together.  The workgroup's per-lane partial sums
// are added into acc[slot][lane] (zeroed beforehand; two's-complement adds of
// signed partials), then energy_finish turns A into E.
constexpr int kEnergyGroup = 8;
constexpr int kEnergyInterleave = 1;

__global__ __launch_bounds__(kBlock) void energy_partial_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin,
    const uint64_t* __restrict__ spins, uint32_t n, uint32_t chunk, uint32_t blocks_per_slot,
    unsigned long long* __restrict__ acc) {
    __shared__ int64_t part[kWavesPerBlock][kWave];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wv = uniform(threadIdx.x >> 6);
    const uint32_t slot = blockIdx.x / blocks_per_slot;
    const uint64_t p0_64 = ((uint64_t)(blockIdx.x % blocks_per_slot) * kWavesPerBlock + wv) * chunk;
    const uint32_t p0 = (uint32_t)min<uint64_t>(p0_64, n);
    const uint32_t p1 = min(n, p0 + chunk);
    const uint64_t* S = spins + (uint64_t)slot * n;
    int64_t a = 0;
    fields_groups<kEnergyGroup, kEnergyInterleave>(
        rp, ci, wt, lin, reinterpret_cast<const uint32_t*>(lin), S, p0, p1, lane,
        [&](uint32_t, uint64_t x, int32_t h, uint32_t l) {
            a += lane_bit(x) ? (int64_t)h + (int32_t)l : 0;
        });
    part[wv][lane] = a;
    __syncthreads();
    if (wv == 0) {
        int64_t t = 0;
#pragma unroll
        for (uint32_t w = 0; w < kWavesPerBlock; ++w) t += part[w][lane];
        if (t) atomicAdd(acc + (uint64_t)slot * kWave + lane, (unsigned long long)t);
    }
}

__global__ void energy_zero_kernel(int64_t* out, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = 0;
}

__global__ void energy_finish_kernel(int64_t* out, uint32_t count, int64_t offset) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = offset + out[i] / 2;
}

__global__ __launch_bounds__(kBlock) void fields_kernel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
    const int32_t* __restrict__ wt, const int32_t* __restrict__ lin,
    const uint32_t* __restrict__ order, const uint64_t* __restrict__ spins, uint32_t n,
    uint32_t cells, int32_t* out) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t w = uniform(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    if (w >= (uint64_t)n * cells) return;
    const uint32_t slot = (uint32_t)(w / n);
    const uint32_t p = (uint32_t)(w % n);
    const uint64_t* S = spins + (uint64_t)slot * n;
    int32_t h = lin[p];
    for (uint32_t e = rp[p]; e < rp[p + 1]; ++e) h += lane_bit(S[ci[e]]) ? wt[e] : 0;
    out[((uint64_t)slot * n + order[p]) * kWave + lane] = h;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_clock_set(uint32_t* clock, uint32_t sweep, uint32_t round, uint32_t row,
                            hipStream_t stream) {
    hipLaunchKernelGGL(clock_kernel, dim3(1), dim3(64), 0, stream, clock, sweep, round, row, 0u);
    return hipGetLastError();
}

hipError_t launch_clock_advance(uint32_t* clock, uint32_t dsweep, uint32_t dround, uint32_t drow,
                                hipStream_t stream) {
    hipLaunchKernelGGL(clock_kernel, dim3(1), dim3(64), 0, stream, clock, dsweep, dround, drow,
                       1u);
    return hipGetLastError();
}

hipError_t launch_thr_table(const double* beta, uint32_t* thr, uint32_t cells, uint32_t dtab,
                            hipStream_t stream, const uint32_t* clock) {
    const uint64_t total = ((uint64_t)dtab + 1) * cells;
    hipLaunchKernelGGL(thr_table_kernel, dim3(ceil_div(total, kBlock)), dim3(kBlock), 0, stream,
                       beta, thr, cells, dtab, clock);
    return hipGetLastError();
}

hipError_t launch_sweep_generic(const DevModel& m, const DevState& s, uint32_t c_begin,
                                uint32_t c_end, uint32_t sweep, uint64_t seed,
                                const uint32_t* thr, uint32_t dtab, hipStream_t stream) {
    if (c_end <= c_begin) return hipSuccess;
    const uint32_t len = c_end - c_begin;
    const uint64_t chunk = lane_per_replica_chunk(len, s.cells);
    const uint32_t waves_per_slot = ceil_div(len, chunk);
    const uint64_t waves = (uint64_t)waves_per_slot * s.cells;
    const dim3 grid(ceil_div(waves, kWavesPerBlock));
    const uint32_t temps = s.temps ? s.temps : s.cells;
    const uint64_t lds_bytes = (uint64_t)temps * (dtab + 1) * sizeof(uint32_t);
    const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);
    if (lds_bytes <= 32768) {
        hipLaunchKernelGGL(sweep_generic_kernel<true>, grid, dim3(kBlock), (uint32_t)lds_bytes,
                           stream, m.rp, m.ci, m.wt, m.lin, s.spins, s.spins, m.n, s.cells,
                           temps, s.temp_of, thr, dtab, s.energy, s.accepted, c_begin, c_end,
                           (uint32_t)chunk, waves_per_slot, sweep, key0, key1, s.clock);
    } else {
        hipLaunchKernelGGL(sweep_generic_kernel<false>, grid, dim3(kBlock), 0, stream, m.rp,
                           m.ci, m.wt, m.lin, s.spins, s.spins, m.n, s.cells, temps, s.temp_of,
                           thr, dtab, s.energy, s.accepted, c_begin, c_end, (uint32_t)chunk,
                           waves_per_slot, sweep, key0, key1, s.clock);
    }
    return hipGetLastError();
}

hipError_t launch_heatbath_table(const double* beta, uint64_t* hb, uint32_t cells, uint32_t dtab,
                                 hipStream_t stream, const uint32_t* clock) {
    const uint64_t total = (2ull * dtab + 1) * cells;
    hipLaunchKernelGGL(heatbath_table_kernel, dim3(ceil_div(total, kBlock)), dim3(kBlock), 0,
                       stream, beta, hb, cells, dtab, clock);
    return hipGetLastError();
}

hipError_t launch_sweep_heatbath(const DevModel& m, const DevState& s, uint32_t c_begin,
                                 uint32_t c_end, uint32_t sweep, uint64_t seed, const uint64_t* hb,
                                 uint32_t dtab, hipStream_t stream) {
    if (c_end <= c_begin) return hipSuccess;
    const uint32_t len = c_end - c_begin;
    const uint64_t chunk = lane_per_replica_chunk(len, s.cells);
    const uint32_t waves_per_slot = ceil_div(len, chunk);
    const uint64_t waves = (uint64_t)waves_per_slot * s.cells;
    const dim3 grid(ceil_div(waves, kWavesPerBlock));
    const uint32_t temps = s.temps ? s.temps : s.cells;
    const uint64_t lds_bytes = (uint64_t)temps * (2ull * dtab + 1) * sizeof(uint64_t);
    const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);
    if (lds_bytes <= 32768) {
        hipLaunchKernelGGL(sweep_heatbath_kernel<true>, grid, dim3(kBlock), (uint32_t)lds_bytes,
                           stream, m.rp, m.ci, m.wt, m.lin, s.spins, s.spins, m.n, s.cells,
                           temps, s.temp_of, hb, dtab, s.energy, s.accepted, c_begin, c_end,
                           (uint32_t)chunk, waves_per_slot, sweep, key0, key1, s.clock);
    } else {
        hipLaunchKernelGGL(sweep_heatbath_kernel<false>, grid, dim3(kBlock), 0, stream, m.rp,
                           m.ci, m.wt, m.lin, s.spins, s.spins, m.n, s.cells, temps, s.temp_of,
                           hb, dtab, s.energy, s.accepted, c_begin, c_end, (uint32_t)chunk,
                           waves_per_slot, sweep, key0, key1, s.clock);
    }
    return hipGetLastError();
}

// Workgroups of one sweep_multispin_kernel instantiation resident per CU at
// once (occupancy query, once per process and dynamic LDS size) times the CU
// count of the current device: the waves of one round.
template <int B, int W, int DEG>
static uint32_t resident_waves(uint32_t lds) {
    static std::atomic<uint64_t> per_cu{0};  // (lds << 32) | blocks, 0 = not asked yet
    uint64_t v = per_cu.load(std::memory_order_relaxed);
    if (v == 0 || (uint32_t)(v >> 32) != lds) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, reinterpret_cast<const void*>(sweep_multispin_kernel<B, W, DEG>), kBlock,
                lds) != hipSuccess || nb < 1)
            nb = 0;
        v = ((uint64_t)lds << 32) | (uint32_t)nb;
        per_cu.store(v, std::memory_order_relaxed);
    }
    constexpr int kDevices = 64;
    static std::atomic<int> cus[kDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevices) return 0;
    int ncu = cus[dev].load(std::memory_order_relaxed);
    if (ncu == 0) {
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            ncu < 1)
            return 0;
        cus[dev].store(ncu, std::memory_order_relaxed);
    }
    return (uint32_t)v * (uint32_t)ncu * kWavesPerBlock;
}

template <int B, int W, int DEG>
static void launch_multispin_grid(const DevModel& m, const DevState& s, uint32_t c_begin,
                                  uint32_t c_end, uint32_t sweep, uint64_t seed,
                                  const uint32_t* lev_thr, hipStream_t stream) {
    const uint32_t len = c_end - c_begin;
    const uint32_t lds = (m.dmax + 1) * kQStride * sizeof(unsigned long long);
    // Sites per lane: about 8192 waves (eight per SIMD) and at most
    // kMaxPerLane so the bit-sliced energy counters stay narrow; wave w takes
    // kWave * per_lane consecutive positions.
    uint64_t per_lane = ((uint64_t)len * s.cells + (uint64_t)kWave * 8192 - 1) /
                        ((uint64_t)kWave * 8192);
    if (per_lane < 1) per_lane = 1;
    if (per_lane > kMaxPerLane) per_lane = kMaxPerLane;
    uint32_t span = kWave * (uint32_t)per_lane;
    uint32_t waves = ceil_div(len, span);
    if (DEG > 0) {
        // A regular lattice whose waves would not all be resident at once
        // takes one round instead: each wave a contiguous span of up to
        // kWave * kMaxPerLaneRegular positions, a multiple of 8 workgroups per
        // slot (the XCD eighths).  Every wave then starts together and none
        // runs alone behind a partial first round; the sites of a class are
        // independent, so the split does not change any result.
        // It is taken when the grid above would lose the XCD eighths (a row
        // of workgroups not a multiple of 8) or when it costs no more wave
        // iterations (waves x sites per lane); on 3D L=64 x 32 slots the grid
        // above packs every lane with 8 sites and measured faster (56.0 us)
        // than one round of 10-site lanes (58.2 us), while 28 slots went from
        // 56.2 to 50.0 us.
        const uint32_t round = resident_waves<B, W, DEG>(lds);
        constexpr uint32_t kUnit = 8 * kWavesPerBlock;
        if (round > 0 && (uint64_t)waves * s.cells > round) {
            const uint32_t per_slot = round / s.cells / kUnit * kUnit;
            if (per_slot > 0) {
                const uint32_t sp = ceil_div(len, per_slot);
                const uint32_t pl = ceil_div(sp, kWave);
                const uint32_t nw = ceil_div(len, sp);
                const bool eighths = ceil_div(waves, kWavesPerBlock) % 8u == 0;
                if (pl <= kMaxPerLaneRegular &&
                    (!eighths || (uint64_t)nw * pl <= (uint64_t)waves * per_lane)) {
                    span = sp;
                    per_lane = pl;
                    waves = nw;
                }
            }
        }
    }
    const dim3 grid(ceil_div(waves, kWavesPerBlock), s.cells);
    const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);
    hipLaunchKernelGGL((sweep_multispin_kernel<B, W, DEG>), grid, dim3(kBlock), lds, stream, m.rp,
                       m.ci, m.sgn, s.spins, s.spins, m.n, m.dmax, m.w0, lev_thr, s.temp_of,
                       s.energy, s.accepted, c_begin, c_end, (uint32_t)per_lane, span, sweep,
                       key0, key1, s.clock);
}

void prepare_sweep_multispin(const DevModel& m) {
    const uint32_t lds = (m.dmax + 1) * kQStride * sizeof(unsigned long long);
    if (m.regular_degree == 6)
        (void)resident_waves<3, 6, 6>(lds);
    else if (m.regular_degree == 4)
        (void)resident_waves<3, 6, 4>(lds);
}

hipError_t launch_sweep_multispin(const DevModel& m, const DevState& s, uint32_t c_begin,
                                  uint32_t c_end, uint32_t sweep, uint64_t seed,
                                  const uint32_t* lev_thr, hipStream_t stream) {
    if (c_end <= c_begin) return hipSuccess;
    // Counter width W = B + 3: at most kMaxPerLane = 8 sites per lane, each
    // adding k * flip <= 2^B - 1, so every sum stays below 2^(B+3).
    if (m.regular_degree == 6) {
        launch_multispin_grid<3, 6, 6>(m, s, c_begin, c_end, sweep, seed, lev_thr, stream);
    } else if (m.regular_degree == 4) {
        launch_multispin_grid<3, 6, 4>(m, s, c_begin, c_end, sweep, seed, lev_thr, stream);
    } else if (m.dmax < 8) {
        launch_multispin_grid<3, 6, 0>(m, s, c_begin, c_end, sweep, seed, lev_thr, stream);
    } else if (m.dmax < 16) {
        launch_multispin_grid<4, 7, 0>(m, s, c_begin, c_end, sweep, seed, lev_thr, stream);
    } else if (m.dmax < 32) {
        launch_multispin_grid<5, 8, 0>(m, s, c_begin, c_end, sweep, seed, lev_thr, stream);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_level_thresholds(const DevModel& m, const DevState& s, const double* beta,
                                   uint32_t* lev_thr, hipStream_t stream) {
    const uint32_t temps = s.temps ? s.temps : s.cells;
    const uint32_t total = temps * (m.dmax + 1);
    hipLaunchKernelGGL(level_thr_kernel, dim3(ceil_div(total, kBlock)), dim3(kBlock), 0, stream,
                       beta, lev_thr, temps, m.dmax, m.w0, s.clock);
    return hipGetLastError();
}

hipError_t launch_exchange(const DevState& s, const double* beta, uint32_t round, uint64_t seed,
                           hipStream_t stream) {
    const uint32_t T = s.pop_temps ? s.pop_temps : s.cells;  // levels per population
    if (T < 2 || s.cells % T) return s.cells % T ? hipErrorInvalidValue : hipSuccess;
    // per population t = phase, phase+2, ... with t+1 < T; under a clock the
    // phase is the device's, so cover phase 0, which has the most pairs
    const uint32_t phase = s.clock ? 0u : (round & 1u);
    const uint32_t per_pop = (T - phase) / 2;
    if (per_pop == 0) return hipSuccess;
    const uint32_t pairs = per_pop * (s.cells / T);
    hipLaunchKernelGGL(exchange_kernel, dim3(pairs), dim3(kWave), 0, stream, beta, s.energy,
                       s.temp_of, s.slot_of, s.pair_stats, s.cells, T, per_pop, round,
                       (uint32_t)seed, (uint32_t)(seed >> 32), s.clock);
    return hipGetLastError();
}

hipError_t launch_identity_perm(const DevState& s, hipStream_t stream) {
    hipLaunchKernelGGL(identity_perm_kernel, dim3(ceil_div((uint64_t)s.cells * kWave, kBlock)),
                       dim3(kBlock), 0, stream, s.temp_of, s.slot_of, s.cells, s.replica_view);
    return hipGetLastError();
}

hipError_t launch_replica_index_perm(const DevState& s, hipStream_t stream) {
    hipLaunchKernelGGL(replica_index_perm_kernel,
                       dim3(ceil_div((uint64_t)s.cells * kWave, kBlock)), dim3(kBlock), 0, stream,
                       s.temp_of, s.slot_of, s.cells);
    return hipGetLastError();
}

hipError_t launch_exchange_replicas(const DevState& s, const double* beta, uint32_t phase,
                                    uint32_t round, uint64_t seed, hipStream_t stream) {
    const uint32_t replicas = s.cells * kWave;
    const uint32_t pairs = (replicas - (phase & 1u)) / 2;
    if (pairs == 0) return hipSuccess;
    hipLaunchKernelGGL(exchange_replicas_kernel, dim3(ceil_div(pairs, kBlock)), dim3(kBlock), 0,
                       stream, beta, s.energy, s.temp_of, s.slot_of, s.pair_stats, replicas,
                       phase & 1u, round, (uint32_t)seed, (uint32_t)(seed >> 32));
    return hipGetLastError();
}

hipError_t launch_init_spins(const DevModel& m, const DevState& s, uint64_t seed,
                             hipStream_t stream) {
    const uint64_t total = (uint64_t)m.n * s.cells;
    hipLaunchKernelGGL(init_spins_kernel, dim3(ceil_div(total, kBlock)), dim3(kBlock), 0, stream,
                       m.order, s.spins, m.n, s.cells, (uint32_t)seed, (uint32_t)(seed >> 32));
    return hipGetLastError();
}

hipError_t launch_energy(const DevModel& m, const DevState& s, int64_t* out, hipStream_t stream) {
    const uint32_t count = s.cells * kWave;
    hipLaunchKernelGGL(energy_zero_kernel, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, stream,
                       out, count);
    // About 16K waves, positions per wave a multiple of one interleaved gather.
    const uint64_t unit = kEnergyGroup * kEnergyInterleave;
    uint64_t chunk = ((uint64_t)m.n * s.cells + 16383) / 16384;
    chunk = (chunk + unit - 1) / unit * unit;
    if (chunk < unit) chunk = unit;
    if (chunk > 4096) chunk = 4096;
    const uint32_t blocks_per_slot = ceil_div(m.n, chunk * kWavesPerBlock);
    hipLaunchKernelGGL(energy_partial_kernel, dim3((uint32_t)((uint64_t)blocks_per_slot * s.cells)),
                       dim3(kBlock), 0, stream, m.rp, m.ci, m.wt, m.lin, s.spins, m.n,
                       (uint32_t)chunk, blocks_per_slot, reinterpret_cast<unsigned long long*>(out));
    hipLaunchKernelGGL(energy_finish_kernel, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, stream,
                       out, count, m.offset);
    return hipGetLastError();
}

hipError_t launch_fields(const DevModel& m, const DevState& s, int32_t* out, hipStream_t stream) {
    const uint64_t waves = (uint64_t)m.n * s.cells;
    hipLaunchKernelGGL(fields_kernel, dim3(ceil_div(waves, kWavesPerBlock)), dim3(kBlock), 0,
                       stream, m.rp, m.ci, m.wt, m.lin, m.order, s.spins, m.n, s.cells, out);
    return hipGetLastError();
}

}  // namespace ige
