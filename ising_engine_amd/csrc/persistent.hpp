// persistent.hpp -- the persistent multi-instance kernel shared by the batch
// API and the serving loop.
//
// An instance is either an anneal (one cell of 64 replicas on a geometric beta
// schedule) or a parallel-tempering solve (UltimateSolver's static-ladder
// loop, src/solver/ultimate.rs:234-336: `cells` cells of 64 replicas on a fixed
// ladder, a DEO exchange after every sweep, the incumbent updated every
// track_every sweeps), each with its own CSR (position space) and seed.  A workgroup pulls the next instance from
// a device work queue, advances it by up to `round` sweeps from its stored
// progress, with workgroup barriers between colour classes, and pulls again:
// no grid-wide synchronisation.  Per-sweep Metropolis thresholds are rebuilt
// by the workgroup in LDS (or in a per-instance global scratch table when the
// field range does not fit).
//
// Uniforms are keyed by (instance seed, position, slot 0, sweep), so an
// instance's trajectory does not depend on the workgroup, the co-resident
// instances, the round boundaries or the launch mode that ran it.
#pragma once

#include <hip/hip_runtime.h>

#include "gang_sync.h"
#include "ige_math.h"
#include "kernels.hpp"
#include "sweep_generic.h"

namespace ige {

struct InstDesc {
    const uint32_t* rp;    // n+1 absolute indices into ci / wt
    const uint32_t* ci;    // neighbour positions
    const int32_t* wt;     // couplings
    const int32_t* lin;    // n linear terms
    const uint32_t* cs;    // ncol+1 colour class starts
    uint64_t* spins;       // n MSC words (position order)
    int64_t* energy;       // 64 tracked energies
    unsigned long long* accepted;
    uint32_t* progress;    // sweeps done so far
    uint32_t n, ncol, sweeps, dtab, key0, key1;
    uint32_t round;        // sweeps to advance in this launch (0: the launch's default)
    uint32_t thr_mode;     // kThrLds or kThrBand
    uint32_t gang;         // workgroups cooperating on this instance (1..kMaxGang)
    uint32_t split;        // PT: members per cell (1: each member owns whole cells)
    uint32_t* bar;         // gang barrier counter, zero at launch start (gang > 1)
    uint64_t* ticks;       // if set: wall-clock ticks member 0 spent on this launch's sweeps
    BestRecord* best;      // if set with track_every: the best-seen replica's record
    uint8_t* best_x;       // n bytes, position order: its spins
    uint32_t track_every;  // incumbent update after every track_every-th sweep and the last
    double beta_hi;
    double c;  // beta_k = beta_hi * det_exp(-(sweeps - 1 - k) * c)
    int64_t offset;
    // Parallel tempering (pt != 0): spins [cells][n], energy [cells][64];
    // lane r of temperature index t sits in slot slot_of[t * 64 + r], and
    // temp_of is its inverse per lane, as in a single state.
    uint32_t cells;         // 1 for an anneal
    uint32_t pt;            // 1: fixed ladder, DEO exchange after every sweep
    const double* ladder;   // [cells] betas, hottest first
    uint32_t* temp_of;      // [cells][64]
    uint32_t* slot_of;      // [cells][64]
    // Deadline rounds (PT, round_ticks > 0): the solve sweeps until its member
    // 0 has spent round_ticks of device wall clock in this launch (at least
    // one sweep, at most `round`), deciding after each sweep through *stop;
    // *swept receives the sweeps done.
    uint64_t round_ticks;
    uint32_t* stop;
    uint32_t* swept;
};

constexpr uint32_t kMaxPtCells = 64;  // temperature cells per solve

constexpr uint32_t kLdsThrEntries = 4096;
constexpr uint32_t kMaxGang = 128;       // queue codes carry the member index in 7 bits
constexpr uint32_t kGangBits = 7;
constexpr uint32_t kGangSites = 2048;    // sites per workgroup before a gang is formed
// How a workgroup decides u < thr(beta, dE) for dE <= dtab.
constexpr uint32_t kThrLds = 0;   // dtab < kLdsThrEntries: per-sweep table in LDS
constexpr uint32_t kThrBand = 1;  // larger ranges: threshold_band, no table

__host__ __device__ __forceinline__ uint32_t choose_thr_mode(uint32_t dtab, uint32_t rows = 1) {
    return (uint64_t)rows * ((uint64_t)dtab + 1) <= kLdsThrEntries ? kThrLds : kThrBand;
}

__host__ __device__ __forceinline__ double anneal_beta(double beta_hi, double c, uint32_t sweeps,
                                                       uint32_t k) {
    return sweeps > 1 ? beta_hi * det_exp(-(double)(sweeps - 1 - k) * c) : beta_hi;
}

__host__ __device__ __forceinline__ uint32_t choose_gang(uint32_t n, uint32_t max_gang,
                                                        uint32_t gang_sites) {
    const uint32_t g = n / (gang_sites ? gang_sites : kGangSites);
    const uint32_t cap = max_gang < kMaxGang ? max_gang : kMaxGang;
    return g < 2 ? 1u : (g < cap ? g : cap);
}

// Incumbent update of UltimateSolver::update_incumbent (src/solver/ultimate.rs:397-431)
// after `sweep` sweeps, for one member of the instance's gang.  The member's
// energy changes are published first, so every member then reads the same 64
// energies and takes the same decision: lowest energy, lowest lane on a tie,
// replacing the record only when strictly lower.  A barrier separates every
// member's read of the record from member 0's write.  Each member copies the
// winning lane's bits of exactly the positions it updates in a sweep, so no
// other member can change them first.
template <int WAVES>
__device__ __forceinline__ void track_point(const InstDesc& I, uint32_t sweep, uint32_t member,
                                            uint32_t gang, uint32_t lane, uint32_t wave,
                                            int64_t& eacc, int64_t (*part)[64],
                                            uint32_t& bar_target, uint32_t spin_limit,
                                            uint32_t* err) {
    part[wave][lane] = eacc;
    eacc = 0;
    __syncthreads();
    if (wave == 0) {
        int64_t tot = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) tot += part[w][lane];
        if (gang > 1)
            atomicAdd(reinterpret_cast<unsigned long long*>(I.energy + lane),
                      (unsigned long long)tot);
        else
            I.energy[lane] += tot;
    }
    if (gang > 1)
        gang_barrier(I.bar, bar_target, gang, spin_limit, err);
    else
        __syncthreads();
    int64_t e = I.energy[lane];
    uint32_t l = lane;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t elo = __shfl_xor((uint32_t)e, off), ehi = __shfl_xor((uint32_t)(e >> 32), off);
        const int64_t e2 = (int64_t)(((uint64_t)ehi << 32) | elo);
        const uint32_t l2 = __shfl_xor(l, off);
        if (e2 < e || (e2 == e && l2 < l)) {
            e = e2;
            l = l2;
        }
    }
    const bool take = !I.best->valid || e < I.best->energy;
    if (gang > 1)
        gang_barrier(I.bar, bar_target, gang, spin_limit, err);
    else
        __syncthreads();
    if (!take) return;
    const uint32_t stride = gang * WAVES;
    const uint32_t lane_of_gang = member * WAVES + wave;
    for (uint32_t c = 0; c < I.ncol; ++c) {
        const uint32_t cb = I.cs[c], ce = I.cs[c + 1];
        if (ce <= cb) continue;
        for (uint32_t g = (cb >> 2) + lane_of_gang; g <= ((ce - 1) >> 2); g += stride) {
            const uint32_t p0 = max(g * 4u, cb), p1 = min(g * 4u + 4u, ce);
            if (lane < p1 - p0) I.best_x[p0 + lane] = (uint8_t)((I.spins[p0 + lane] >> l) & 1u);
        }
    }
    if (member == 0 && threadIdx.x == 0) {
        I.best->energy = e;
        I.best->valid = 1;
        I.best->sweep = sweep;
        I.best->temp = 0;
        I.best->lane = l;
    }
}

// Sweeps [k0, k1) of one instance by one member of its gang, with the
// thresholds of mode MODE (kThrLds: a per-sweep table in thr_lds; kThrBand:
// threshold_band, no table).  Adds the member's energy changes to eacc (per
// lane) and its accepted count to nacc.
template <int WAVES, uint32_t MODE>
__device__ __forceinline__ void anneal_sweeps(const InstDesc& I, uint32_t k0, uint32_t k1,
                                              uint32_t member, uint32_t gang, uint32_t lane,
                                              uint32_t wave, uint32_t* thr_lds,
                                              int64_t (*part)[64], uint32_t& bar_target,
                                              uint32_t spin_limit, uint32_t* err, int64_t& eacc,
                                              uint32_t& nacc) {
    for (uint32_t k = k0; k < k1; ++k) {
        const double beta = anneal_beta(I.beta_hi, I.c, I.sweeps, k);
        const uint32_t dlim = threshold_limit(beta, I.dtab);
        if constexpr (MODE == kThrLds) {
            for (uint32_t d = threadIdx.x; d <= dlim; d += WAVES * 64)
                thr_lds[d] = d == 0 ? 0xFFFFFFFFu : metropolis_threshold(beta, (int64_t)d);
            __syncthreads();
        }
        const double nbd = -beta * 1.4426950408889634;
        auto accept = [&](uint32_t uu, int32_t de) -> bool {
            if ((uint32_t)de > dlim) return false;
            if constexpr (MODE == kThrLds)
                return uu < thr_lds[de];
            else
                return band_accept(uu, de, nbd, beta);
        };
        const uint32_t stride = gang * WAVES;
        const uint32_t lane_of_gang = member * WAVES + wave;
        for (uint32_t c = 0; c < I.ncol; ++c) {
            const uint32_t cb = __builtin_amdgcn_readfirstlane(I.cs[c]);
            const uint32_t ce = __builtin_amdgcn_readfirstlane(I.cs[c + 1]);
            if (ce > cb) {
                for (uint32_t g = (cb >> 2) + lane_of_gang; g <= ((ce - 1) >> 2); g += stride) {
                    const uint32_t p0 = max(g * 4u, cb), p1 = min(g * 4u + 4u, ce);
                    generic_group(I.rp, I.ci, I.wt, I.lin, I.spins, I.spins, p0, p1, 0u, k,
                                  I.key0, I.key1, lane, accept, eacc, nacc);
                }
            }
            if (gang > 1)
                gang_barrier(I.bar, bar_target, gang, spin_limit, err);
            else
                __syncthreads();
        }
        if (I.track_every && ((k + 1) % I.track_every == 0 || k + 1 == I.sweeps))
            track_point<WAVES>(I, k + 1, member, gang, lane, wave, eacc, part, bar_target,
                               spin_limit, err);
    }
}

// Barrier of a gang of one workgroup: the solve's data is written and read by
// this workgroup only within a launch (the waves share the CU's L1, and the
// energy atomics and load_agent reads meet at L2), so the workgroup barrier
// orders it; device-scope fences here (an L2 write-back and invalidation per
// wave, twice a sweep) cost the serving loop's PT solves several microseconds
// a sweep.
__device__ __forceinline__ void block_sync_device() { __syncthreads(); }

__device__ __forceinline__ int64_t load_agent(const int64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The DEO exchange after sweep number `round` (engine.rs:713-818, the
// exchange_kernel of a single state): pairs (t, t + 1) with t = round mod 2;
// lane r of temperature t against lane r of t + 1; accept iff
// delta = (beta_t - beta_{t+1}) (E_t - E_{t+1}) >= 0 or u < floor(e^delta 2^32),
// u = Philox(t, r >> 2, round, TAG_EXCH)[r & 3]; accepted lanes swap
// temperature indices.  One wave; the pairs are disjoint.
__device__ __forceinline__ void pt_exchange(const InstDesc& I, uint32_t round, uint32_t lane) {
    for (uint32_t t = round & 1u; t + 1 < I.cells; t += 2) {
        const uint32_t s1 = I.slot_of[t * 64 + lane], s2 = I.slot_of[(t + 1) * 64 + lane];
        const double delta =
            (I.ladder[t] - I.ladder[t + 1]) *
            (double)(load_agent(I.energy + s1 * 64 + lane) - load_agent(I.energy + s2 * 64 + lane));
        const U32x4 u = philox4x32_10(t, lane >> 2, round, kTagExch << 24, I.key0, I.key1);
        const uint32_t q = lane & 3u;
        const uint32_t uu = q == 0 ? u.v[0] : q == 1 ? u.v[1] : q == 2 ? u.v[2] : u.v[3];
        if (delta >= 0.0 || uu < exchange_threshold(delta)) {
            I.slot_of[t * 64 + lane] = s2;
            I.slot_of[(t + 1) * 64 + lane] = s1;
            I.temp_of[s1 * 64 + lane] = t + 1;
            I.temp_of[s2 * 64 + lane] = t;
        }
    }
}

// UltimateSolver::update_incumbent (ultimate.rs:397-431) after `sweep` sweeps
// of a PT solve: the lowest tracked energy, ties to the lowest (temperature
// index, lane), replaces the record when strictly lower.  Every member takes
// the same decision from the same energies; a barrier separates the reads of
// the record from member 0's write, and another one the copy of the winner's
// bits (each member a share of the positions) from the next sweep.
template <int WAVES>
__device__ __forceinline__ void pt_track_point(const InstDesc& I, uint32_t sweep, uint32_t member,
                                               uint32_t gang, uint32_t lane, uint32_t wave,
                                               uint32_t& bar_target, uint32_t spin_limit,
                                               uint32_t* err) {
    int64_t e = 0;
    uint32_t key = 0xFFFFFFFFu, q = 0;
    for (uint32_t s = 0; s < I.cells; ++s) {
        const int64_t es = load_agent(I.energy + s * 64 + lane);
        const uint32_t ks = I.temp_of[s * 64 + lane] * 64 + lane;
        if (key == 0xFFFFFFFFu || es < e || (es == e && ks < key)) {
            e = es;
            key = ks;
            q = s * 64 + lane;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t elo = __shfl_xor((uint32_t)e, off), ehi = __shfl_xor((uint32_t)(e >> 32), off);
        const int64_t e2 = (int64_t)(((uint64_t)ehi << 32) | elo);
        const uint32_t k2 = __shfl_xor(key, off), q2 = __shfl_xor(q, off);
        if (e2 < e || (e2 == e && k2 < key)) {
            e = e2;
            key = k2;
            q = q2;
        }
    }
    const bool take = !I.best->valid || e < I.best->energy;
    if (gang > 1)
        gang_barrier(I.bar, bar_target, gang, spin_limit, err);
    else
        block_sync_device();
    if (take) {
        const uint64_t* src = I.spins + (uint64_t)(q / 64) * I.n;
        const uint32_t l = q % 64;
        for (uint32_t p = (member * WAVES + wave) * 64 + lane; p < I.n; p += gang * WAVES * 64)
            I.best_x[p] = (uint8_t)((src[p] >> l) & 1u);
        if (member == 0 && threadIdx.x == 0) {
            I.best->energy = e;
            I.best->valid = 1;
            I.best->sweep = sweep;
            I.best->temp = key / 64;
            I.best->lane = l;
        }
    }
    if (gang > 1)
        gang_barrier(I.bar, bar_target, gang, spin_limit, err);
    else
        block_sync_device();
}

// Sweeps [k0, k1) of a PT solve by one member of its gang.  The gang's
// members form gang / split owners of split members each; owner o sweeps
// cells o, o + owners, ... (cells are independent within a sweep) class by
// class, its split members taking every split-th group of positions of a
// class, with a workgroup barrier between classes (split 1) or a gang barrier
// (split > 1: every owner then runs the same number of cell iterations, idle
// on a missing cell, so that the gang's barriers line up).  Each lane accepts
// with its own temperature (temp_of, ladder), and each member adds its energy
// changes to the cell's tracked energies by device-scope atomics.  After the
// sweep: a barrier across the gang, the exchange (member 0, wave 0), another
// barrier, and every track_every sweeps the incumbent.  The sweep's uniforms
// are keyed by (position, cell, sweep), its exchange by the sweep number: a
// single state with the same seed, ladder and IGE_RUN_EXCHANGE follows the
// same path.
// Returns the end of the sweeps done (k1, or earlier at a deadline).
template <int WAVES, uint32_t MODE>
__device__ __forceinline__ uint32_t pt_sweeps(const InstDesc& I, uint32_t k0, uint32_t k1,
                                              uint32_t member, uint32_t gang, uint32_t lane,
                                              uint32_t wave, uint32_t* thr_lds,
                                              uint32_t& bar_target, uint32_t spin_limit,
                                              uint32_t* err, uint32_t& nacc, uint64_t t_start) {
    const uint32_t C = I.cells;
    if constexpr (MODE == kThrLds) {
        // the ladder is fixed: one table [C][dtab + 1] for the whole item
        const uint32_t row = I.dtab + 1;
        for (uint32_t i = threadIdx.x; i < C * row; i += WAVES * 64) {
            const uint32_t t = i / row, d = i % row;
            thr_lds[i] = d == 0 ? 0xFFFFFFFFu : metropolis_threshold(I.ladder[t], (int64_t)d);
        }
        __syncthreads();
    }
    const uint32_t split = I.split > 1 ? I.split : 1u;
    const uint32_t owners = gang / split, owner = member / split, sub = member % split;
    const uint32_t per_owner = (C + owners - 1) / owners;
    for (uint32_t k = k0; k < k1; ++k) {
        for (uint32_t j = 0; j < per_owner; ++j) {
            const uint32_t cell = owner + j * owners;
            const bool active = cell < C;
            if (!active && split == 1) break;
            const uint32_t tr = active ? I.temp_of[cell * 64 + lane] : 0u;
            const double beta = I.ladder[tr];
            const double nbd = -beta * 1.4426950408889634;
            const uint32_t dlim = threshold_limit(beta, I.dtab);
            auto accept = [&](uint32_t uu, int32_t de) -> bool {
                if ((uint32_t)de > dlim) return false;
                if constexpr (MODE == kThrLds)
                    return uu < thr_lds[tr * (I.dtab + 1) + (uint32_t)de];
                else
                    return band_accept(uu, de, nbd, beta);
            };
            uint64_t* S = I.spins + (uint64_t)cell * I.n;
            int64_t eacc = 0;
            for (uint32_t c = 0; c < I.ncol; ++c) {
                const uint32_t cb = __builtin_amdgcn_readfirstlane(I.cs[c]);
                const uint32_t ce = __builtin_amdgcn_readfirstlane(I.cs[c + 1]);
                if (active && ce > cb) {
                    for (uint32_t g = (cb >> 2) + sub * WAVES + wave; g <= ((ce - 1) >> 2);
                         g += split * WAVES) {
                        const uint32_t p0 = max(g * 4u, cb), p1 = min(g * 4u + 4u, ce);
                        generic_group(I.rp, I.ci, I.wt, I.lin, S, S, p0, p1, cell, k, I.key0,
                                      I.key1, lane, accept, eacc, nacc);
                    }
                }
                if (split > 1)
                    gang_barrier(I.bar, bar_target, gang, spin_limit, err);
                else
                    __syncthreads();
            }
            if (eacc != 0)
                atomicAdd(reinterpret_cast<unsigned long long*>(I.energy + cell * 64 + lane),
                          (unsigned long long)eacc);
        }
        // the energy atomics complete at device scope before anyone reads them
        if (gang > 1)
            gang_barrier(I.bar, bar_target, gang, spin_limit, err);
        else
            block_sync_device();
        const bool timed = I.round_ticks != 0;
        if (C > 1 || timed) {
            if (member == 0 && wave == 0) {
                if (C > 1) pt_exchange(I, k, lane);
                if (timed && lane == 0)
                    __hip_atomic_store(I.stop,
                                       wall_clock64() - t_start >= I.round_ticks ? 1u : 0u,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (gang > 1)
                gang_barrier(I.bar, bar_target, gang, spin_limit, err);
            else
                block_sync_device();
        }
        if (I.track_every && ((k + 1) % I.track_every == 0 || k + 1 == I.sweeps))
            pt_track_point<WAVES>(I, k + 1, member, gang, lane, wave, bar_target, spin_limit, err);
        // every member reads the flag member 0 wrote before the last barrier
        if (timed && __hip_atomic_load(I.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
            return k + 1;
    }
    return k1;
}

// queue[t] = instance << kGangBits | member.  A gang's tickets are consecutive, so at
// most one gang is incomplete at any time and it is completed by the next
// workgroups to finish their work: with every workgroup of the launch
// resident and gang <= grid size, every wait ends.
// kPacked (WAVES >= 4): registers capped for 8 waves per SIMD (4 workgroups
// of 8 waves per CU instead of 3; SGPRs spill to VGPR lanes outside the
// entry loops).  The batch runs this form (+7% on the 4096-instance QUBO
// batch); the serving loop the other, whose PT gangs lost 7% with it.  The
// packed form runs anneals only (the batch has no PT solves), so it compiles
// no PT path and carries none of its uniform values.
template <int WAVES, bool kPacked>
__global__ __launch_bounds__(WAVES * 64)
__attribute__((amdgpu_waves_per_eu(kPacked && WAVES >= 4 ? 8 : 1))) void persistent_anneal_kernel(
    const InstDesc* __restrict__ insts, const uint32_t* __restrict__ queue, uint32_t count,
    unsigned int* next, uint32_t round, uint32_t spin_limit, uint32_t* err) {
    __shared__ uint32_t thr_lds[kLdsThrEntries];
    __shared__ int64_t part[WAVES][64];
    __shared__ uint32_t s_item;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(next, 1u);
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= count) break;
        const uint32_t code = queue[item];
        const InstDesc I = insts[code >> kGangBits];
        const uint32_t member = code & (kMaxGang - 1u);
        const uint32_t gang = I.gang;
        uint32_t bar_target = 0;
        const uint32_t mode = I.thr_mode;
        const uint64_t t_start = wall_clock64();
        const uint32_t k0 = __builtin_amdgcn_readfirstlane(*I.progress);
        const uint32_t k1 = min(I.sweeps, k0 + (I.round ? I.round : round));
        int64_t eacc = 0;
        uint32_t nacc = 0;
        uint32_t k_end = k1;
        // One code path per kind and threshold mode, so none carries another's
        // registers.  A PT solve keeps its energies up to date itself (eacc stays 0).
        if (!kPacked && I.pt) {
            if (mode == kThrLds)
                k_end = pt_sweeps<WAVES, kThrLds>(I, k0, k1, member, gang, lane, wave, thr_lds,
                                                  bar_target, spin_limit, err, nacc, t_start);
            else
                k_end = pt_sweeps<WAVES, kThrBand>(I, k0, k1, member, gang, lane, wave, thr_lds,
                                                   bar_target, spin_limit, err, nacc, t_start);
        } else if (mode == kThrLds) {
            anneal_sweeps<WAVES, kThrLds>(I, k0, k1, member, gang, lane, wave, thr_lds, part,
                                          bar_target, spin_limit, err, eacc, nacc);
        } else {
            anneal_sweeps<WAVES, kThrBand>(I, k0, k1, member, gang, lane, wave, thr_lds, part,
                                           bar_target, spin_limit, err, eacc, nacc);
        }
        // nacc is already the wave's total (a popcount of the accept ballot).
        if (lane == 0 && nacc) atomicAdd(I.accepted, (unsigned long long)nacc);
        if (!I.pt) {
            part[wave][lane] = eacc;
            __syncthreads();
            if (wave == 0) {
                int64_t tot = 0;
#pragma unroll
                for (int w = 0; w < WAVES; ++w) tot += part[w][lane];
                if (gang > 1)
                    atomicAdd(reinterpret_cast<unsigned long long*>(I.energy + lane),
                              (unsigned long long)tot);
                else
                    I.energy[lane] += tot;
            }
        }
        if (gang > 1) gang_barrier(I.bar, bar_target, gang, spin_limit, err);
        if (member == 0 && threadIdx.x == 0) {
            *I.progress = k_end;
            if (I.ticks) *I.ticks = wall_clock64() - t_start;
            if (I.swept) *I.swept = k_end - k0;
        }
        __syncthreads();
    }
}

// Default initial words (Philox (site, 0, 0, TAG_INIT)) and exact energies,
// one workgroup per listed instance.
static __global__ __launch_bounds__(256) void instance_init_kernel(const InstDesc* __restrict__ insts,
                                                            const uint32_t* __restrict__ list,
                                                            const uint32_t* const* orders) {
    const InstDesc I = insts[list[blockIdx.x]];
    const uint32_t* order = orders[list[blockIdx.x]];
    const uint32_t cells = I.pt ? I.cells : 1u;
    for (uint32_t s = 0; s < cells; ++s)
        for (uint32_t p = threadIdx.x; p < I.n; p += blockDim.x) {
            const U32x4 o = philox4x32_10(order[p], s, 0u, kTagInit << 24, I.key0, I.key1);
            I.spins[(uint64_t)s * I.n + p] = (uint64_t)o.v[0] | ((uint64_t)o.v[1] << 32);
        }
    if (I.pt)
        for (uint32_t i = threadIdx.x; i < cells * 64; i += blockDim.x) {
            I.temp_of[i] = i / 64;
            I.slot_of[i] = i / 64;
        }
    if (threadIdx.x == 0) {
        *I.progress = 0;
        *I.accepted = 0;
        if (I.best) I.best->valid = 0;
    }
}

// Exact energies of the listed instances from their words: into their tracked
// energies, or (out != NULL) into out[list index][cells][64], leaving them
// untouched.
static __global__ __launch_bounds__(256) void instance_energy_kernel(const InstDesc* __restrict__ insts,
                                                              const uint32_t* __restrict__ list,
                                                              int64_t* out) {
    __shared__ int64_t part[4][64];
    const InstDesc I = insts[list[blockIdx.x]];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t cells = I.pt ? I.cells : 1u;
    for (uint32_t s = 0; s < cells; ++s) {
        const uint64_t* S = I.spins + (uint64_t)s * I.n;
        int64_t a = 0;
        for (uint32_t p = wv; p < I.n; p += 4) {
            int64_t f = 2 * (int64_t)I.lin[p];
            for (uint32_t e = I.rp[p]; e < I.rp[p + 1]; ++e) f += lane_bit(S[I.ci[e]]) ? I.wt[e] : 0;
            a += lane_bit(S[p]) ? f : 0;
        }
        part[wv][lane] = a;
        __syncthreads();
        if (threadIdx.x < 64) {
            const int64_t e =
                I.offset + (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]) / 2;
            if (out)
                out[((uint64_t)blockIdx.x * cells + s) * 64 + lane] = e;
            else
                I.energy[s * 64 + lane] = e;
        }
        __syncthreads();
    }
}

}  // namespace ige
