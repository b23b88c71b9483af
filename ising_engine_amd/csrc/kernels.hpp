// kernels.hpp -- device data views and host-side launchers for the gfx950
// kernels.  All launchers are asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ige {

// Model in colour-ordered position space: position p holds site order[p];
// colour class c occupies positions [class_start[c], class_start[c+1]).
struct DevModel {
    uint32_t n = 0;
    const uint32_t* rp = nullptr;   // n+1 row pointers (position space)
    const uint32_t* ci = nullptr;   // nnz neighbour positions
    const int32_t* wt = nullptr;    // nnz couplings q_ij
    const int32_t* lin = nullptr;   // n linear terms
    const uint32_t* order = nullptr;  // n: position -> site
    const uint32_t* sgn = nullptr;  // n: bit e set iff entry e of the row has q < 0
    int32_t w0 = 0;                 // common |q| for the multispin kernel
    uint32_t dmax = 0;              // max degree
    uint32_t regular_degree = 0;    // d when every site has degree d and row p starts at p*d
    int64_t offset = 0;
};

// Dense couplings for the matrix-core path: q = scale * a, a int8.
struct DevDense {
    uint32_t n = 0, n_pad = 0;      // n_pad: n rounded up to a multiple of 512
    const int8_t* a = nullptr;      // n_pad^2 position space, zero padded, in 32x32
                                    // operand tiles (see block_fields)
    const int32_t* wdiag = nullptr; // [n_pad / 64][64][64] scale * a inside each diagonal block
    const int8_t* hdiag = nullptr;  // [n_pad / 32][32][32] a inside each diagonal 32 x 32
                                    // tile, row-major (32 bytes per row)
    const int32_t* lin = nullptr;   // n_pad linear terms, zero padded
    const uint32_t* order = nullptr;  // n: position -> site
    int32_t scale = 1;
    uint32_t dtab = 0;              // bound on |dE| (max local field)
};

// Step clock of a captured sweep loop (ige_run's graphs): kernels add
// clock[kClockSweep] to their sweep number, clock[kClockRound] to their
// exchange round and clock[kClockRow] * cells to their beta pointer, and a
// final node of each replay advances the three.  NULL for eager launches,
// whose arguments are absolute.
enum : uint32_t { kClockSweep = 0, kClockRound = 1, kClockRow = 2, kClockWords = 4 };

struct DevState {
    uint32_t cells = 0;
    // Temperature rows (entries of beta and of the threshold tables): cells, or
    // cells * 64 in the replica view, where temp_of holds replica indices
    // rho = slot * 64 + lane (identity at the switch) and slot_of[rho] the
    // position slot * 64 + lane of replica rho.
    uint32_t temps = 0;
    uint32_t replica_view = 0;
    uint64_t* spins = nullptr;      // [cells][n] MSC words, position order
    int64_t* energy = nullptr;      // [cells][64]
    uint32_t* temp_of = nullptr;    // [cells][64] temperature (replica) index of (slot, lane)
    uint32_t* slot_of = nullptr;    // [cells][64] slot of (temperature, lane), or position of
                                    // each replica in the replica view
    unsigned long long* accepted = nullptr;    // 1 counter
    unsigned long long* pair_stats = nullptr;  // [cells-1][2] attempts, accepts
    uint32_t pop_temps = 0;  // levels per population (0: one population of all cells)
    const uint32_t* clock = nullptr;  // kClockWords step counters, or NULL
};

// clock[kClockSweep, kClockRound, kClockRow] = (sweep, round, row), and
// clock[...] += (dsweep, dround, drow): one thread each, no memset (these
// run inside captured graphs).
hipError_t launch_clock_set(uint32_t* clock, uint32_t sweep, uint32_t round, uint32_t row,
                            hipStream_t stream);
hipError_t launch_clock_advance(uint32_t* clock, uint32_t dsweep, uint32_t dround, uint32_t drow,
                                hipStream_t stream);

// thr[t*(dtab+1) + d] = metropolis_threshold(beta[t], d) for d in [1, dtab].
// clock: NULL, or a step clock whose row advances beta by rows of `cells`.
hipError_t launch_thr_table(const double* beta, uint32_t* thr, uint32_t cells, uint32_t dtab,
                            hipStream_t stream, const uint32_t* clock = nullptr);

// Generic sweep of one colour class (lane = replica).
hipError_t launch_sweep_generic(const DevModel& m, const DevState& s, uint32_t c_begin,
                                uint32_t c_end, uint32_t sweep, uint64_t seed,
                                const uint32_t* thr, uint32_t dtab, hipStream_t stream);

// Heat-bath thresholds hb[t*(2 dtab+1) + h + dtab] = heatbath_threshold(beta[t], h)
// for h in [-dtab, dtab], and the heat-bath sweep of one colour class
// (lane = replica, the generic kernel's uniforms and mapping).
hipError_t launch_heatbath_table(const double* beta, uint64_t* hb, uint32_t cells, uint32_t dtab,
                                 hipStream_t stream, const uint32_t* clock = nullptr);
hipError_t launch_sweep_heatbath(const DevModel& m, const DevState& s, uint32_t c_begin,
                                 uint32_t c_end, uint32_t sweep, uint64_t seed, const uint64_t* hb,
                                 uint32_t dtab, hipStream_t stream);

// Multispin sweep of one colour class (lane = site, bit-sliced +-J).
// lev_thr: [cells][dmax+1] thresholds per temperature and level
// (launch_level_thresholds); each workgroup turns its slot's into bit planes
// through temp_of.
hipError_t launch_sweep_multispin(const DevModel& m, const DevState& s, uint32_t c_begin,
                                  uint32_t c_end, uint32_t sweep, uint64_t seed,
                                  const uint32_t* lev_thr, hipStream_t stream);
// Asks (once) how many workgroups of the model's multispin kernel fit a CU,
// which sizes its grids; called at state creation, outside any capture.
void prepare_sweep_multispin(const DevModel& m);

// Uniforms of the multispin kernel: kEagerPlanes leading bits as bit planes
// (two per Philox call), the rest from a tail word drawn only on a tie.
constexpr uint32_t kEagerPlanes = 8;
// Per (slot, level m), in the sweep's LDS: the threshold's kEagerPlanes
// leading bit planes, the nonzero mask and the 64 thresholds (two per u64).
constexpr uint32_t kQStride = kEagerPlanes + 1 + 32;

// Whole runs on small regular lattices (degree 4 or 6, n <= kResidentMaxSites,
// no exchange): one workgroup per slot keeps the slot's words in LDS for all
// `sweeps` sweeps (numbered sweep0, sweep0 + 1, ...), the same updates as the
// colour-class launches.  class_start: ncol + 1 class bounds on the device.
// Sweep k reads its thresholds from lev_rows + k * row_stride, laid out as
// lev_thr below (row_stride 0: one row for every sweep).
constexpr uint32_t kResidentMaxSites = 4096;
hipError_t launch_resident_sweeps(const DevModel& m, const DevState& s,
                                  const uint32_t* class_start, uint32_t ncol,
                                  const uint32_t* lev_rows, uint32_t row_stride, uint32_t sweep0,
                                  uint32_t sweeps, uint64_t seed, hipStream_t stream);
// Colour classes [c0, c1) of sweep `sweep` in one workgroup per slot (any
// multispin model), one class after the other: the same updates as
// launch_sweep_multispin class by class.  For trailing classes too small to
// be worth a launch each: every fused class is one pass of the workgroup's
// kFuseMaxClassSites lanes (a class that needs several passes is a chain of
// dependent memory round trips per pass and runs no faster fused).
// lev_thr as for launch_sweep_multispin.
constexpr uint32_t kFuseMaxClassSites = 512;

// Sweeps [sweep0, sweep0 + sweeps) of a multispin state of irregular degree
// (dmax < 32) without exchange in one launch (a few when a barrier counter
// would overflow): every class of every sweep, a gang of workgroups per slot
// meeting at the slot's barrier bars[slot] (bars: 2 * cells words, zeroed
// here; err raised when a wait expires) after each class, on one XCD when
// dispatch puts them there; the same updates as launch_sweep_multispin
// class by class.  Sweep k's level thresholds are lev + k * lev_stride
// (lev_thr layout).  max_class: positions in the largest class, at most
// kChainMaxSites (a lane pair per position: two passes of kChainMaxSites /
// 256 workgroups).
// Needs cells <= chain_ms_capacity_blocks(dmax).
constexpr uint32_t kChainMaxSites = 32768;
hipError_t launch_sweep_chain_ms(const DevModel& m, const DevState& s, const uint32_t* cstart,
                                 uint32_t ncol, uint32_t max_class, const uint32_t* lev,
                                 uint32_t lev_stride, uint32_t sweep0, uint32_t sweeps,
                                 uint64_t seed, uint32_t* bars, uint32_t* err, uint32_t spin_limit,
                                 hipStream_t stream);
uint32_t chain_ms_capacity_blocks(uint32_t dmax);
// launch_sweep_chain_ms without barriers, for one slot: a site waits for
// its neighbours' per-site sweep counts.  units: 2 n u64 of scratch (the
// words with their counts during the launch); err raised when a wait
// expires.  The same bits as launch_sweep_chain_ms.
hipError_t launch_sweep_flow_ms(const DevModel& m, const DevState& s, const uint32_t* cstart,
                                uint32_t ncol, uint32_t max_class, const uint32_t* lev,
                                uint32_t lev_stride, uint32_t sweep0, uint32_t sweeps,
                                uint64_t seed, uint64_t* units, uint32_t* err,
                                uint32_t spin_limit, hipStream_t stream);
uint32_t flow_ms_capacity_blocks(uint32_t dmax);
hipError_t launch_classes_sweep(const DevModel& m, const DevState& s,
                                const uint32_t* class_start, uint32_t c0, uint32_t c1,
                                uint32_t sweep, uint64_t seed, const uint32_t* lev_thr,
                                hipStream_t stream);
// lev_thr rows for `rows` ladders beta_rows[k * cells + t] into out (rows x
// cells x (dmax + 1)).
hipError_t launch_level_rows(const DevModel& m, uint32_t cells, const double* beta_rows,
                             uint32_t rows, uint32_t* out, hipStream_t stream);

// lev_thr[t*(dmax+1) + m] = metropolis_threshold(beta[t], m*w0/2) for m in
// [1, dmax], 0 for m = 0.  It depends on the ladder alone, so an exchange
// (which moves temperatures between slots) needs no refresh.
hipError_t launch_level_thresholds(const DevModel& m, const DevState& s, const double* beta,
                                   uint32_t* lev_thr, hipStream_t stream);

// DEO exchange number `round`: pairs (t, t+1) with t = round mod 2, inside
// each population of s.pop_temps temperature indices (engine.rs:713-716).
hipError_t launch_exchange(const DevState& s, const double* beta, uint32_t round, uint64_t seed,
                           hipStream_t stream);
// One ReplicaExchange round in the replica view (replica_exchange.rs:61-109):
// pairs of replica indices (i, i + 1) for i = phase, phase + 2, ... < cells*64 - 1,
// uniforms keyed on `round`; beta [cells * 64].  Pair i's attempts and accepts
// are added to pair_stats[2i], [2i + 1].
hipError_t launch_exchange_replicas(const DevState& s, const double* beta, uint32_t phase,
                                    uint32_t round, uint64_t seed, hipStream_t stream);

// Dense sweep: all positions in order, one workgroup per slot, fields of each
// block of 64 positions recomputed on int8 matrix cores.  bounds != NULL
// (dense_bounds_words(d, cells) int32 of scratch): proposals are decided by
// per-(position, replica) accept bounds computed before the sweep on the whole
// GPU (dense_bounds_kernel, then sweep_dense_bound_kernel; requires
// dense_bounds_eligible and every beta the sweep reads >= kDenseBoundMinBeta);
// NULL: by the threshold band inside the sequential chain.
// gang: NULL, or buffers for field helpers (dense_gang_helpers): other
// workgroups on the slot's XCD that compute part of each block's fields.
struct DevDenseGang {
    uint32_t helpers = 0;       // helper workgroups per slot (<= kMaxDenseHelpers)
    uint32_t* sync = nullptr;   // kDenseGangSyncWords per slot, zeroed by the launch
    uint32_t* err = nullptr;    // raised when a wait exceeds spin_limit polls
    uint64_t* xpub = nullptr;   // dense_gang_xpub_words: finished blocks' spins per replica
    int32_t* part = nullptr;    // dense_gang_part_words: the helpers' partial fields
    uint32_t spin_limit = 0;
};
constexpr uint32_t kMaxDenseHelpers = 3;
constexpr uint32_t kDenseGangSyncWords = 8;
hipError_t launch_sweep_dense(const DevDense& d, const DevState& s, uint32_t sweep, uint64_t seed,
                              const double* beta, int32_t* bounds, const DevDenseGang* gang,
                              hipStream_t stream);
size_t dense_bounds_words(const DevDense& d, uint32_t cells);
// Field helpers per slot for the accept-bound sweep of `cells` slots (0: the
// slot's workgroup computes every field itself).
uint32_t dense_gang_helpers(const DevDense& d, uint32_t cells);
size_t dense_gang_xpub_words(const DevDense& d, uint32_t cells);
size_t dense_gang_part_words(uint32_t cells, uint32_t helpers);
// The accept-bound kernel keeps local fields, accept bounds and their sums in
// 24-bit multiplies and needs beta >= kDenseBoundMinBeta for its error bound.
constexpr double kDenseBoundMinBeta = 1e-6;
bool dense_bounds_eligible(const DevDense& d);
// Local fields of every (slot, site, lane) on the matrix cores, layout as
// launch_fields.
hipError_t launch_dense_fields(const DevDense& d, const DevState& s, int32_t* out,
                               hipStream_t stream);
size_t dense_lds_bytes_host(uint32_t n_pad);

hipError_t launch_init_spins(const DevModel& m, const DevState& s, uint64_t seed,
                             hipStream_t stream);
// Exact per-(slot, lane) energies from scratch into `out` ([cells][64]).
hipError_t launch_energy(const DevModel& m, const DevState& s, int64_t* out, hipStream_t stream);
// Local fields, out[(slot*n + site)*64 + lane] (site order).
hipError_t launch_fields(const DevModel& m, const DevState& s, int32_t* out, hipStream_t stream);
hipError_t launch_identity_perm(const DevState& s, hipStream_t stream);
// temp_of[q] = temp_of[q] * 64 + q % 64 and slot_of its inverse: the
// slot-ladder permutation expressed as replica indices (the switch to the
// replica view).
hipError_t launch_replica_index_perm(const DevState& s, hipStream_t stream);

// Steepest descent (every replica flips its most negative-dE site per step).
// ctl: kDescentCtlWords counters, cleared by launch_descent_reset with the
// keys ([cells][64]); after a step without a flip ctl[kDescentDone] is 1 and
// further steps do nothing.
enum : uint32_t {
    kDescentDone = 0,       // 1 once a step flipped nothing
    kDescentStepFlips = 1,  // flips of the step in progress
    kDescentSteps = 2,      // steps that flipped at least one replica
    kDescentFlips = 3,      // flips of all finished steps
    kDescentCtlWords = 4
};
hipError_t launch_descent_reset(unsigned long long* keys, uint32_t cells, unsigned long long* ctl,
                                hipStream_t stream);
// site_aux[p] = order[p] | (clamped[order[p]] != 0) << 31 for the n positions
// (order: position -> site; clamped: n bytes in site order, on the device).
hipError_t launch_descent_aux(const uint32_t* order, const uint8_t* clamped, uint32_t n,
                              uint32_t* site_aux, hipStream_t stream);
// One step.  dense_fields: NULL (dE from the CSR rows) or [cells][n][64]
// fields of the current spins in site order (launch_dense_fields).
// pos_of: n, site -> position.  clamped: NULL or n bytes in site order (read
// by the dense scan).  site_aux: n, position -> site with bit 31 set for a
// clamped site (read by the sparse scan; the model's order when nothing is
// clamped).  flips: [cells][64], one added per flip.
hipError_t launch_descent_step(const DevModel& m, const DevState& s, const int32_t* dense_fields,
                               const uint32_t* pos_of, const uint8_t* clamped,
                               const uint32_t* site_aux, unsigned long long* keys, uint32_t* flips,
                               unsigned long long* ctl, hipStream_t stream);

// Isoenergetic cluster moves for pairs [pair0, pair0 + pairs) of the
// cells * 32 pairs (temperature t = pair / 32, lanes r = pair % 32 and r + 32),
// ICM sweep number `round`.  scratch: pairs * icm_scratch_words(n) u32.
// flipped: NULL or a counter the cluster sizes are added to.
uint64_t icm_scratch_words(uint32_t n);
hipError_t launch_icm(const DevModel& m, const DevState& s, const uint32_t* pos_of,
                      uint32_t pair0, uint32_t pairs, uint32_t round, uint64_t seed,
                      uint32_t* scratch, unsigned long long* flipped, hipStream_t stream);

// Population-annealing resampling: spins and energies gathered from the
// snapshot (old_spins, old_energy) by member parents (pops*64 entries, with
// pops = cells / s.pop_temps).
hipError_t launch_resample(const DevState& s, uint32_t n, const uint64_t* old_spins,
                           const int64_t* old_energy, const uint32_t* parents,
                           hipStream_t stream);

// Best-seen replica: energy, where it was found, and (separately) its bits.
struct BestRecord {
    int64_t energy;
    uint32_t valid;
    uint32_t sweep;  // sweeps completed when it was found
    uint32_t temp;   // temperature index it held
    uint32_t lane;
};
// take: 2 words of scratch (decision and winning slot * 64 + lane).
hipError_t launch_track_best(const DevModel& m, const DevState& s, uint32_t sweep,
                             BestRecord* best, uint32_t* take, uint8_t* best_x,
                             hipStream_t stream);

}  // namespace ige
