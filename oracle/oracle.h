/*
 * oracle.h -- CPU restatement of the reference's multispin Metropolis sweep
 * and parallel-tempering exchange, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library. The product (ising_engine_amd/)
 * never links, loads or calls it.
 *
 * What each function restates (paths relative to the reference checkout):
 *   or_ir_from_pairs     src/engine_v2/ir.rs:30-69   (CSR, both triangles,
 *                        ascending columns, duplicates summed, zeros dropped)
 *   or_ir_energy         src/engine_v2/ir.rs:85-103  (canonical scorer)
 *   or_greedy_coloring   src/engine_v2/ir.rs:114-139
 *   or_color_order       src/engine_v2/operators/metropolis_sweep.rs:40-58
 *   or_fields            src/engine_v2/backends/sparse_bitslice.rs:88-112
 *   or_energies          src/engine_v2/backends/sparse_bitslice.rs:114-137
 *   or_run (sweep part)  src/engine_v2/operators/metropolis_sweep.rs:95-146
 *                        with ΔE = (1-2x)·h from sparse_bitslice.rs:247-255
 *   or_run (OR_HEATBATH) src/engine_v2/operators/gibbs_color_sweep.rs:119-174
 *   or_resample          src/solver/population_annealing.rs:229-286
 *   or_run (exchange)    src/solver/engine.rs:713-818 (DEO even/odd phase,
 *                        lane r of temperature t against lane r of t+1)
 *   or_run_temps         or_run with one temperature per replica
 *                        (metropolis_sweep.rs:115, T = temps[rep % len])
 *   or_exchange_replicas src/engine_v2/operators/replica_exchange.rs:61-109
 *                        (adjacent replica indices, start = round & 1)
 *   or_geometric_ladder  src/engine_v2/decision.rs:133-142
 *   or_suggested_range   src/solver/schedule.rs:21-43
 *   or_steepest_descent  src/engine_v2/operators/steepest_descent.rs:80-124
 *   or_polish_1opt       src/solver/local_search.rs:19-60
 *   or_update_best       src/solver/ultimate.rs:397-431
 *   or_icm_sweep         src/solver/icm.rs:91-185 (seed site and round keyed
 *                        on Philox; see the function)
 *
 * Parity mode departs from the reference in exactly two declared places,
 * because the reference's ChaCha8 streams are a third-party dependency that
 * cannot be built here (rand_chacha 0.3.1, rand 0.8.5):
 *   - randoms come from Philox4x32-10 (Salmon et al., SC'11) keyed by
 *     (seed) with counters (position, slot, sweep, tag|index), read either
 *     as one 32-bit word per replica (word view) or as eight bit planes over
 *     the 64 replicas followed by a per-replica tail word (plane view);
 *   - acceptance compares a 32-bit uniform u with
 *     thr = min(floor(exp(-beta*dE) * 2^32), 2^32-1), exp evaluated by
 *     or_det_exp (IEEE double, no contraction), so that the device and this
 *     file give identical bits.
 * Replica exchange keeps spins in place and swaps temperature indices per
 * lane (temp_of / slot_of), which is the same Markov chain as the
 * reference's spin copy observed in temperature space.
 */
#ifndef ISING_ORACLE_H
#define ISING_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG and deterministic exp ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_det_exp(double x);
uint32_t or_threshold(double beta, int64_t de);      /* de >= 1 */
uint32_t or_exchange_threshold(double delta);        /* delta < 0 */
uint64_t or_heatbath_threshold(double beta, int64_t h); /* in [0, 2^32] */

/* ---- model ---- */
/* Returns nnz (>= 0) or -1 on invalid pairs. Outputs sized n+1 / 2*npairs. */
int64_t or_ir_from_pairs(uint32_t n, uint64_t npairs, const uint32_t *pi,
                         const uint32_t *pj, const int64_t *q,
                         uint32_t *row_ptr, uint32_t *col_idx, int64_t *w);
int64_t or_ir_energy(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                     const int64_t *w, const int64_t *linear, int64_t offset,
                     const uint8_t *x);
void or_greedy_coloring(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                        uint32_t *colors);
/* order[pos] = site, colour-major then ascending site; returns #colours. */
uint32_t or_color_order(uint32_t n, const uint32_t *colors, uint32_t *order,
                        uint32_t *class_start /* ncolours+1, caller sizes n+1 */);

/* fields[(slot*n + site)*64 + lane] = linear + sum_j q_ij x_j */
void or_fields(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
               const int64_t *w, const int64_t *linear, uint32_t cells,
               const uint64_t *words, int64_t *fields);
/* energies[slot*64 + lane], exact int64 */
void or_energies(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                 const int64_t *w, const int64_t *linear, int64_t offset,
                 uint32_t cells, const uint64_t *words, int64_t *energies);

/* Default initial words: Philox(ctr=(site, slot, 0, TAG_INIT<<24)). */
void or_init_words(uint32_t n, uint32_t cells, uint64_t seed, uint64_t *words);

/* ---- the run ---- */
/* Metropolis with word- or plane-view uniforms, or the heat-bath rule with
 * word-view uniforms. */
enum { OR_VIEW_WORD = 0, OR_VIEW_PLANE = 1, OR_HEATBATH = 2 };

/*
 * Run `sweeps` colour-ordered Metropolis sweeps over `cells` slots of 64
 * lanes, optionally followed by one DEO exchange per sweep; sweep0 and
 * xround0 are the sweep and exchange counters the uniforms are keyed on.
 *   words      in/out [cells][n], site order; bit r = lane r
 *   temp_of    in/out [cells][64], temperature index held by (slot, lane)
 *   energies   in/out [cells][64], tracked int64 energies
 *   betas      [cells] (betas_per_sweep == 0) or [sweeps][cells]
 *   stats      out [2 + 2*(cells-1)]: accepted, proposed, then per pair
 *              (attempts, accepts); added to, not overwritten
 * Returns 0, or -1 on an invalid argument.
 */
int or_run(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
           const int64_t *w, const int64_t *linear, uint32_t cells, uint64_t seed,
           uint32_t view, uint32_t exchange, uint32_t sweep0, uint32_t xround0,
           uint32_t sweeps, const double *betas, uint32_t betas_per_sweep,
           uint32_t pop_temps /* levels per population, 0 = cells */,
           uint64_t *words, uint32_t *temp_of, int64_t *energies, uint64_t *stats);
/* or_run with `temps` betas per row: temps == cells (a temperature per slot
 * index, as or_run), or temps == cells * 64 (the replica view: temp_of holds
 * replica indices and replica rho runs at beta[rho]; exchange must be 0). */
int or_run_temps(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const int64_t *w,
                 const int64_t *linear, uint32_t cells, uint32_t temps, uint64_t seed,
                 uint32_t view, uint32_t exchange, uint32_t sweep0, uint32_t xround0,
                 uint32_t sweeps, const double *betas, uint32_t betas_per_sweep,
                 uint32_t pop_temps, uint64_t *words, uint32_t *temp_of, int64_t *energies,
                 uint64_t *stats);
/* One ReplicaExchange round in the replica view (pairs from `phase`, uniforms
 * keyed on xround); beta [cells*64]; stats [2 + 2*(cells*64-1)]. */
void or_exchange_replicas(uint32_t cells, uint64_t seed, uint32_t xround, uint32_t phase,
                          const double *beta, const int64_t *energies, uint32_t *temp_of,
                          uint64_t *stats);

/* Exchange number `round` (phase = round & 1; engine.rs:713-818) on tracked
 * energies; stats as or_run. */
void or_exchange(uint32_t cells, uint64_t seed, uint32_t round, const double *beta,
                 const int64_t *energies, uint32_t *temp_of, uint64_t *stats);
/* The same inside each population of pop_temps levels (engine.rs:714-716). */
void or_exchange_pops(uint32_t cells, uint32_t pop_temps, uint64_t seed, uint32_t round,
                      const double *beta, const int64_t *energies, uint32_t *temp_of,
                      uint64_t *stats);
/* apply_resample (population_annealing.rs:229-286); parents [cells/T*64]. */
void or_resample(uint32_t n, uint32_t cells, uint32_t pop_temps, const uint32_t *parents,
                 uint64_t *words, const uint32_t *temp_of, int64_t *energies);

/* ---- ladders (reference helpers) ---- */
void or_geometric_ladder(uint32_t k, double t_hi, double t_lo, double *temps);
void or_suggested_range(uint32_t n, const uint32_t *row_ptr, const int64_t *w,
                        const int64_t *linear, double *t_max, double *t_min);

/* ---- deterministic descent and the incumbent ---- */
/* Every replica of every slot flips its most negative-dE site (lowest site on
 * ties) once per step; stops after max_steps or a step without a flip.
 * clamped: NULL or [n] bytes (nonzero = never flipped).  words / energies in
 * and out; flips[cells*64] added to.  Returns the steps that flipped. */
uint32_t or_steepest_descent(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                             const int64_t *w, const int64_t *linear, const uint8_t *clamped,
                             uint32_t cells, uint32_t max_steps, uint64_t *words,
                             int64_t *energies, uint32_t *flips);
/* One state of n bytes to a 1-opt local minimum; returns the improvement. */
int64_t or_polish_1opt(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                       const int64_t *w, const int64_t *linear, const uint8_t *clamped,
                       uint8_t *x, uint64_t *flips);
/* One ICM sweep (icm.rs:163-185) over every temperature's 32 lane pairs. */
uint32_t or_icm_sweep(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                      const int64_t *w, const int64_t *linear, uint32_t cells, uint64_t seed,
                      uint32_t round, uint64_t *words, const uint32_t *temp_of,
                      int64_t *energies, uint32_t *flipped);

/* Incumbent update from tracked energies; words [cells][n] site order. */
int or_update_best(uint32_t n, uint32_t cells, const uint64_t *words, const uint32_t *temp_of,
                   const int64_t *energies, int64_t *best_e, int32_t *valid, uint8_t *best_x);
/* The same with ties broken by replica index in the replica view. */
int or_update_best_view(uint32_t n, uint32_t cells, uint32_t replica_view, const uint64_t *words,
                        const uint32_t *temp_of, const int64_t *energies, int64_t *best_e,
                        int32_t *valid, uint8_t *best_x);

#ifdef __cplusplus
}
#endif
#endif
