/*
 * cpu_baseline.h -- port of the reference's Rayon+SIMD production sweep
 * (src/solver/engine.rs::step, lines 573-819) used as the reported CPU
 * baseline.  TEST/BENCH INFRASTRUCTURE: only bench.py's cpu_baseline leg and
 * tests load it.
 */
#ifndef ISING_CPU_BASELINE_H
#define ISING_CPU_BASELINE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/*
 * spins:    [cells][n][64] int8 (0/1), the QuantumField layout at one slice
 * energies: [cells][64] f64, tracked
 * temps:    [cells] temperatures (engine.rs: beta = 1/temps[t])
 * Runs `steps` calls of step(): Phase 1 sweep of every cell in ascending site
 * order (cells in parallel over `threads` OpenMP threads), Phase 2 DEO exchange.
 * *deo_odd_phase carries the DEO phase across calls.
 */
int cb_run(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const double *w,
           const double *linear, uint32_t cells, const double *temps, uint32_t steps,
           uint64_t seed, int threads, int8_t *spins, double *energies, int *deo_odd_phase);

/* cb_run, adding per pair t (t, t+1) the exchange attempts and accepts to
 * pair_stats[2t], pair_stats[2t+1]. */
int cb_run_stats(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const double *w,
                 const double *linear, uint32_t cells, const double *temps, uint32_t steps,
                 uint64_t seed, int threads, int8_t *spins, double *energies, int *deo_odd_phase,
                 uint64_t *pair_stats);

/* calculate_replica_energies_local (engine.rs:333-418) for every cell. */
void cb_energies(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const double *w,
                 const double *linear, uint32_t cells, const int8_t *spins, double *energies);

/* fast_exp (engine.rs:181-205), exported for its accuracy test. */
double cb_fast_exp(double x);

#ifdef __cplusplus
}
#endif
#endif
