/*
 * cpu_baseline.c -- port of src/solver/engine.rs::step (lines 573-819), the
 * reference's production multispin sweep: i8 lane-per-replica spins, f64 dE
 * over the CSR (calculate_delta_e_local, engine.rs:218-306), per-cell
 * Xoshiro256++x8 streams (engine.rs:81-172), fast_exp acceptance
 * (engine.rs:181-205), DEO even/odd exchange every step (engine.rs:713-818).
 * Cells run in parallel (OpenMP here, rayon par_chunks_mut there).
 *
 * Declared departures: the per-step master draw and the exchange draws use
 * SplitMix64 instead of ChaCha8 (rand_chacha is not available here); cost
 * per step is one draw plus at most 64 per pair, negligible next to the sweep.
 * The Trotter path (num_slices > 1) is not ported: the production default is
 * one slice, where engine.rs takes its fast path.
 *
 * Reported baseline only; never used to check the GPU.
 */
#include "cpu_baseline.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define NR 64

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

typedef struct {
    uint64_t s[4][8];
} xo8_t;

/* engine.rs:90-107: SplitMix64 expansion, lane-major over the four words. */
static void xo8_init(xo8_t *g, uint64_t seed) {
    uint64_t state = seed;
    for (int lane = 0; lane < 8; ++lane)
        for (int w = 0; w < 4; ++w) {
            state += 0x9e3779b97f4a7c15ull;
            uint64_t z = state;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            g->s[w][lane] = z ^ (z >> 31);
        }
}

/* engine.rs:111-130: one step of all 8 lanes, mantissa-trick uniforms. */
static inline void xo8_next(xo8_t *g, double out[8]) {
    for (int l = 0; l < 8; ++l) {
        uint64_t r = rotl(g->s[0][l] + g->s[3][l], 23) + g->s[0][l];
        uint64_t bits = 0x3FF0000000000000ull | (r >> 12);
        double d;
        memcpy(&d, &bits, 8);
        out[l] = d - 1.0;
    }
    for (int l = 0; l < 8; ++l) {
        uint64_t t = g->s[1][l] << 17;
        g->s[2][l] ^= g->s[0][l];
        g->s[3][l] ^= g->s[1][l];
        g->s[1][l] ^= g->s[2][l];
        g->s[0][l] ^= g->s[3][l];
        g->s[2][l] ^= t;
        g->s[3][l] = rotl(g->s[3][l], 45);
    }
}

/* engine.rs:181-205 */
double cb_fast_exp(double x) {
    x = x < -20.0 ? -20.0 : (x > 0.0 ? 0.0 : x);
    double r = x * 0.0625;
    double r2 = r * r;
    double r3 = r2 * r;
    double p = 1.0 + r + 0.5 * r2 + (1.0 / 6.0) * r3 + (1.0 / 24.0) * r2 * r2 +
               (1.0 / 120.0) * r2 * r3 + (1.0 / 720.0) * r3 * r3;
    double p2 = p * p;
    double p4 = p2 * p2;
    double p8 = p4 * p4;
    return p8 * p8;
}

static uint64_t splitmix_next(uint64_t *state) {
    *state += 0x9e3779b97f4a7c15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static double splitmix_f64(uint64_t *state) {
    return (double)(splitmix_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* One cell's Phase 1 (engine.rs:623-670, fast path). */
static void sweep_cell(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                       const double *w, const double *linear, double beta, uint64_t seed,
                       int8_t *spins, double *es) {
    xo8_t g;
    xo8_init(&g, seed);
    double u[NR], delta[NR];
    int8_t sign_v[NR];
    for (uint32_t v = 0; v < n; ++v) {
        for (int b = 0; b < 8; ++b) xo8_next(&g, u + 8 * b);
        int8_t *xv = spins + (uint64_t)v * NR;
        /* calculate_delta_e_local (engine.rs:218-306), 2-body part */
        for (int r = 0; r < NR; ++r) sign_v[r] = (int8_t)(1 - 2 * xv[r]);
        double wl = linear[v];
        for (int r = 0; r < NR; ++r) delta[r] = wl * (double)sign_v[r];
        for (uint32_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
            const int8_t *xj = spins + (uint64_t)col_idx[e] * NR;
            double we = w[e];
            for (int r = 0; r < NR; ++r) {
                int8_t t = (int8_t)((xj[r] ^ xv[r]) - xv[r]);
                delta[r] += we * (double)t;
            }
        }
        for (int r = 0; r < NR; ++r) {
            double d = delta[r];
            double ev = cb_fast_exp(-d * beta);
            int accept = (d <= 0.0) | (u[r] < ev);
            int8_t flip = (int8_t)accept;
            xv[r] ^= flip;
            es[r] += d * (double)flip;
        }
    }
}

static int run_impl(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                    const double *w, const double *linear, uint32_t cells, const double *temps,
                    uint32_t steps, uint64_t seed, int threads, int8_t *spins, double *energies,
                    int *deo_odd_phase, uint64_t *pair_stats) {
    if (cells == 0 || threads < 1) return -1;
    uint64_t master = seed;
    const uint64_t cell_len = (uint64_t)n * NR;
    for (uint32_t step = 0; step < steps; ++step) {
        uint64_t base_seed = splitmix_next(&master);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
        for (uint32_t c = 0; c < cells; ++c)
            sweep_cell(n, row_ptr, col_idx, w, linear, 1.0 / temps[c], base_seed + c,
                       spins + c * cell_len, energies + (uint64_t)c * NR);
        /* Phase 2: DEO (engine.rs:713-789) */
        uint32_t start = *deo_odd_phase ? 1u : 0u;
        for (uint32_t t = start; t + 1 < cells; t += 2) {
            double *e1 = energies + (uint64_t)t * NR, *e2 = energies + (uint64_t)(t + 1) * NR;
            double dbeta = 1.0 / temps[t] - 1.0 / temps[t + 1];
            int8_t mask[NR];
            int any = 0;
            for (int r = 0; r < NR; ++r) {
                double ex = dbeta * (e1[r] - e2[r]);
                int s = ex >= 0.0 || splitmix_f64(&master) < exp(ex);
                mask[r] = (int8_t)(-s);
                any |= s;
                if (pair_stats) {
                    pair_stats[2 * t] += 1;
                    pair_stats[2 * t + 1] += (uint64_t)s;
                }
            }
            if (any) {
                int8_t *c1 = spins + t * cell_len, *c2 = spins + (t + 1) * cell_len;
                for (uint64_t v = 0; v < n; ++v)
                    for (int r = 0; r < NR; ++r) {
                        int8_t d = (int8_t)((c1[v * NR + r] ^ c2[v * NR + r]) & mask[r]);
                        c1[v * NR + r] ^= d;
                        c2[v * NR + r] ^= d;
                    }
                for (int r = 0; r < NR; ++r)
                    if (mask[r]) {
                        double tmp = e1[r];
                        e1[r] = e2[r];
                        e2[r] = tmp;
                    }
            }
        }
        *deo_odd_phase = !*deo_odd_phase;
    }
    return 0;
}

int cb_run(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const double *w,
           const double *linear, uint32_t cells, const double *temps, uint32_t steps,
           uint64_t seed, int threads, int8_t *spins, double *energies, int *deo_odd_phase) {
    return run_impl(n, row_ptr, col_idx, w, linear, cells, temps, steps, seed, threads, spins,
                    energies, deo_odd_phase, NULL);
}

int cb_run_stats(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const double *w,
                 const double *linear, uint32_t cells, const double *temps, uint32_t steps,
                 uint64_t seed, int threads, int8_t *spins, double *energies, int *deo_odd_phase,
                 uint64_t *pair_stats) {
    return run_impl(n, row_ptr, col_idx, w, linear, cells, temps, steps, seed, threads, spins,
                    energies, deo_odd_phase, pair_stats);
}

void cb_energies(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const double *w,
                 const double *linear, uint32_t cells, const int8_t *spins, double *energies) {
    for (uint32_t c = 0; c < cells; ++c) {
        const int8_t *s = spins + (uint64_t)c * n * NR;
        double *es = energies + (uint64_t)c * NR;
        for (int r = 0; r < NR; ++r) es[r] = 0.0;
        for (uint32_t v = 0; v < n; ++v) {
            for (int r = 0; r < NR; ++r) es[r] += linear[v] * (double)s[(uint64_t)v * NR + r];
            for (uint32_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
                uint32_t j = col_idx[e];
                if (v < j)
                    for (int r = 0; r < NR; ++r)
                        es[r] += w[e] * (double)(s[(uint64_t)v * NR + r] & s[(uint64_t)j * NR + r]);
            }
        }
    }
}
