/*
 * oracle.c -- CPU restatement of the reference sweep/exchange path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product.  Each function cites
 * the reference lines it restates (paths relative to the reference checkout).
 *
 * Build with -ffp-contract=off: or_det_exp must round exactly like the
 * device code, which is compiled the same way.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Multipliers and Weyl key increments are the    */
/* published constants; pinned by the paper's known-answer vectors in        */
/* tests/test_oracle_kat.py.                                                  */
/* ------------------------------------------------------------------------- */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Stream tags in the top byte of counter word 3. */
#define TAG_WORD 1u
#define TAG_PLANE 2u
#define TAG_EXCH 3u
#define TAG_INIT 4u
#define TAG_TAIL 5u
#define TAG_ICM 6u
#define TAG_REPX 8u
/* Bit planes drawn for every replica before its tail word is consulted. */
#define EAGER_PLANES 8u

static void philox_seeded(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                          uint32_t out[4]) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------------- */
/* Deterministic exp for x <= 0: Cody-Waite reduction by ln2 with the fdlibm  */
/* split constants, degree-13 Taylor polynomial in Horner form, exact power   */
/* of two scaling.  Results below exp(-40) are returned as 0 (they map to a   */
/* zero threshold anyway).  Pinned against libm exp in the tests.             */
/* ------------------------------------------------------------------------- */
double or_det_exp(double x) {
    if (!(x > -40.0)) return 0.0;
    if (x >= 0.0) return 1.0;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = rint(x * inv_ln2);
    double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 6227020800.0;           /* 1/13! */
    p = p * r + 1.0 / 479001600.0;           /* 1/12! */
    p = p * r + 1.0 / 39916800.0;            /* 1/11! */
    p = p * r + 1.0 / 3628800.0;             /* 1/10! */
    p = p * r + 1.0 / 362880.0;              /* 1/9!  */
    p = p * r + 1.0 / 40320.0;               /* 1/8!  */
    p = p * r + 1.0 / 5040.0;                /* 1/7!  */
    p = p * r + 1.0 / 720.0;                 /* 1/6!  */
    p = p * r + 1.0 / 120.0;                 /* 1/5!  */
    p = p * r + 1.0 / 24.0;                  /* 1/4!  */
    p = p * r + 1.0 / 6.0;                   /* 1/3!  */
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)kd);
}

static uint32_t prob_to_u32(double e) {
    double v = e * 4294967296.0;
    if (v >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

uint32_t or_threshold(double beta, int64_t de) {
    return prob_to_u32(or_det_exp(-(beta * (double)de)));
}

uint32_t or_exchange_threshold(double delta) { return prob_to_u32(or_det_exp(delta)); }

/* gibbs_color_sweep.rs:131-158: P(x_i = 1) = 1 / (1 + exp(beta h)) as a threshold on a
 * 32-bit uniform, floor(P 2^32) in [0, 2^32]; exp is evaluated only at -beta |h| <= 0:
 * P = e / (1 + e) for h >= 0 and 1 / (1 + e) for h < 0, e = exp(-beta |h|).  beta -> inf
 * (T = 0) gives 0, 1/2 or 1 as the reference's limit does. */
uint64_t or_heatbath_threshold(double beta, int64_t h) {
    const double a = beta * (double)(h < 0 ? -h : h);
    const double e = or_det_exp(-a);
    const double p = h >= 0 ? e / (1.0 + e) : 1.0 / (1.0 + e);
    return (uint64_t)(p * 4294967296.0);
}

/* ------------------------------------------------------------------------- */
/* ProblemIR (src/engine_v2/ir.rs)                                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint32_t col;
    int64_t q;
} entry_t;

static int cmp_entry(const void *a, const void *b) {
    uint32_t x = ((const entry_t *)a)->col, y = ((const entry_t *)b)->col;
    return (x > y) - (x < y);
}

/* ir.rs:30-69: both triangles, rows by ascending column, duplicates summed,
 * zero sums dropped.  The reference asserts i < j < n; here that is -1. */
int64_t or_ir_from_pairs(uint32_t n, uint64_t npairs, const uint32_t *pi, const uint32_t *pj,
                         const int64_t *q, uint32_t *row_ptr, uint32_t *col_idx, int64_t *w) {
    uint64_t *cnt = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    for (uint64_t k = 0; k < npairs; ++k) {
        if (!(pi[k] < pj[k] && pj[k] < n)) {
            free(cnt);
            return -1;
        }
        cnt[pi[k] + 1]++;
        cnt[pj[k] + 1]++;
    }
    for (uint32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    entry_t *adj = (entry_t *)malloc(sizeof(entry_t) * (size_t)(2 * npairs + 1));
    uint64_t *fill = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)n + 1));
    memcpy(fill, cnt, sizeof(uint64_t) * ((size_t)n + 1));
    for (uint64_t k = 0; k < npairs; ++k) {
        adj[fill[pi[k]]++] = (entry_t){pj[k], q[k]};
        adj[fill[pj[k]]++] = (entry_t){pi[k], q[k]};
    }
    int64_t nnz = 0;
    row_ptr[0] = 0;
    for (uint32_t i = 0; i < n; ++i) {
        entry_t *row = adj + cnt[i];
        uint64_t len = cnt[i + 1] - cnt[i];
        qsort(row, (size_t)len, sizeof(entry_t), cmp_entry);
        uint64_t k = 0;
        while (k < len) {
            uint32_t j = row[k].col;
            int64_t acc = 0;
            while (k < len && row[k].col == j) acc += row[k++].q;
            if (acc != 0) {
                col_idx[nnz] = j;
                w[nnz] = acc;
                nnz++;
            }
        }
        row_ptr[i + 1] = (uint32_t)nnz;
    }
    free(adj);
    free(fill);
    free(cnt);
    return nnz;
}

/* ir.rs:85-103: E = offset + sum_i lin_i x_i + 1/2 sum_i x_i sum_row q x_j.
 * The pair sum over both triangles is even, so /2 is exact in integers. */
int64_t or_ir_energy(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                     const int64_t *w, const int64_t *linear, int64_t offset, const uint8_t *x) {
    int64_t lin = 0, pair = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!x[i]) continue;
        lin += linear[i];
        for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
            if (x[col_idx[e]]) pair += w[e];
    }
    return offset + lin + pair / 2;
}

/* ir.rs:114-139: ascending sites, smallest colour not used by an already
 * coloured neighbour. */
void or_greedy_coloring(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                        uint32_t *colors) {
    uint32_t *seen = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 2));
    for (uint32_t i = 0; i < n + 2; ++i) seen[i] = UINT32_MAX;
    for (uint32_t i = 0; i < n; ++i) colors[i] = UINT32_MAX;
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
            uint32_t c = colors[col_idx[e]];
            if (c != UINT32_MAX && c <= n) seen[c] = i;
        }
        uint32_t c = 0;
        while (seen[c] == i) c++;
        colors[i] = c;
    }
    free(seen);
}

/* metropolis_sweep.rs:40-58: colour 0 sites ascending, then colour 1, ... */
uint32_t or_color_order(uint32_t n, const uint32_t *colors, uint32_t *order,
                        uint32_t *class_start) {
    uint32_t ncol = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (colors[i] + 1 > ncol) ncol = colors[i] + 1;
    uint32_t pos = 0;
    for (uint32_t c = 0; c < ncol; ++c) {
        class_start[c] = pos;
        for (uint32_t i = 0; i < n; ++i)
            if (colors[i] == c) order[pos++] = i;
    }
    class_start[ncol] = pos;
    return ncol;
}

/* ------------------------------------------------------------------------- */
/* Fields and energies (sparse_bitslice.rs:88-137)                            */
/* ------------------------------------------------------------------------- */
static inline int bit(uint64_t word, uint32_t lane) { return (int)((word >> lane) & 1u); }

static int64_t field_of(const uint32_t *row_ptr, const uint32_t *col_idx, const int64_t *w,
                        const int64_t *linear, const uint64_t *slot_words, uint32_t site,
                        uint32_t lane) {
    int64_t h = linear[site];
    for (uint32_t e = row_ptr[site]; e < row_ptr[site + 1]; ++e)
        if (bit(slot_words[col_idx[e]], lane)) h += w[e];
    return h;
}

void or_fields(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const int64_t *w,
               const int64_t *linear, uint32_t cells, const uint64_t *words, int64_t *fields) {
    for (uint32_t s = 0; s < cells; ++s)
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t r = 0; r < 64; ++r)
                fields[((uint64_t)s * n + i) * 64 + r] =
                    field_of(row_ptr, col_idx, w, linear, words + (uint64_t)s * n, i, r);
}

void or_energies(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const int64_t *w,
                 const int64_t *linear, int64_t offset, uint32_t cells, const uint64_t *words,
                 int64_t *energies) {
    uint8_t *x = (uint8_t *)malloc((size_t)n + 1);
    for (uint32_t s = 0; s < cells; ++s)
        for (uint32_t r = 0; r < 64; ++r) {
            for (uint32_t i = 0; i < n; ++i) x[i] = (uint8_t)bit(words[(uint64_t)s * n + i], r);
            energies[s * 64 + r] = or_ir_energy(n, row_ptr, col_idx, w, linear, offset, x);
        }
    free(x);
}

void or_init_words(uint32_t n, uint32_t cells, uint64_t seed, uint64_t *words) {
    uint32_t o[4];
    for (uint32_t s = 0; s < cells; ++s)
        for (uint32_t i = 0; i < n; ++i) {
            philox_seeded(seed, i, s, 0, TAG_INIT << 24, o);
            words[(uint64_t)s * n + i] = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
        }
}

/* ------------------------------------------------------------------------- */
/* Uniforms for one (position, slot, sweep): 64 lanes.                        */
/* ------------------------------------------------------------------------- */
static void uniforms_word(uint64_t seed, uint32_t pos, uint32_t slot, uint32_t sweep,
                          uint32_t u[64]) {
    uint32_t o[4];
    for (uint32_t r = 0; r < 64; ++r) {
        philox_seeded(seed, pos >> 2, slot, sweep, (TAG_WORD << 24) | r, o);
        u[r] = o[pos & 3u];
    }
}

/* Plane view: the top EAGER_PLANES bits of u_r are bit r of planes
 * P_0..P_7 (P_{2j}, P_{2j+1} = words 0|1 and 2|3 of Philox(pos, slot, sweep,
 * TAG_PLANE<<24 | j)), most significant first; the remaining 24 bits are the
 * top bits of the replica's tail word Philox(pos, slot, sweep,
 * TAG_TAIL<<24 | r>>2)[r & 3]. */
static void uniforms_plane(uint64_t seed, uint32_t pos, uint32_t slot, uint32_t sweep,
                           uint32_t u[64]) {
    uint64_t plane[EAGER_PLANES];
    uint32_t o[4];
    for (uint32_t j = 0; j < EAGER_PLANES / 2; ++j) {
        philox_seeded(seed, pos, slot, sweep, (TAG_PLANE << 24) | j, o);
        plane[2 * j] = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
        plane[2 * j + 1] = (uint64_t)o[2] | ((uint64_t)o[3] << 32);
    }
    for (uint32_t r = 0; r < 64; ++r) {
        uint32_t v = 0;
        for (uint32_t b = 0; b < EAGER_PLANES; ++b) v |= (uint32_t)bit(plane[b], r) << (31u - b);
        philox_seeded(seed, pos, slot, sweep, (TAG_TAIL << 24) | (r >> 2), o);
        v |= o[r & 3u] >> EAGER_PLANES;
        u[r] = v;
    }
}

static uint32_t uniform_exchange(uint64_t seed, uint32_t t, uint32_t lane, uint32_t round) {
    uint32_t o[4];
    philox_seeded(seed, t, lane >> 2, round, TAG_EXCH << 24, o);
    return o[lane & 3u];
}

/* engine.rs:713-789 + :818: exchange number `round` uses the pairs (t, t+1)
 * with t = round mod 2 (the phase toggles after every exchange); lane r of
 * temperature t against lane r of temperature t+1; accept iff
 * delta = (beta_t - beta_{t+1}) (E_t - E_{t+1}) >= 0 or u < exp(delta).  The
 * configurations stay in their slots and the two lanes swap temperatures. */
void or_exchange_pops(uint32_t cells, uint32_t pop_temps, uint64_t seed, uint32_t round,
                      const double *beta, const int64_t *energies, uint32_t *temp_of,
                      uint64_t *stats) {
    const uint32_t phase = round & 1u;
    const uint32_t T = pop_temps ? pop_temps : cells;
    uint32_t *slot_of = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cells * 64);
    for (uint32_t s = 0; s < cells; ++s)
        for (uint32_t r = 0; r < 64; ++r) slot_of[temp_of[s * 64 + r] * 64 + r] = s;
    /* engine.rs:714-716: every population p, its pairs t = phase, phase + 2, ... */
    for (uint32_t base = 0; base + T <= cells; base += T)
    for (uint32_t t = base + phase; t + 1 < base + T; t += 2) {
        double dbeta = beta[t] - beta[t + 1];
        for (uint32_t r = 0; r < 64; ++r) {
            uint32_t s1 = slot_of[t * 64 + r], s2 = slot_of[(t + 1) * 64 + r];
            double delta = dbeta * (double)(energies[s1 * 64 + r] - energies[s2 * 64 + r]);
            uint32_t uu = uniform_exchange(seed, t, r, round);
            int accept = (delta >= 0.0) || (uu < or_exchange_threshold(delta));
            stats[2 + 2 * t] += 1;
            if (accept) {
                stats[3 + 2 * t] += 1;
                temp_of[s1 * 64 + r] = t + 1;
                temp_of[s2 * 64 + r] = t;
                slot_of[t * 64 + r] = s2;
                slot_of[(t + 1) * 64 + r] = s1;
            }
        }
    }
    free(slot_of);
}

void or_exchange(uint32_t cells, uint64_t seed, uint32_t round, const double *beta,
                 const int64_t *energies, uint32_t *temp_of, uint64_t *stats) {
    or_exchange_pops(cells, cells, seed, round, beta, energies, temp_of, stats);
}

/* apply_resample (population_annealing.rs:229-286) on MSC words: member
 * m = p*64 + r is lane r of population p over its levels t < T (temperature
 * index p*T + t, found through temp_of); afterwards the replica of member m
 * at level t holds what member parents[m] held there, spins and energy.
 * words [cells][n] site order. */
void or_resample(uint32_t n, uint32_t cells, uint32_t pop_temps, const uint32_t *parents,
                 uint64_t *words, const uint32_t *temp_of, int64_t *energies) {
    const uint32_t T = pop_temps ? pop_temps : cells;
    uint32_t *slot_of = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cells * 64);
    for (uint32_t s = 0; s < cells; ++s)
        for (uint32_t r = 0; r < 64; ++r) slot_of[temp_of[s * 64 + r] * 64 + r] = s;
    uint64_t *old = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)cells * n + 8);
    int64_t *old_e = (int64_t *)malloc(sizeof(int64_t) * (size_t)cells * 64);
    memcpy(old, words, sizeof(uint64_t) * (size_t)cells * n);
    memcpy(old_e, energies, sizeof(int64_t) * (size_t)cells * 64);
    for (uint32_t d = 0; d < cells; ++d)
        for (uint32_t r = 0; r < 64; ++r) {
            const uint32_t g = temp_of[d * 64 + r], p = g / T, t = g % T;
            const uint32_t m = parents[p * 64 + r], rs = m % 64;
            const uint32_t src = slot_of[((m / 64) * T + t) * 64 + rs];
            energies[d * 64 + r] = old_e[src * 64 + rs];
            for (uint32_t i = 0; i < n; ++i) {
                const uint64_t b = (old[(uint64_t)src * n + i] >> rs) & 1u;
                words[(uint64_t)d * n + i] = (words[(uint64_t)d * n + i] & ~(1ull << r)) | (b << r);
            }
        }
    free(slot_of);
    free(old);
    free(old_e);
}

/* ------------------------------------------------------------------------- */
/* The run: metropolis_sweep.rs:95-146 in colour order, then engine.rs:713-818 */
/* With view OR_HEATBATH the sweep is gibbs_color_sweep.rs:119-174 instead:    */
/* the new spin is u < or_heatbath_threshold(beta, h), word-view uniforms.     */
/* ------------------------------------------------------------------------- */
int or_run_temps(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const int64_t *w,
                 const int64_t *linear, uint32_t cells, uint32_t temps, uint64_t seed,
                 uint32_t view, uint32_t exchange, uint32_t sweep0, uint32_t xround0,
                 uint32_t sweeps, const double *betas, uint32_t betas_per_sweep,
                 uint32_t pop_temps, uint64_t *words, uint32_t *temp_of, int64_t *energies,
                 uint64_t *stats) {
    if (cells == 0 || view > OR_HEATBATH || (temps != cells && temps != cells * 64u)) return -1;
    if (temps != cells && exchange) return -1; /* the replica view exchanges separately */
    for (uint32_t i = 0; i < cells * 64u; ++i)
        if (temp_of[i] >= temps) return -1;
    uint32_t *colors = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    uint32_t *cstart = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 2));
    or_greedy_coloring(n, row_ptr, col_idx, colors);
    uint32_t ncol = or_color_order(n, colors, order, cstart);
    uint32_t u[64];

    for (uint32_t k = 0; k < sweeps; ++k) {
        uint32_t sweep = sweep0 + k;
        const double *beta = betas + (betas_per_sweep ? (uint64_t)k * temps : 0);
        for (uint32_t c = 0; c < ncol; ++c) {
            for (uint32_t pos = cstart[c]; pos < cstart[c + 1]; ++pos) {
                uint32_t site = order[pos];
                for (uint32_t s = 0; s < cells; ++s) {
                    uint64_t *sw = words + (uint64_t)s * n;
                    if (view == OR_VIEW_PLANE)
                        uniforms_plane(seed, pos, s, sweep, u);
                    else
                        uniforms_word(seed, pos, s, sweep, u);
                    uint64_t flip = 0;
                    for (uint32_t r = 0; r < 64; ++r) {
                        int64_t h = field_of(row_ptr, col_idx, w, linear, sw, site, r);
                        int x = bit(sw[site], r);
                        int64_t de = x ? -h : h;
                        double b = beta[temp_of[s * 64 + r]];
                        int accept;
                        if (view == OR_HEATBATH)
                            accept = ((uint64_t)u[r] < or_heatbath_threshold(b, h)) != x;
                        else
                            accept = (de <= 0) || (u[r] < or_threshold(b, de));
                        if (accept) {
                            flip |= 1ull << r;
                            energies[s * 64 + r] += de;
                        }
                    }
                    sw[site] ^= flip;
                    stats[0] += (uint64_t)__builtin_popcountll(flip);
                    stats[1] += 64;
                }
            }
        }
        if (exchange && cells >= 2)
            or_exchange_pops(cells, pop_temps, seed, xround0 + k, beta, energies, temp_of, stats);
    }
    free(colors);
    free(order);
    free(cstart);
    return 0;
}

int or_run(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const int64_t *w,
           const int64_t *linear, uint32_t cells, uint64_t seed, uint32_t view, uint32_t exchange,
           uint32_t sweep0, uint32_t xround0, uint32_t sweeps, const double *betas,
           uint32_t betas_per_sweep, uint32_t pop_temps, uint64_t *words, uint32_t *temp_of,
           int64_t *energies, uint64_t *stats) {
    return or_run_temps(n, row_ptr, col_idx, w, linear, cells, cells, seed, view, exchange,
                        sweep0, xround0, sweeps, betas, betas_per_sweep, pop_temps, words,
                        temp_of, energies, stats);
}

/* replica_exchange.rs:61-109 in the replica view: R = cells * 64 replicas,
 * replica rho at temperature 1 / beta[rho]; the configuration of replica rho
 * sits at position q = slot * 64 + lane with temp_of[q] = rho.  One round with
 * start `phase`: pairs (i, i + 1) for i = phase, phase + 2, ..., i + 1 < R,
 * delta = (beta_i - beta_{i+1}) (E_i - E_{i+1}), accept iff delta >= 0 or
 * u < exp(delta) with u = Philox(i >> 2, 0, xround, TAG_REPX << 24)[i & 3];
 * an accepted pair swaps the two configurations' replica indices (the
 * reference's swap_replicas, seen through the permutation).  The pairs of a
 * round are disjoint, so their order does not matter.
 * stats[2 + 2i] / [3 + 2i]: attempts / accepts of pair i. */
void or_exchange_replicas(uint32_t cells, uint64_t seed, uint32_t xround, uint32_t phase,
                          const double *beta, const int64_t *energies, uint32_t *temp_of,
                          uint64_t *stats) {
    const uint32_t R = cells * 64u;
    uint32_t *loc = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)R);
    for (uint32_t q = 0; q < R; ++q) loc[temp_of[q]] = q;
    for (uint32_t i = phase & 1u; i + 1 < R; i += 2) {
        const uint32_t qi = loc[i], qj = loc[i + 1];
        const double delta = (beta[i] - beta[i + 1]) * (double)(energies[qi] - energies[qj]);
        uint32_t o[4];
        philox_seeded(seed, i >> 2, 0, xround, TAG_REPX << 24, o);
        const uint32_t uu = o[i & 3u];
        const int accept = (delta >= 0.0) || (uu < or_exchange_threshold(delta));
        stats[2 + 2 * (uint64_t)i] += 1;
        if (accept) {
            stats[3 + 2 * (uint64_t)i] += 1;
            temp_of[qi] = i + 1;
            temp_of[qj] = i;
            loc[i] = qj;
            loc[i + 1] = qi;
        }
    }
    free(loc);
}

/* ------------------------------------------------------------------------- */
/* Ladders                                                                   */
/* ------------------------------------------------------------------------- */
/* Rust's f64::powi lowers to compiler-rt __powidf2 (binary exponentiation). */
static double powi_rt(double a, int b) {
    int recip = b < 0;
    double r = 1.0;
    while (1) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

/* decision.rs:133-142 */
void or_geometric_ladder(uint32_t k, double t_hi, double t_lo, double *temps) {
    double hi = t_hi > t_lo ? t_hi : t_lo;
    double lo = t_hi < t_lo ? t_hi : t_lo;
    if (lo < 1e-6) lo = 1e-6;
    if (k <= 1) {
        temps[0] = hi;
        return;
    }
    double ratio = pow(lo / hi, 1.0 / ((double)k - 1.0));
    for (uint32_t i = 0; i < k; ++i) temps[i] = hi * powi_rt(ratio, (int)i);
}

/* schedule.rs:21-43 over the CSR rows (both triangles, full q). */
void or_suggested_range(uint32_t n, const uint32_t *row_ptr, const int64_t *w,
                        const int64_t *linear, double *t_max, double *t_min) {
    double max_local = 0.0, min_nonzero = INFINITY;
    for (uint32_t i = 0; i < n; ++i) {
        double local = fabs((double)linear[i]);
        if (linear[i] != 0) min_nonzero = fmin(min_nonzero, fabs((double)linear[i]));
        for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
            local += fabs((double)w[e]);
            if (w[e] != 0) min_nonzero = fmin(min_nonzero, fabs((double)w[e]));
        }
        if (local > max_local) max_local = local;
    }
    if (max_local == 0.0 || !isfinite(min_nonzero)) {
        *t_max = 1.0;
        *t_min = 0.1;
        return;
    }
    double tmax = max_local / 0.69314718055994530942;
    double tmin = min_nonzero / log(100.0);
    if (tmax * 0.5 < tmin) tmin = tmax * 0.5;
    *t_max = tmax;
    *t_min = tmin;
}

/* ------------------------------------------------------------------------- */
/* Deterministic descent and the incumbent                                    */
/* ------------------------------------------------------------------------- */

/* dE of flipping site i of replica `lane`: (1 - 2 x_i) * field_i. */
static int64_t flip_gain(const uint32_t *row_ptr, const uint32_t *col_idx, const int64_t *w,
                         const int64_t *linear, const uint64_t *slot_words, uint32_t site,
                         uint32_t lane) {
    const int64_t h = field_of(row_ptr, col_idx, w, linear, slot_words, site, lane);
    return bit(slot_words[site], lane) ? -h : h;
}

/* steepest_descent.rs:80-124 (SteepestDescent::apply) over every replica of
 * every slot: per step each replica flips the site of its most negative dE
 * (strict <, so the lowest site wins a tie and dE = 0 never moves); the loop
 * ends after `max_steps` steps or after a step in which no replica moved.
 * Clamped sites (local_search.rs is_clamped) are never chosen. */
uint32_t or_steepest_descent(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                             const int64_t *w, const int64_t *linear, const uint8_t *clamped,
                             uint32_t cells, uint32_t max_steps, uint64_t *words,
                             int64_t *energies, uint32_t *flips) {
    const uint64_t reps = (uint64_t)cells * 64;
    int64_t *best_de = (int64_t *)malloc(reps * sizeof(int64_t));
    uint32_t *best_site = (uint32_t *)malloc(reps * sizeof(uint32_t));
    uint32_t steps = 0;
    for (uint32_t k = 0; k < max_steps; ++k) {
        for (uint64_t q = 0; q < reps; ++q) {
            best_de[q] = 0;
            best_site[q] = UINT32_MAX;
        }
        for (uint32_t s = 0; s < cells; ++s) {
            const uint64_t *sw = words + (uint64_t)s * n;
            for (uint32_t i = 0; i < n; ++i) {
                if (clamped && clamped[i]) continue;
                for (uint32_t r = 0; r < 64; ++r) {
                    const int64_t de = flip_gain(row_ptr, col_idx, w, linear, sw, i, r);
                    if (de < best_de[s * 64 + r]) {
                        best_de[s * 64 + r] = de;
                        best_site[s * 64 + r] = i;
                    }
                }
            }
        }
        int any = 0;
        for (uint32_t s = 0; s < cells; ++s)
            for (uint32_t r = 0; r < 64; ++r) {
                const uint64_t q = (uint64_t)s * 64 + r;
                if (best_site[q] == UINT32_MAX) continue;
                words[(uint64_t)s * n + best_site[q]] ^= 1ull << r;
                energies[q] += best_de[q];
                flips[q] += 1;
                any = 1;
            }
        if (!any) break;
        ++steps;
    }
    free(best_de);
    free(best_site);
    return steps;
}

/* local_search.rs:19-60 (steepest_descent_1opt) on one state of bytes:
 * gains g_i = (1 - 2x_i) field_i kept incrementally, the most negative gain
 * flipped (lowest index on ties) until none is negative.  Returns the total
 * improvement (>= 0); *flips receives the number of flips. */
int64_t or_polish_1opt(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                       const int64_t *w, const int64_t *linear, const uint8_t *clamped,
                       uint8_t *x, uint64_t *flips) {
    int64_t *g = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    for (uint32_t i = 0; i < n; ++i) {
        int64_t h = linear[i];
        for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
            if (x[col_idx[e]]) h += w[e];
        g[i] = x[i] ? -h : h;
    }
    int64_t improvement = 0;
    uint64_t count = 0;
    for (;;) {
        uint32_t best = UINT32_MAX;
        int64_t best_g = 0;
        for (uint32_t i = 0; i < n; ++i)
            if (!(clamped && clamped[i]) && g[i] < best_g) {
                best = i;
                best_g = g[i];
            }
        if (best == UINT32_MAX) break;
        improvement -= best_g;
        ++count;
        const int64_t dir = x[best] ? -1 : 1; /* +1 for 0 -> 1 */
        x[best] ^= 1u;
        g[best] = -best_g;
        for (uint32_t e = row_ptr[best]; e < row_ptr[best + 1]; ++e) {
            const uint32_t j = col_idx[e];
            g[j] += (x[j] ? -1 : 1) * w[e] * dir;
        }
    }
    free(g);
    if (flips) *flips = count;
    return improvement;
}

/* ultimate.rs:397-431 (update_incumbent): the replica with the lowest
 * tracked energy, scanned in (temperature, lane) order with strict <, replaces
 * the incumbent when its energy is strictly lower (or no incumbent exists).
 * words are [cells][n] in site order; best_x receives n bytes.  Returns 1 if
 * the incumbent changed. */
int or_update_best_view(uint32_t n, uint32_t cells, uint32_t replica_view, const uint64_t *words,
                        const uint32_t *temp_of, const int64_t *energies, int64_t *best_e,
                        int32_t *valid, uint8_t *best_x) {
    uint64_t arg = UINT64_MAX;
    int64_t e_min = 0;
    /* the replica index of the current minimum: temperature * 64 + lane, or
     * in the replica view temp_of itself */
    uint64_t key_min = UINT64_MAX;
    for (uint32_t s = 0; s < cells; ++s)
        for (uint32_t r = 0; r < 64; ++r) {
            const uint64_t q = (uint64_t)s * 64 + r;
            const uint64_t key = replica_view ? (uint64_t)temp_of[q] : (uint64_t)temp_of[q] * 64 + r;
            if (arg == UINT64_MAX || energies[q] < e_min ||
                (energies[q] == e_min && key < key_min)) {
                arg = q;
                e_min = energies[q];
                key_min = key;
            }
        }
    if (arg == UINT64_MAX || (*valid && e_min >= *best_e)) return 0;
    const uint32_t s = (uint32_t)(arg / 64), r = (uint32_t)(arg % 64);
    for (uint32_t i = 0; i < n; ++i) best_x[i] = (uint8_t)bit(words[(uint64_t)s * n + i], r);
    *best_e = e_min;
    *valid = 1;
    return 1;
}

int or_update_best(uint32_t n, uint32_t cells, const uint64_t *words, const uint32_t *temp_of,
                   const int64_t *energies, int64_t *best_e, int32_t *valid, uint8_t *best_x) {
    return or_update_best_view(n, cells, 0, words, temp_of, energies, best_e, valid, best_x);
}

/* ------------------------------------------------------------------------- */
/* Isoenergetic cluster moves (icm.rs:91-185)                                 */
/* ------------------------------------------------------------------------- */

/* icm_sweep (icm.rs:163-185): for every temperature t and r < 32 a Houdayer
 * move on the replicas (t, r) and (t, r + 32), found through temp_of (the
 * configurations stay in their slots).  The seed is the floor(u |D| / 2^32)-th
 * disagreeing site in site order, u = Philox(t, r, round, TAG_ICM<<24)[0];
 * the cluster is the seed's connected component of disagreeing sites; both
 * replicas flip it, E_A += dE_A and E_B -= dE_A (icm.rs:91-160).  words:
 * [cells][n] site order.  flipped[cells*32] (or NULL) receives each pair's
 * cluster size.  Returns the number of moves made. */
uint32_t or_icm_sweep(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                      const int64_t *w, const int64_t *linear, uint32_t cells, uint64_t seed,
                      uint32_t round, uint64_t *words, const uint32_t *temp_of,
                      int64_t *energies, uint32_t *flipped) {
    uint32_t *slot_of = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cells * 64);
    for (uint32_t s = 0; s < cells; ++s)
        for (uint32_t r = 0; r < 64; ++r) slot_of[temp_of[s * 64 + r] * 64 + r] = s;
    uint8_t *in_c = (uint8_t *)malloc((size_t)n + 1);
    uint8_t *dis = (uint8_t *)malloc((size_t)n + 1);
    uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    uint32_t moves = 0;
    for (uint32_t t = 0; t < cells; ++t)
        for (uint32_t r = 0; r < 32; ++r) {
            const uint32_t sa = slot_of[t * 64 + r], sb = slot_of[t * 64 + r + 32];
            uint64_t *A = words + (uint64_t)sa * n, *B = words + (uint64_t)sb * n;
            uint32_t count = 0;
            for (uint32_t i = 0; i < n; ++i) {
                dis[i] = (uint8_t)(bit(A[i], r) ^ bit(B[i], r + 32));
                in_c[i] = 0;
                count += dis[i];
            }
            if (flipped) flipped[t * 32 + r] = 0;
            if (count == 0) continue;
            uint32_t o[4];
            philox_seeded(seed, t, r, round, TAG_ICM << 24, o);
            uint32_t k = (uint32_t)(((uint64_t)o[0] * count) >> 32), sd = 0;
            for (uint32_t i = 0; i < n; ++i)
                if (dis[i] && k-- == 0) {
                    sd = i;
                    break;
                }
            uint32_t top = 0, size = 0;
            stack[top++] = sd;
            in_c[sd] = 1;
            while (top) {
                const uint32_t i = stack[--top];
                ++size;
                for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
                    const uint32_t j = col_idx[e];
                    if (dis[j] && !in_c[j]) {
                        in_c[j] = 1;
                        stack[top++] = j;
                    }
                }
            }
            /* icm.rs:55-82: dE of the cluster flip in replica A */
            int64_t de = 0;
            for (uint32_t i = 0; i < n; ++i) {
                if (!in_c[i]) continue;
                const int64_t xi = bit(A[i], r);
                de += linear[i] * (1 - 2 * xi);
                for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
                    const uint32_t j = col_idx[e];
                    const int64_t xj = bit(A[j], r);
                    if (in_c[j]) {
                        if (i < j) de += w[e] * ((1 - xi) * (1 - xj) - xi * xj);
                    } else {
                        de += w[e] * xj * (1 - 2 * xi);
                    }
                }
            }
            for (uint32_t i = 0; i < n; ++i)
                if (in_c[i]) {
                    A[i] ^= 1ull << r;
                    B[i] ^= 1ull << (r + 32);
                }
            energies[sa * 64 + r] += de;
            energies[sb * 64 + r + 32] -= de;
            if (flipped) flipped[t * 32 + r] = size;
            ++moves;
        }
    free(slot_of);
    free(in_c);
    free(dis);
    free(stack);
    return moves;
}
