/*
 * ising_gpu.h -- C ABI of the MI355X multispin Metropolis / parallel-tempering
 * engine (gfx950, one GPU).
 *
 * Conventions
 *   - Every call returns int32 status: IGE_OK (0) or a negative IGE_ERR_*.
 *     ige_last_error() returns the message of the calling thread's last
 *     failure.  No C++ exception crosses this boundary.
 *   - The library owns device buffers; the caller owns host buffers.
 *   - A state is externally synchronised: one host thread at a time.  Distinct
 *     states run on distinct HIP streams and may be driven concurrently.
 *   - A "cell" (slot) holds 64 replicas of the model, one bit each, packed as
 *     one uint64 per site ("MSC word": bit r = replica r).
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   ige_model_from_pairs    ProblemIR::from_pairs      src/engine_v2/ir.rs:30-69
 *   ige_model_coloring      ProblemIR::greedy_coloring src/engine_v2/ir.rs:114-139
 *   ige_model_csr,          ProblemIR row_ptr / col_idx / weights, linear, offset
 *   ige_model_linear                                   src/engine_v2/ir.rs:14-24
 *   ige_model_order         MetropolisSweep::ensure_order
 *                                                      src/engine_v2/operators/metropolis_sweep.rs:40-58
 *   ige_model_get_info      ProblemIR::is_integral + DecisionEngine backend choice
 *                                                      src/engine_v2/ir.rs:143-148,
 *                                                      src/engine_v2/decision.rs:115-124
 *   ige_state_create        SparseBitSlice::new        src/engine_v2/backends/sparse_bitslice.rs:42-81
 *                           QuantumField::new + random init
 *                                                      src/solver/types.rs:23-58,
 *                                                      src/solver/ultimate.rs:271-275
 *   ige_run                 MetropolisSweep::apply     src/engine_v2/operators/metropolis_sweep.rs:95-146
 *                           engine::step (sweep + DEO) src/solver/engine.rs:573-819
 *                           with IGE_RUN_HEATBATH: GibbsColorSweep::apply
 *                                                      src/engine_v2/operators/gibbs_color_sweep.rs:119-174
 *   ige_exchange            engine::step Phase 2 (DEO schedule)
 *                                                      src/solver/engine.rs:713-818
 *   ige_set_replica_temperatures
 *                           RuntimeView.temperatures as MetropolisSweep reads them,
 *                           T = temps[rep % temps.len()]
 *                                                      src/engine_v2/operators/metropolis_sweep.rs:109-116
 *   ige_exchange_replicas   ReplicaExchange::apply     src/engine_v2/operators/replica_exchange.rs:61-109
 *   ige_read_energies       SpinState::energies_into   src/engine_v2/backends/sparse_bitslice.rs:273-278
 *   ige_read_fields         SparseBitSlice fields ledger; on a dense state the fields
 *                           are recomputed on int8 matrix cores (the dense SK / QUBO
 *                           instances of src/bin/sk_parisi_benchmark.rs:56-110)
 *                                                      src/engine_v2/backends/sparse_bitslice.rs:26-27
 *   ige_recompute_energies  SpinState::audit / energy_of
 *                                                      src/engine_v2/backends/sparse_bitslice.rs:114-137,317-346
 *   ige_read_spins          SpinState::spin / extract_into
 *                                                      src/engine_v2/backends/sparse_bitslice.rs:280-289
 *   ige_write_spins         SparseBitSlice::rebuild    src/engine_v2/backends/sparse_bitslice.rs:70-95
 *   ige_server_*            POST /api/v1/solve request loop
 *                                                      src/bin/server_api.rs:111-213;
 *                           ige_server_solution: the solution UltimateSolver::solve returns
 *                                                      src/solver/ultimate.rs:299-374
 *   ige_ultimate_ladder     UltimateSolver's geometric ladder
 *                                                      src/solver/ultimate.rs:237-246
 *   ige_batch_*             no single counterpart: many UltimateSolver-style anneals
 *                           (src/solver/ultimate.rs:234-336) in one launch
 *   ige_read_stats          Report{accepted,proposed} + StepScratch swap statistics
 *                                                      src/engine_v2/operator.rs:34-52,
 *                                                      src/solver/engine.rs:454-557
 *   ige_problem_parse       parse_rudy + maxcut_to_qubo src/engine_v2/frontend.rs:40-84;
 *                           parse_rudy_maxcut / parse_orlib_bqp / parse_biqmac_sparse /
 *                           parse_qplib                src/benchmark/instances.rs:103-344
 *   ige_problem_get_info    ProblemIR::is_integral     src/engine_v2/ir.rs:143-148
 *   ige_set_populations,    PA-PT hybrid of UltimateSolver::solve_population_annealing
 *   ige_resample, ige_pa_*                             src/solver/ultimate.rs:559-852,
 *                           apply_resample and helpers src/solver/population_annealing.rs:45-286
 *   ige_icm, IGE_RUN_ICM    icm_sweep / icm_pair_move  src/solver/icm.rs:91-185
 *                           (called per exchange round, src/solver/ultimate.rs:326-328)
 *   ige_steepest_descent    SteepestDescent::apply     src/engine_v2/operators/steepest_descent.rs:80-124
 *   ige_model_polish        steepest_descent_1opt      src/solver/local_search.rs:19-60
 *   IGE_RUN_TRACK_BEST,     UltimateSolver::update_incumbent and the anytime
 *   ige_update_best,        best-state return          src/solver/ultimate.rs:299-336,
 *   ige_read_best                                      src/solver/ultimate.rs:369-373,397-431
 */
#ifndef ISING_GPU_H
#define ISING_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IGE_OK 0
#define IGE_ERR_INVALID (-1)     /* bad argument or inconsistent input      */
#define IGE_ERR_HIP (-2)         /* a HIP runtime call failed               */
#define IGE_ERR_NOMEM (-3)       /* device or host allocation failed        */
#define IGE_ERR_UNSUPPORTED (-4) /* model/kernel combination not supported  */
#define IGE_ERR_RANGE (-5)       /* value outside the exact integer range   */

#define IGE_KERNEL_AUTO 0      /* multispin when eligible, else dense when the
                                  model is dense and eligible, else generic */
#define IGE_KERNEL_GENERIC 1   /* lane = replica, any integer CSR model     */
#define IGE_KERNEL_MULTISPIN 2 /* lane = site, bit-sliced +-J arithmetic    */
#define IGE_KERNEL_DENSE 3     /* dense q = scale*int8: positions in order,
                                  fields recomputed on int8 matrix cores    */

#define IGE_RUN_EXCHANGE 1u   /* DEO replica exchange after every sweep     */
#define IGE_RUN_TRACK_BEST 2u /* incumbent update after every sweep (and its
                                 exchange), see ige_update_best             */
#define IGE_RUN_ICM 8u        /* one ICM sweep after every sweep (and its
                                 exchange), before the incumbent update      */
#define IGE_RUN_HEATBATH 4u   /* heat-bath (Gibbs) sweeps instead of Metropolis:
                                 new spin = 1 with P = 1/(1+exp(beta h)),
                                 lane-per-replica kernel on any state       */
#define IGE_RUN_EAGER 16u     /* one launch per colour class and sweep from
                                 the host: no captured sweep-loop graph and
                                 no resident run; the results are the same
                                 bits either way                            */

#define IGE_STEPS_UNLIMITED 0xFFFFFFFFu /* descend until no replica can improve */

typedef struct ige_ctx ige_ctx;
typedef struct ige_model ige_model;
typedef struct ige_state ige_state;

typedef struct {
    uint32_t n;
    uint32_t num_colors;
    uint64_t nnz;
    uint32_t max_degree;
    int32_t multispin_eligible; /* 1 when every |q_ij| == w0 and the Ising field is 0 */
    int32_t coupling_magnitude; /* w0 when eligible, else 0                        */
    int32_t dense_scale;     /* s when every q_ij = s * int8 and n <= 8192, else 0 */
    int64_t max_local_field; /* max_i |lin_i| + sum_j |q_ij|: bound on |dE|      */
} ige_model_info;

typedef struct {
    uint64_t sweeps;    /* sweeps run on this state                      */
    uint64_t exchanges; /* DEO exchange rounds run on this state         */
    uint64_t accepted;  /* accepted single-spin proposals                */
    uint64_t proposed; /* proposals (= sweeps * n * 64 * cells)         */
    uint32_t phase;    /* DEO phase of the next exchange (= exchanges & 1) */
    uint32_t kernel;   /* IGE_KERNEL_GENERIC, _MULTISPIN or _DENSE      */
    uint64_t graph_sweeps;    /* sweeps run by replays of the captured loop */
    uint64_t graph_fallbacks; /* ige_run calls whose capture was invalidated
                                 (another thread synchronised the device
                                 meanwhile) and that ran eagerly instead    */
    uint64_t resident_sweeps; /* sweeps run with the slot's words in LDS for
                                 the whole call (small regular lattices,
                                 no exchange)                               */
    uint64_t chain_sweeps;    /* sweeps run by one launch over every colour
                                 class of the call (one-slot multispin states
                                 of irregular degree, no exchange,
                                 incumbent, heat bath or ICM)               */
} ige_stats;

/* ---- context ---- */
int32_t ige_context_create(int32_t device, ige_ctx **out);
int32_t ige_context_destroy(ige_ctx *ctx);
const char *ige_last_error(void);
const char *ige_version(void);

/* ---- model (integer QUBO: E = offset + sum lin_i x_i + sum_{i<j} q_ij x_i x_j) ---- */
int32_t ige_model_from_pairs(ige_ctx *ctx, uint32_t n, int64_t offset, const int32_t *linear,
                             uint64_t npairs, const uint32_t *pi, const uint32_t *pj,
                             const int32_t *q, ige_model **out);
int32_t ige_model_destroy(ige_model *model);
int32_t ige_model_get_info(const ige_model *model, ige_model_info *info);
int32_t ige_model_coloring(const ige_model *model, uint32_t *colors /* n */);
/* The sweep's position order (MetropolisSweep::ensure_order,
 * src/engine_v2/operators/metropolis_sweep.rs:40-58): order[p] = the site
 * visited p-th (colour 0's sites ascending, then colour 1's, ...);
 * class_start: num_colors + 1 class bounds in position space (or NULL). */
int32_t ige_model_order(const ige_model *model, uint32_t *order /* n */,
                        uint32_t *class_start /* num_colors+1 */);
int32_t ige_model_csr(const ige_model *model, uint32_t *row_ptr /* n+1 */,
                      uint32_t *col_idx /* nnz */, int32_t *w /* nnz */);
/* ProblemIR's linear terms and offset (src/engine_v2/ir.rs:14-24); either
 * pointer may be NULL. */
int32_t ige_model_linear(const ige_model *model, int32_t *linear /* n */, int64_t *offset);

/* ---- instance formats (host only, no device) ---- */
typedef struct ige_problem ige_problem;

#define IGE_FORMAT_RUDY 0u   /* rudy MaxCut as the engine_v2 frontend reads it:
                                `u v [w]`, w = 1 when absent, bad lines skipped */
#define IGE_FORMAT_GSET 1u   /* rudy MaxCut as the benchmark parser reads it
                                (G-Set, Biq Mac MaxCut): `i j w`, '#'/'%'
                                comments, out-of-range vertices an error     */
#define IGE_FORMAT_ORLIB 2u  /* OR-Library BQP, maximise x'Qx; index = problem */
#define IGE_FORMAT_BIQMAC 3u /* Biq Mac sparse BQP, minimise x'Qx            */
#define IGE_FORMAT_QPLIB 4u  /* QPLIB, binary unconstrained (QBN) only        */

typedef struct {
    uint32_t n;
    uint32_t problems_in_file; /* OR-Library files hold several             */
    uint64_t npairs;           /* distinct pairs i < j with a nonzero sum     */
    int32_t maximize;          /* native objective = -energy                  */
    int32_t integral;          /* every coefficient an integer (is_integral)  */
    double offset;
} ige_problem_info;

/* Parses `len` bytes of text into the minimisation form E = offset + lin.x +
 * sum q_ij x_i x_j (f64, the reference's conventions per format).  A format
 * error is IGE_ERR_INVALID with the parser's message. */
int32_t ige_problem_parse(const char *text, uint64_t len, uint32_t format, uint32_t index,
                          ige_problem **out);
int32_t ige_problem_destroy(ige_problem *problem);
int32_t ige_problem_get_info(const ige_problem *problem, ige_problem_info *info);
/* The integer arrays ige_model_from_pairs takes (pairs sorted, i < j):
 * IGE_ERR_RANGE when a coefficient is not an integer or exceeds int32. */
int32_t ige_problem_pairs(const ige_problem *problem, int64_t *offset, int32_t *linear /* n */,
                          uint32_t *pi, uint32_t *pj, int32_t *q /* npairs each */);
/* f64 energy of x (n bytes 0/1). */
int32_t ige_problem_energy(const ige_problem *problem, const uint8_t *x, double *energy);

/* ---- replica state ---- */
/* init_words: [cells][n] in site order, or NULL for the seeded default. */
int32_t ige_state_create(ige_ctx *ctx, const ige_model *model, uint32_t cells, uint64_t seed,
                         const uint64_t *init_words, uint32_t kernel, ige_state **out);
int32_t ige_state_destroy(ige_state *state);
/* beta[t] for temperature index t (cells entries, finite, >= 0). */
int32_t ige_set_ladder(ige_state *state, const double *beta);
/* beta_schedule: NULL (use the ladder) or [sweeps][cells]; the ladder is left
 * at the last row.  Asynchronous on the state's stream.  A multispin state of
 * a regular lattice (degree 4 or 6, n <= 4096) run without exchange,
 * incumbent, heat-bath or ICM does the whole call in one resident launch (one
 * workgroup per slot, words in LDS).  Otherwise runs of 32 sweeps or more
 * replay a hipGraph of 32 captured steps (kept with the state and captured
 * again when the flags, populations or schedule buffer change); the rest,
 * IGE_RUN_ICM, profiling and the dense kernel launch eagerly. */
int32_t ige_run(ige_state *state, uint32_t sweeps, uint32_t flags, const double *beta_schedule);
/* `rounds` DEO exchanges without sweeps, on the current ladder and tracked
 * energies.  Exchange number k uses pairs (t, t+1) with t = k mod 2 and is
 * keyed on k, whether it runs here or after a sweep in ige_run. */
int32_t ige_exchange(ige_state *state, uint32_t rounds);

/* The replica view, the engine_v2 operators' own semantics.  Replica
 * rho runs at T = temps[rho % len] (finite, > 0), as MetropolisSweep::apply
 * reads RuntimeView.temperatures.  The first call on a state in the
 * slot-ladder view switches it: the configuration at (temperature index t,
 * lane) becomes replica t * 64 + lane (slot * 64 + lane on a state that never
 * exchanged), and temp_of (ige_read_temp_of) then holds replica indices;
 * ige_run sweeps on those temperatures (no
 * schedule, IGE_RUN_EXCHANGE or IGE_RUN_ICM), ige_update_best breaks energy
 * ties by replica index and ige_read_stats reports the cells*64 - 1 replica
 * pairs.  Later calls change the temperatures only.  ige_set_ladder switches
 * back (temperature indices = slot indices again, identity permutation);
 * ICM, populations and resampling need the slot-ladder view. */
int32_t ige_set_replica_temperatures(ige_state *state, const double *temps, uint32_t len);
/* ReplicaExchange::apply with budget.sweeps = rounds, in the replica view:
 * round k of this call pairs replicas (i, i + 1) for i = (k & 1), (k & 1) + 2,
 * ...; delta = (beta_i - beta_{i+1}) (E_i - E_{i+1}); accept iff delta >= 0 or
 * u < floor(exp(delta) 2^32), u = Philox4x32-10(i >> 2, 0, x, 8 << 24)[i & 3]
 * with x the state's exchange counter (one per round, shared with
 * ige_exchange); an accepted pair swaps the two configurations' replica
 * indices.  Asynchronous. */
int32_t ige_exchange_replicas(ige_state *state, uint32_t rounds);
int32_t ige_synchronize(ige_state *state);

int32_t ige_read_spins(ige_state *state, uint64_t *words /* [cells][n] */);
/* Replaces every slot's words ([cells][n], site order) and recomputes the
 * tracked energies exactly (SparseBitSlice::rebuild,
 * src/engine_v2/backends/sparse_bitslice.rs:70-95); the temperature
 * permutation, counters and ladder are kept.  Synchronous. */
int32_t ige_write_spins(ige_state *state, const uint64_t *words);
int32_t ige_read_energies(ige_state *state, int64_t *energies /* [cells][64] */);
int32_t ige_read_temp_of(ige_state *state, uint32_t *temp_of /* [cells][64] */);
int32_t ige_read_fields(ige_state *state, int32_t *fields /* [cells][n][64] */);
int32_t ige_recompute_energies(ige_state *state, int64_t *energies /* [cells][64] */);
/* pair_attempts / pair_accepts: one entry per adjacent pair of temperature
 * indices, cells - 1 (cells*64 - 1 in the replica view) each, or NULL. */
int32_t ige_read_stats(ige_state *state, ige_stats *stats, uint64_t *pair_attempts,
                       uint64_t *pair_accepts);

/* ---- deterministic descent and the incumbent ---- */
typedef struct {
    uint32_t steps;     /* steps in which at least one replica flipped          */
    uint32_t converged; /* 1 if a step found no improving flip (local minimum)  */
    uint64_t flips;     /* flips over all replicas                              */
    uint64_t proposed;  /* (steps + converged) * n * 64 * cells, as the
                           reference counts one scan of every (site, replica) */
} ige_descent_report;

/* Steepest descent of every replica: per step each replica flips the site
 * with the most negative dE (lowest site on a tie); at most max_steps steps
 * (IGE_STEPS_UNLIMITED: until a step finds nothing).  clamped: NULL or n bytes
 * in site order, nonzero = never flipped.  Tracked energies stay exact;
 * sweep/exchange counters and Metropolis statistics are not touched.
 * Synchronous. */
int32_t ige_steepest_descent(ige_state *state, uint32_t max_steps, const uint8_t *clamped,
                             ige_descent_report *report);
/* ige_steepest_descent without any allocation, for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference.  Same arguments, steps
 * and results.  Its scratch (per-replica argmin keys, control words, flip
 * counts, the clamp map and, when the model's dense matrix already exists,
 * the matrix-core fields) lives in one fixed 16 MiB device workspace per GPU,
 * held from library load until the process ends; a state needing more is
 * IGE_ERR_INVALID.  Calls on one GPU share that workspace: calls from host
 * threads of this process take turns, and each call has finished all its
 * device work when it returns. */
int32_t ige_steepest_descent_noalloc(ige_state *state, uint32_t max_steps, const uint8_t *clamped,
                                     ige_descent_report *report);

/* `rounds` ICM sweeps (Houdayer moves, icm.rs:163-185): for every temperature t
 * and r < 32 the replicas (t, r) and (t, r + 32) flip, both, the connected
 * cluster of sites where they disagree that contains a random seed site;
 * tracked energies move by +-dE exactly.  Move number k of a pair is keyed on
 * the state's ICM counter.  flipped: NULL or the total cluster size added up. */
int32_t ige_icm(ige_state *state, uint32_t rounds, uint64_t *flipped);

/* ---- population annealing (PA, and the PA-PT hybrid) ---- */
/* Splits the cells into `pops` populations of T = cells / pops temperature
 * indices each (population p owns p*T .. p*T+T-1; set the ladder to the same
 * T betas for every population).  Exchange then pairs (t, t+1) only inside a
 * population (engine.rs:713-716).  pops = 1 (the default) is plain PT. */
int32_t ige_set_populations(ige_state *state, uint32_t pops);
/* apply_resample (population_annealing.rs:229-286): member m = p*64 + r is
 * lane r of population p across its T levels; afterwards the replica of
 * member m at level t holds what member parents[m] held at level t, spins and
 * tracked energy.  parents: pops*64 entries, each < pops*64. */
int32_t ige_resample(ige_state *state, const uint32_t *parents);
/* Host-side PA arithmetic (population_annealing.rs:45-124, :237-246), on
 * count = pops*64 members.  normalize: log-sum-exp weights and ESS = 1/sum w^2;
 * free_energy_increment: logsumexp(log_w) - ln count; systematic_resample:
 * offspring counts from one offset u in [0, 1); parents: the parent list of
 * the counts, shuffled (Fisher-Yates with Philox uniforms keyed on seed and
 * the resampling event number); family_entropy: rho = count * sum_f (n_f /
 * count)^2 and the number of surviving families. */
int32_t ige_pa_normalize(const double *log_w, uint32_t count, double *w_norm, double *ess);
int32_t ige_pa_free_energy_increment(const double *log_w, uint32_t count, double *out);
int32_t ige_pa_systematic_resample(const double *w_norm, uint32_t count, double u,
                                   uint32_t *counts);
int32_t ige_pa_parents(const uint32_t *counts, uint32_t count, uint64_t seed, uint32_t event,
                       uint32_t *parents);
int32_t ige_pa_family_entropy(const uint32_t *family, uint32_t count, double *rho,
                              uint32_t *families);
/* adapt_spacings (population_annealing.rs:137-148): out_t = d_t sqrt(max(A_t,
 * 0.02)) rescaled so that the spacings keep their total (sums in index order);
 * uniform acceptance is a fixed point. */
int32_t ige_pa_adapt_spacings(const double *spacings, const double *acceptance, uint32_t count,
                              double *out);
/* The acceptance-uniform ladder tuning of the PA-PT solve (ultimate.rs:599-660,
 * LadderMode::AcceptanceUniform), on a throwaway state whose populations are
 * set: `iters` times, the ladder T_t = temp_max (temp_min / temp_max)^alpha_t
 * (t < T = cells / pops, the same for every population), `steps` sweeps with a
 * DEO exchange each, the swap acceptance of pair (t, t + 1) summed over the
 * populations, and alpha_1.. = the running sums of ige_pa_adapt_spacings of
 * the spacings alpha_{t+1} - alpha_t, alpha_{T-1} = 1.  alphas: [T], in (the
 * starting profile, normally t / (T - 1)) and out.  T < 3, iters == 0 or
 * steps == 0 leave it unchanged, as the reference does. */
int32_t ige_pa_tune_ladder(ige_state *state, double temp_max, double temp_min, uint32_t iters,
                           uint32_t steps, double *alphas);

/* Incumbent: the replica with the lowest tracked energy, ties to the lowest
 * (temperature index, lane), replaces the best-seen state when strictly lower
 * (or when there is none yet).  ige_update_best does one such update now;
 * ige_run with IGE_RUN_TRACK_BEST does one after every sweep.  Asynchronous. */
int32_t ige_update_best(ige_state *state);
/* x: n bytes in site order (or NULL); energy: its exact energy; sweep: the
 * sweeps completed when it was found (or NULL).  IGE_ERR_INVALID before the
 * first update. */
int32_t ige_read_best(ige_state *state, uint8_t *x, int64_t *energy, uint32_t *sweep);

/* Polish `count` states of n bytes each ([count][n], site order, in/out) to
 * 1-opt local minima, all at once (64 states per cell).  clamped: NULL or n
 * bytes.  energy: [count] exact energies after, flips: [count] (or NULL). */
int32_t ige_model_polish(const ige_model *model, uint32_t count, uint8_t *x,
                         const uint8_t *clamped, int64_t *energy, uint32_t *flips);

/* ---- batches: thousands of independent instances in one persistent launch ---- */
typedef struct ige_batch ige_batch;

typedef struct {
    uint32_t n;
    int64_t offset;
    const int32_t *linear; /* n                                    */
    uint64_t npairs;
    const uint32_t *pi;    /* npairs, pi < pj                      */
    const uint32_t *pj;
    const int32_t *q;
    uint64_t seed;
    uint32_t sweeps;       /* anneal length                        */
    double beta_lo;        /* schedule of ige_anneal_schedule      */
    double beta_hi;
    /* Serving only (a batch ignores both): incumbent update after every
     * track_every-th sweep and after the last (UltimateSolver's update after
     * each exchange round, ultimate.rs:329-335); 0 = none.  polish != 0:
     * the solution is polished to a 1-opt local minimum (ultimate.rs:374). */
    uint32_t track_every;
    uint32_t polish;
    /* Serving only: num_temps >= 1 runs UltimateSolver's static-ladder loop
     * (ultimate.rs:234-374) instead of the anneal: num_temps cells of 64
     * replicas at the ladder of ige_ultimate_ladder(temp_max, temp_min,
     * num_temps), a DEO exchange after every sweep, the incumbent updated
     * after every track_every-th sweep (sweeps_per_exchange; it must divide
     * sweeps = sweeps_per_exchange * exchanges), and the reference's solution:
     * the incumbent when strictly below the best final replica, else that
     * replica, then polish.  beta_lo and beta_hi are not used then.  0: the
     * anneal (a batch takes only anneals). */
    uint32_t num_temps;
    double temp_max;
    double temp_min;
} ige_instance_desc;

#define IGE_BATCH_PERSISTENT 0u /* one hipGraph launch (counter resets and
                                   one persistent kernel over a device work
                                   queue), captured once per (waves, k0, k1) */
#define IGE_BATCH_PLAIN 1u      /* per instance, per sweep, per colour launches */

/* beta_k = beta_hi * exp(-(sweeps-1-k) c), c = ln(beta_hi/beta_lo)/(sweeps-1),
 * exp evaluated deterministically; geometric from beta_lo to beta_hi.  A batch
 * instance follows exactly this schedule, so ige_run with it on a one-cell
 * state of the same seed reproduces the instance bit for bit. */
int32_t ige_anneal_schedule(double beta_lo, double beta_hi, uint32_t sweeps, double *beta_out);
/* UltimateSolver's temperature ladder (ultimate.rs:237-246):
 * T_t = temp_max * f^t, f = (temp_min / temp_max)^(1 / (num_temps - 1))
 * (f = 1 for one temperature), f^t by binary exponentiation as Rust's powi;
 * beta_out[t] = 1 / T_t, hottest first. */
int32_t ige_ultimate_ladder(double temp_max, double temp_min, uint32_t num_temps,
                            double *beta_out);
int32_t ige_batch_create(ige_ctx *ctx, uint32_t count, const ige_instance_desc *desc,
                         ige_batch **out);
int32_t ige_batch_destroy(ige_batch *batch);
/* Back to the seeded initial spins, exact energies, zero counters. */
int32_t ige_batch_reset(ige_batch *batch);
/* Sweeps [k0, k1) of every instance's schedule (clamped to its length).
 * waves: workgroup size in wavefronts per instance (1, 2, 4, 8, 16), 0 = tuned. */
int32_t ige_batch_run(ige_batch *batch, uint32_t mode, uint32_t waves, uint32_t k0, uint32_t k1);
/* Times the persistent kernel at 1, 2, 4, 8, 16 waves per instance on
 * `probe_sweeps` sweeps, keeps the fastest, then resets the batch. */
int32_t ige_batch_autotune(ige_batch *batch, uint32_t probe_sweeps, uint32_t *waves_out,
                           double *ms_out /* 5 entries, or NULL */);
/* Large instances run on a gang of workgroups (instance sites / gang_sites of
 * them, at most max_gang <= 128) that meet at a device-wide barrier after every
 * colour class.  spin_limit bounds each barrier wait; an expired wait makes
 * the next synchronize/read fail with IGE_ERR_HIP.  Batch defaults: 1 (no gangs:
 * throughput), 2048, 2^24.  The server sizes gangs per round from its load:
 * a PT request gets one workgroup per temperature cell (at most max_gang of
 * them, owning several cells each when fewer), an anneal one; CUs left
 * without any workgroup are shared out by sites to split requests further
 * (an anneal up to min(max_gang, n / gang_sites) members, a PT request with
 * one workgroup per cell up to s <= min(max_gang, n / gang_sites, 128 / cells)
 * per cell); when the base gangs exceed the resident workgroups each request
 * gets a share by sites instead.  Server defaults: 16, 1024, 2^24 (a gang
 * barrier per colour class costs more than a few hundred sites of work). */
int32_t ige_batch_configure(ige_batch *batch, uint32_t max_gang, uint32_t gang_sites,
                            uint32_t spin_limit);
int32_t ige_batch_synchronize(ige_batch *batch);
int32_t ige_batch_info(const ige_batch *batch, uint64_t *total_sites, uint64_t *total_attempts);
int32_t ige_batch_read(ige_batch *batch, uint32_t index, uint64_t *words /* n, site order */,
                       int64_t *energies /* 64 */, uint64_t *accepted);
/* Every instance's 64 tracked energies (instance i at i*64) and, computed
 * afresh on the device from its words and CSR, its exact energies
 * (SpinState::audit's comparison, sparse_bitslice.rs:317-346, for a whole
 * batch at once).  Either pointer may be NULL. */
int32_t ige_batch_energies(ige_batch *batch, int64_t *tracked /* [count][64] */,
                           int64_t *recomputed /* [count][64] */);
/* launch_ms: duration of the last IGE_BATCH_PERSISTENT run, from HIP events
 * recorded on the batch's stream around its launch.  ms: device time of each
 * instance's sweeps in the last persistent run that advanced it (wall clock
 * of its first workgroup, first sweep to last); the batch's critical path is
 * the largest of them.  Either pointer may be NULL.  No reference counterpart
 * (the reference times whole solves, src/bin/host_timing_calibration.rs). */
int32_t ige_batch_timing(ige_batch *batch, double *launch_ms, double *ms /* [count] */);

/* ---- serving: continuous batching of solve requests ---- */
/* Request statuses. */
#define IGE_REQ_QUEUED 0
#define IGE_REQ_RUNNING 1
#define IGE_REQ_DONE 2
#define IGE_REQ_FAILED (-1)

typedef struct ige_server ige_server;

typedef struct {
    uint64_t id;
    int32_t status;        /* IGE_REQ_*                                     */
    uint32_t n;
    int64_t best_energy;   /* lowest of the final replica energies          */
    uint32_t best_replica; /* its slot * 64 + lane                          */
    uint32_t cells;        /* cells of 64 replicas (num_temps, 1 for an anneal) */
    uint64_t accepted;
    double latency_ms;     /* submit -> result ready                        */
    double queue_ms;       /* submit -> admitted to the device              */
    int64_t solution_energy; /* energy of the solution (ige_server_solution) */
    uint32_t solution_sweep; /* sweeps done when the solution's state was seen */
    uint32_t polish_flips;   /* flips of the 1-opt polish                   */
} ige_request_info;

typedef struct {
    uint64_t queued, live, completed, rounds, attempts;
    uint64_t arena_used, arena_capacity, max_live;
    double busy_ms;
} ige_server_counters;

/* Starts the serving loop on a background thread with an HBM arena of
 * hbm_bytes.  Each round advances every live request by as many sweeps as fit
 * in about round_ms; waves = workgroup size per request (0 = 8). */
int32_t ige_server_create(ige_ctx *ctx, uint64_t hbm_bytes, double round_ms, uint32_t waves,
                          ige_server **out);
int32_t ige_server_destroy(ige_server *server);
/* Gang settings as ige_batch_configure (per-round sizing described there);
 * only before the first submit. */
int32_t ige_server_configure(ige_server *server, uint32_t max_gang, uint32_t gang_sites,
                             uint32_t spin_limit);
/* Queues one request (validated like POST /api/v1/solve, server_api.rs:70-99);
 * thread-safe.  The schedule is that of ige_anneal_schedule. */
int32_t ige_server_submit(ige_server *server, const ige_instance_desc *request, uint64_t *id);
/* Status and, once done, the final words of every cell ([cells][n], site
 * order; cells in info) and their energies ([cells][64]).  release != 0
 * forgets a finished request. */
int32_t ige_server_poll(ige_server *server, uint64_t id, int32_t release, ige_request_info *info,
                        uint64_t *words /* cells*n, or NULL */,
                        int64_t *energies /* cells*64, or NULL */);
int32_t ige_server_wait(ige_server *server, uint64_t id, double timeout_ms,
                        ige_request_info *info);
/* The request's solution as UltimateSolver::solve returns it
 * (ultimate.rs:336-374): the lowest-energy final replica (ties to the lowest
 * temperature index, then lane), replaced by the best state seen at the
 * incumbent updates (track_every > 0) when that is strictly lower, then
 * 1-opt polished when the request asked for it.  With num_temps >= 1 the
 * request ran UltimateSolver's loop itself; with num_temps = 0 it ran a
 * one-cell anneal.  x: n bytes in site order. */
int32_t ige_server_solution(ige_server *server, uint64_t id, uint8_t *x, int64_t *energy);
int32_t ige_server_stats(ige_server *server, ige_server_counters *stats);

/* Per-launch timing of the sweep kernels with HIP events on the state's
 * stream (off by default).  ms_total is the summed duration of `launches`
 * sweep-kernel launches since profiling was enabled. */
int32_t ige_profile_enable(ige_state *state, int32_t on);
int32_t ige_profile_read(ige_state *state, double *ms_total, uint64_t *launches,
                         double *ms_other, uint64_t *other_launches);

#ifdef __cplusplus
}
#endif
#endif
