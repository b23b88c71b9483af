// api.cpp -- the C ABI declared in include/ising_gpu.h.
//
// Host side of the engine: owns device buffers, uploads the model in
// colour-ordered position space, and sequences the per-sweep launches
// (threshold refresh, one sweep launch per colour class, DEO exchange) on
// the state's own stream.  No C++ exception crosses the ABI.
#include "../../include/ising_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "abi_util.hpp"
#include "descent_host.hpp"
#include "ige_math.h"
#include "kernels.hpp"
#include "noalloc.hpp"
#include "problem_ir.hpp"

namespace {

using ige::abi::check;
using ige::abi::DevBuf;
using ige::abi::Failure;
using ige::abi::g_last_error;
using ige::abi::guarded;
using ige::abi::require;

}  // namespace

struct ige_model {
    ige_ctx* ctx = nullptr;
    ige::ProblemIR ir;
    std::vector<uint32_t> colors;
    ige::ColorOrder co;
    ige::MultispinShape ms;
    ige::DenseShape ds;
    int64_t max_local = 0;
    uint32_t dmax = 0;
    uint32_t regular_degree = 0;
    // Multispin sweeps run classes [fuse_from, num_colors) in one launch (each
    // at most kFuseMaxClassSites sites, at least two classes); num_colors: none.
    uint32_t fuse_from = 0;
    uint32_t max_class = 0;   // positions in the largest class
    DevBuf<uint32_t> rp, ci, order, pos_of, sgn, cstart;  // cstart: colour class bounds
    DevBuf<int32_t> wt, lin;
    // Dense int8 matrix, built on the first dense state of this model.
    mutable std::mutex dense_mu;
    mutable DevBuf<int8_t> dense_a, dense_hdiag;
    mutable DevBuf<int32_t> dense_lin, dense_wdiag;
    mutable uint32_t n_pad = 0;
    bool dense_ready() const {
        std::lock_guard<std::mutex> lock(dense_mu);
        return n_pad != 0;
    }
    void ensure_dense() const {
        std::lock_guard<std::mutex> lock(dense_mu);
        if (n_pad) return;
        const uint32_t np = (ir.n + 511) / 512 * 512;
        const std::vector<int8_t> a = ige::dense_matrix(ir, co, np, ds.scale);
        std::vector<int32_t> l(np, 0);
        for (uint32_t p = 0; p < ir.n; ++p) l[p] = ir.linear[co.order[p]];
        std::vector<int32_t> wd(static_cast<size_t>(np / 64) * 4096);
        for (uint32_t b = 0; b < np / 64; ++b)
            for (uint32_t p = 0; p < 64; ++p)
                for (uint32_t k = 0; k < 64; ++k)
                    wd[(static_cast<size_t>(b) * 64 + p) * 64 + k] =
                        ds.scale * a[static_cast<size_t>(b * 64 + p) * np + b * 64 + k];
        // diagonal 32 x 32 tiles, row-major: the accept-bound chain's coupling rows
        std::vector<int8_t> hd(static_cast<size_t>(np / 32) * 1024);
        for (uint32_t t = 0; t < np / 32; ++t)
            for (uint32_t p = 0; p < 32; ++p)
                for (uint32_t k = 0; k < 32; ++k)
                    hd[(static_cast<size_t>(t) * 32 + p) * 32 + k] =
                        a[static_cast<size_t>(t * 32 + p) * np + t * 32 + k];
        // matrix-core operand order: tile (g, c) = rows 32g.., columns 32c.., byte
        // 16 l + b = row 32g + l % 32, column 32c + 16 (l / 32) + b
        std::vector<int8_t> at(a.size());
        const uint32_t nt = np / 32;
        for (uint32_t g = 0; g < nt; ++g)
            for (uint32_t c = 0; c < nt; ++c)
                for (uint32_t l = 0; l < 64; ++l)
                    for (uint32_t bb = 0; bb < 16; ++bb)
                        at[(static_cast<size_t>(g) * nt + c) * 1024 + 16 * l + bb] =
                            a[static_cast<size_t>(32 * g + l % 32) * np + 32 * c + 16 * (l / 32) + bb];
        dense_a.alloc(at.size());
        dense_lin.alloc(np);
        dense_wdiag.alloc(wd.size());
        dense_hdiag.alloc(hd.size());
        dense_hdiag.upload(hd.data(), hd.size(), nullptr);
        dense_a.upload(at.data(), at.size(), nullptr);
        dense_lin.upload(l.data(), np, nullptr);
        dense_wdiag.upload(wd.data(), wd.size(), nullptr);
        check(hipStreamSynchronize(nullptr), "dense upload");
        n_pad = np;
    }
    ige::DevDense dense_view() const {
        ige::DevDense d;
        d.n = ir.n;
        d.n_pad = n_pad;
        d.a = dense_a.p;
        d.wdiag = dense_wdiag.p;
        d.hdiag = dense_hdiag.p;
        d.lin = dense_lin.p;
        d.order = order.p;
        d.scale = ds.scale;
        d.dtab = static_cast<uint32_t>(max_local);
        return d;
    }
    ige::DevModel view() const {
        ige::DevModel m;
        m.n = ir.n;
        m.rp = rp.p;
        m.ci = ci.p;
        m.wt = wt.p;
        m.lin = lin.p;
        m.order = order.p;
        m.sgn = sgn.p;
        m.w0 = ms.w0;
        m.dmax = dmax;
        m.regular_degree = regular_degree;
        m.offset = ir.offset;
        return m;
    }
};

struct ige_state {
    ige_ctx* ctx = nullptr;
    const ige_model* model = nullptr;
    uint32_t cells = 0;
    uint64_t seed = 0;
    uint32_t kernel = IGE_KERNEL_GENERIC;
    hipStream_t stream = nullptr;
    uint32_t sweep = 0;   // sweeps run so far: the counter sweep uniforms are keyed on
    uint32_t xround = 0;  // exchanges run so far: keys exchange uniforms, phase = xround & 1
    uint32_t icm_round = 0;  // ICM sweeps run so far: keys the cluster seeds
    uint32_t pops = 1;       // populations (each a ladder of cells / pops levels)
    uint32_t dtab = 0;
    bool ladder_set = false;
    // Replica view (ige_set_replica_temperatures): one temperature per replica
    // index instead of one per slot index.
    bool replica_view = false;
    uint32_t temps() const { return replica_view ? cells * 64u : cells; }
    std::vector<double> ladder;
    DevBuf<uint64_t> spins;
    DevBuf<int64_t> energy, scratch_e;
    DevBuf<uint32_t> temp_of, slot_of, thr;
    DevBuf<unsigned long long> counters;  // accepted, then pair stats
    DevBuf<uint32_t> lev_thr;             // multispin thresholds per temperature and level
    DevBuf<double> beta;                             // current ladder (+ schedule rows)
    DevBuf<int32_t> dense_bounds;                    // accept bounds of a dense sweep
    // Dense field helpers: [0] error flag, then kDenseGangSyncWords per slot;
    // finished blocks' spins; the helpers' partial sums (on first use).
    DevBuf<uint32_t> dense_sync;
    DevBuf<uint64_t> dense_xpub;
    DevBuf<int32_t> dense_part;
    DevBuf<uint64_t> hb;                             // heat-bath thresholds, on first use
    DevBuf<uint64_t> flow_units;  // words with per-site sweep counts (chained runs)
    DevBuf<int32_t> fields;
    DevBuf<ige::BestRecord> best;  // allocated by the first incumbent update
    DevBuf<uint32_t> icm_scratch;  // ICM bitmaps and queues, on first use
    DevBuf<uint8_t> best_x;        // n bytes, site order
    DevBuf<uint32_t> best_take;    // decision of the last update
    // Captured sweep loop: kGraphSteps steps of ige_run replayed as one graph
    // launch, its counters read from `clock`.  Recaptured when `graph_key`
    // (flags, populations, buffers the launches name) changes.
    DevBuf<uint32_t> clock;
    DevBuf<double> sched_buf;  // schedule rows of the running ige_run, grown as needed
    hipGraphExec_t graph = nullptr;
    uint64_t graph_key[5] = {0, 0, 0, 0, 0};
    uint64_t graph_fallbacks = 0;  // captures that failed (those calls ran eagerly)
    uint64_t graph_sweeps = 0;     // sweeps run by graph replays
    uint64_t resident_sweeps = 0;  // sweeps run by resident (LDS) runs
    uint64_t chain_sweeps = 0;     // sweeps run by chained launches
    DevBuf<uint32_t> lev_rows;     // per-sweep level thresholds of a resident run
    DevBuf<uint32_t> chain_sync;   // chained runs: error flag, a barrier counter per slot, the
                                   // XCDs of each slot's gang
    bool profile = false;
    double prof_ms = 0.0, prof_other_ms = 0.0;
    uint64_t prof_launches = 0, prof_other_launches = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    ige::DevState view() {
        ige::DevState s;
        s.cells = cells;
        s.temps = temps();
        s.replica_view = replica_view ? 1u : 0u;
        s.spins = spins.p;
        s.energy = energy.p;
        s.temp_of = temp_of.p;
        s.slot_of = slot_of.p;
        s.accepted = counters.p;
        s.pair_stats = counters.p + 1;
        s.pop_temps = cells / pops;
        return s;
    }
    ~ige_state() {
        if (stream) (void)hipStreamSynchronize(stream);  // no replay still in flight
        if (graph) (void)hipGraphExecDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void set_device(const ige_ctx* ctx) { check(hipSetDevice(ctx->device), "hipSetDevice"); }

void upload_model(ige_model* m) {
    const ige::ProblemIR& ir = m->ir;
    const uint32_t n = ir.n;
    const auto& co = m->co;
    std::vector<uint32_t> rp(static_cast<size_t>(n) + 1, 0), ci(ir.nnz()), sgn(n, 0);
    std::vector<int32_t> wt(ir.nnz()), lin(n);
    uint64_t e_out = 0;
    for (uint32_t p = 0; p < n; ++p) {
        const uint32_t site = co.order[p];
        lin[p] = ir.linear[site];
        uint32_t k = 0;
        for (uint32_t e = ir.row_ptr[site]; e < ir.row_ptr[site + 1]; ++e, ++k) {
            ci[e_out] = co.pos_of[ir.col_idx[e]];
            wt[e_out] = ir.weights[e];
            if (ir.weights[e] < 0 && k < 32) sgn[p] |= 1u << k;
            ++e_out;
        }
        rp[p + 1] = static_cast<uint32_t>(e_out);
    }
    m->rp.alloc(rp.size());
    m->ci.alloc(ci.size());
    m->wt.alloc(wt.size());
    m->lin.alloc(lin.size());
    m->order.alloc(n);
    m->pos_of.alloc(n);
    m->sgn.alloc(n);
    m->rp.upload(rp.data(), rp.size(), nullptr);
    m->ci.upload(ci.data(), ci.size(), nullptr);
    m->wt.upload(wt.data(), wt.size(), nullptr);
    m->lin.upload(lin.data(), lin.size(), nullptr);
    m->order.upload(co.order.data(), n, nullptr);
    m->pos_of.upload(co.pos_of.data(), n, nullptr);
    m->sgn.upload(sgn.data(), n, nullptr);
    m->cstart.alloc(co.class_start.size());
    m->cstart.upload(co.class_start.data(), co.class_start.size(), nullptr);
    check(hipStreamSynchronize(nullptr), "model upload");
}

void finish_model(ige_model* m) {
    if (m->ir.n == 0) throw std::invalid_argument("model needs n >= 1");
    m->max_local = m->ir.max_local_field();
    if (m->max_local > 0x7FFFFFFFll)
        throw ige::RangeError("max local field exceeds int32; fields would not be exact");
    m->dmax = m->ir.max_degree();
    m->colors = m->ir.greedy_coloring();
    m->co = ige::color_order(m->colors);
    {
        const auto& cs = m->co.class_start;
        uint32_t c = m->co.num_colors;
        while (c > 0 && cs[c] - cs[c - 1] <= ige::kFuseMaxClassSites) --c;
        m->fuse_from = m->co.num_colors - c >= 2 ? c : m->co.num_colors;
        m->max_class = 0;
        for (uint32_t k = 0; k < m->co.num_colors; ++k)
            m->max_class = std::max(m->max_class, cs[k + 1] - cs[k]);
    }
    m->ms = ige::multispin_shape(m->ir);
    m->ds = ige::dense_shape(m->ir);
    m->regular_degree = m->dmax;
    for (uint32_t i = 0; i < m->ir.n; ++i)
        if (m->ir.row_ptr[i + 1] - m->ir.row_ptr[i] != m->dmax) m->regular_degree = 0;
    if (m->regular_degree != 4 && m->regular_degree != 6) m->regular_degree = 0;
    upload_model(m);
}

void refresh_thresholds(ige_state* s, const double* dev_beta, const uint32_t* clock = nullptr) {
    const auto m = s->model->view();
    auto st = s->view();
    st.clock = clock;
    if (s->kernel == IGE_KERNEL_GENERIC)
        check(ige::launch_thr_table(dev_beta, s->thr.p, s->temps(), s->dtab, s->stream, clock),
              "thr");
    else if (s->kernel == IGE_KERNEL_MULTISPIN)
        check(ige::launch_level_thresholds(m, st, dev_beta, s->lev_thr.p, s->stream),
              "level thresholds");
    // the dense kernel evaluates its thresholds from beta directly
}

// Switches between the slot-ladder view (one temperature per slot index) and
// the replica view (one per replica index).  The temperature tables, the
// ladder and the pair statistics are resized and the captured sweep loop
// (which names those buffers) is dropped; the accepted count is kept.  Into
// the replica view, the configuration at (temperature t, lane) becomes
// replica t * 64 + lane; back into the slot-ladder view, the permutation
// becomes the identity (a general permutation of replicas cannot be one of
// lane-preserving temperature indices).
void set_view(ige_state* s, bool replica) {
    if (s->replica_view == replica) return;
    check(hipStreamSynchronize(s->stream), "view switch");
    if (s->graph) {
        check(hipGraphExecDestroy(s->graph), "graph destroy");
        s->graph = nullptr;
    }
    unsigned long long accepted = 0;
    s->counters.download(&accepted, 1, s->stream);
    check(hipStreamSynchronize(s->stream), "view switch");
    s->replica_view = replica;
    const size_t T = s->temps();
    s->beta.alloc(T);
    if (s->thr.p) s->thr.alloc(T * (s->dtab + 1));
    if (s->lev_thr.p) s->lev_thr.alloc(T * (s->model->dmax + 1));
    if (s->hb.p) s->hb.alloc(T * (2ull * static_cast<uint64_t>(s->model->max_local) + 1));
    s->counters.alloc(1 + 2 * T);
    std::vector<unsigned long long> zeros(s->counters.count, 0);
    zeros[0] = accepted;
    s->counters.upload(zeros.data(), zeros.size(), s->stream);
    s->ladder.clear();
    s->ladder_set = false;
    if (replica)
        check(ige::launch_replica_index_perm(s->view(), s->stream), "perm");
    else
        check(ige::launch_identity_perm(s->view(), s->stream), "perm");
    check(hipStreamSynchronize(s->stream), "view switch");
}

void ensure_best(ige_state* s) {
    if (s->best.p) return;
    s->best.alloc(1);
    s->best_x.alloc(s->model->ir.n);
    s->best_take.alloc(2);
    ige::BestRecord zero{};
    s->best.upload(&zero, 1, s->stream);
    check(hipStreamSynchronize(s->stream), "best init");
}

void track_best(ige_state* s) {
    check(ige::launch_track_best(s->model->view(), s->view(), s->sweep, s->best.p,
                                 s->best_take.p, s->best_x.p, s->stream),
          "track best");
}

void icm_rounds(ige_state* s, uint32_t rounds, unsigned long long* flipped);

// What one step of ige_run launches; shared by the eager loop and the
// captured graph.
struct StepPlan {
    bool exch = false, track = false, heat = false, icm = false, sched = false;
    bool per_class = false;  // one launch per colour class (no fused tail)
    bool dense_bounds = false;  // dense sweeps decide by accept bounds (sweep_dense_bound_kernel)
    uint32_t hb_dtab = 0;
};

// Steps per replay of the captured sweep loop.
constexpr uint32_t kGraphSteps = 32;
// Polls of one device-wide barrier wait of a chained run before it gives up
// (its workgroups are all resident, so a wait ends long before).
constexpr uint32_t kChainSpinLimit = 1u << 24;
// Polls of one wait in a chained run without barriers (each poll reloads the
// awaited words, about a microsecond with its sleep).
constexpr uint32_t kFlowSpinLimit = 1u << 21;

// Raises the error an expired wait left in flag word *f (and clears it).
void check_wait_flag(ige_state* s, uint32_t* f, const char* what) {
    if (!f) return;
    uint32_t e = 0;
    check(hipMemcpyAsync(&e, f, sizeof(e), hipMemcpyDeviceToHost, s->stream), what);
    check(hipStreamSynchronize(s->stream), what);
    if (e) {
        const uint32_t zero = 0;
        check(hipMemcpyAsync(f, &zero, sizeof(zero), hipMemcpyHostToDevice, s->stream), what);
        check(hipStreamSynchronize(s->stream), what);
        throw Failure(IGE_ERR_HIP, std::string(what) + ": a wait exceeded its spin limit");
    }
}

// Raises the error a chained run's barrier or a dense sweep's field helpers
// left (and clears it).
void check_chain_error(ige_state* s) {
    check_wait_flag(s, s->chain_sync.p, "chained sweep barrier");
    check_wait_flag(s, s->dense_sync.p, "dense sweep field helpers");
}

// One sweep step on s->stream: the thresholds (when `refresh`), the colour
// sweeps, the exchange, an ICM sweep and the incumbent update.  sweep and
// round are absolute, or relative to the step clock when `clock` is set
// (inside a graph capture; no host synchronisation happens then).
void issue_step(ige_state* s, const StepPlan& P, const double* beta, uint32_t sweep,
                uint32_t round, bool refresh, const uint32_t* clock) {
    const auto m = s->model->view();
    auto st = s->view();
    st.clock = clock;
    const auto& cs = s->model->co.class_start;
    if (refresh) {
        if (P.heat)
            check(ige::launch_heatbath_table(beta, s->hb.p, s->temps(), P.hb_dtab, s->stream,
                                             clock),
                  "heat-bath table");
        else
            refresh_thresholds(s, beta, clock);
    }
    const bool prof = s->profile && !clock;
    if (prof) check(hipEventRecord(s->ev0, s->stream), "event");
    const uint32_t launches =
        (s->kernel == IGE_KERNEL_DENSE && !P.heat) ? 1u : s->model->co.num_colors;
    if (P.heat) {
        for (uint32_t c = 0; c < s->model->co.num_colors; ++c)
            check(ige::launch_sweep_heatbath(m, st, cs[c], cs[c + 1], sweep, s->seed, s->hb.p,
                                             P.hb_dtab, s->stream),
                  "sweep_heatbath");
    } else if (s->kernel == IGE_KERNEL_DENSE) {
        require(clock == nullptr, "the dense sweep is not captured");
        const auto dv = s->model->dense_view();
        if (P.dense_bounds && !s->dense_bounds.p)
            s->dense_bounds.alloc(ige::dense_bounds_words(dv, s->cells));
        ige::DevDenseGang gang;
        if (P.dense_bounds && ige::abi::dense_helpers_enabled())
            gang.helpers = ige::dense_gang_helpers(dv, s->cells);
        if (gang.helpers) {
            if (!s->dense_sync.p) {
                const size_t words = 1 + static_cast<size_t>(s->cells) * ige::kDenseGangSyncWords;
                s->dense_sync.alloc(words);
                const std::vector<uint32_t> zero(words, 0u);
                s->dense_sync.upload(zero.data(), zero.size(), s->stream);
                s->dense_xpub.alloc(ige::dense_gang_xpub_words(dv, s->cells));
                s->dense_part.alloc(ige::dense_gang_part_words(s->cells, gang.helpers));
            }
            gang.err = s->dense_sync.p;
            gang.sync = s->dense_sync.p + 1;
            gang.xpub = s->dense_xpub.p;
            gang.part = s->dense_part.p;
            gang.spin_limit = kChainSpinLimit;
        }
        check(ige::launch_sweep_dense(dv, st, sweep, s->seed, beta,
                                      P.dense_bounds ? s->dense_bounds.p : nullptr, &gang,
                                      s->stream),
              "sweep_dense");
    } else if (s->kernel == IGE_KERNEL_GENERIC) {
        for (uint32_t c = 0; c < s->model->co.num_colors; ++c)
            check(ige::launch_sweep_generic(m, st, cs[c], cs[c + 1], sweep, s->seed, s->thr.p,
                                            s->dtab, s->stream),
                  "sweep_generic");
    } else {
        const uint32_t ncol = s->model->co.num_colors;
        const uint32_t fuse = P.per_class ? ncol : s->model->fuse_from;
        for (uint32_t c = 0; c < fuse; ++c)
            check(ige::launch_sweep_multispin(m, st, cs[c], cs[c + 1], sweep, s->seed,
                                              s->lev_thr.p, s->stream),
                  "sweep_multispin");
        check(ige::launch_classes_sweep(m, st, s->model->cstart.p, fuse, ncol, sweep, s->seed,
                                        s->lev_thr.p, s->stream),
              "classes_sweep");
    }
    const auto lap = [&](double& acc, uint64_t& count, uint64_t add) {
        check(hipEventRecord(s->ev1, s->stream), "event");
        check(hipEventSynchronize(s->ev1), "event sync");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, s->ev0, s->ev1), "elapsed");
        acc += ms;
        count += add;
    };
    if (prof) lap(s->prof_ms, s->prof_launches, launches);
    if (P.exch) {
        if (prof) check(hipEventRecord(s->ev0, s->stream), "event");
        check(ige::launch_exchange(st, beta, round, s->seed, s->stream), "exchange");
        if (prof) lap(s->prof_other_ms, s->prof_other_launches, 1);
    }
    if (P.icm) {
        require(clock == nullptr, "cluster moves are not captured");
        icm_rounds(s, 1, nullptr);
    }
    if (P.track)
        check(ige::launch_track_best(m, st, sweep + 1, s->best.p, s->best_take.p, s->best_x.p,
                                     s->stream),
              "track best");
}

// The state's captured sweep loop for plan P: kGraphSteps steps with counters
// relative to the step clock, then a node that advances the clock.  Kernel
// arguments name the state's buffers, so the graph is recaptured when one of
// them (or the plan) changes.
hipGraphExec_t sweep_graph(ige_state* s, const StepPlan& P) {
    if (!s->clock.p) s->clock.alloc(ige::kClockWords);
    const uint64_t key[5] = {
        (P.exch ? 1u : 0u) | (P.track ? 2u : 0u) | (P.heat ? 4u : 0u) | (P.sched ? 8u : 0u) |
            (s->replica_view ? 16u : 0u) | (static_cast<uint64_t>(s->pops) << 32),
        reinterpret_cast<uint64_t>(P.sched ? s->sched_buf.p : nullptr),
        reinterpret_cast<uint64_t>(s->hb.p), reinterpret_cast<uint64_t>(s->best.p), 1};
    if (s->graph && std::equal(key, key + 5, s->graph_key)) return s->graph;
    if (s->graph) {
        check(hipStreamSynchronize(s->stream), "graph retire");  // its replays have ended
        check(hipGraphExecDestroy(s->graph), "graph destroy");
        s->graph = nullptr;
    }
    // A capture can be invalidated from outside (another thread's device-wide
    // synchronisation while it runs); nothing was launched then, so the call
    // falls back to eager launches and the next call captures again.
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    bool recorded = true;
    try {
        const bool refresh = P.sched;
        for (uint32_t k = 0; k < kGraphSteps; ++k) {
            const double* beta =
                P.sched ? s->sched_buf.p + static_cast<size_t>(k) * s->cells : s->beta.p;
            issue_step(s, P, beta, k, k, refresh, s->clock.p);
        }
        check(ige::launch_clock_advance(s->clock.p, kGraphSteps, P.exch ? kGraphSteps : 0,
                                        P.sched ? kGraphSteps : 0, s->stream),
              "clock");
    } catch (const std::exception&) {
        recorded = false;
    }
    const hipError_t end = hipStreamEndCapture(s->stream, &g);
    hipGraphExec_t exec = nullptr;
    const hipError_t inst = (recorded && end == hipSuccess && g)
                                ? hipGraphInstantiate(&exec, g, nullptr, nullptr, 0)
                                : hipErrorStreamCaptureInvalidated;
    if (g) (void)hipGraphDestroy(g);
    if (inst != hipSuccess) {
        (void)hipGetLastError();
        s->graph_fallbacks += 1;
        return nullptr;
    }
    s->graph = exec;
    std::copy(key, key + 5, s->graph_key);
    return exec;
}

using ige::DescentResult;

// Steepest descent of every replica of `st`: dE from the matrix-core fields
// on a dense model, from the CSR rows otherwise.
DescentResult descend(const ige_model* model, const ige::DevState& st, const uint8_t* clamped,
                      uint32_t max_steps, uint32_t* flips, hipStream_t stream) {
    const uint32_t n = model->ir.n;
    DevBuf<int32_t> fields;
    const bool dense = model->ds.eligible && model->ds.dense &&
                       static_cast<uint64_t>(st.cells) * n * 64 * sizeof(int32_t) <= (1ull << 32);
    ige::DevDense dv;
    if (dense) {
        model->ensure_dense();
        dv = model->dense_view();
        fields.alloc(static_cast<size_t>(st.cells) * n * 64);
    }
    return ige::run_descent(model->view(), model->pos_of.p, dense ? &dv : nullptr, fields.p, st,
                            clamped, max_steps, flips, stream);
}

// ICM sweeps (asynchronous).  The scratch, a visited bitmap and a queue per
// move, is kept on the state and capped at 256 MiB; the moves of one sweep run
// in as many launches as that needs (at 3D L=64 x 32, 248 moves of 1 MiB each
// per launch measured faster than all 1024 at once: 70 against 83 ms a sweep).
void icm_rounds(ige_state* s, uint32_t rounds, unsigned long long* flipped) {
    const uint32_t n = s->model->ir.n;
    const uint32_t pairs = s->cells * 32;
    const uint64_t per = ige::icm_scratch_words(n);
    const uint64_t fit = std::max<uint64_t>(1, (256ull << 20) / 4 / per);
    const uint32_t batch = static_cast<uint32_t>(std::min<uint64_t>(pairs, fit));
    if (!s->icm_scratch.p) s->icm_scratch.alloc(static_cast<size_t>(batch) * per);
    for (uint32_t k = 0; k < rounds; ++k) {
        for (uint32_t p0 = 0; p0 < pairs; p0 += batch)
            check(ige::launch_icm(s->model->view(), s->view(), s->model->pos_of.p, p0,
                                  std::min(batch, pairs - p0), s->icm_round, s->seed,
                                  s->icm_scratch.p, flipped, s->stream),
                  "icm");
        s->icm_round += 1;
    }
}

void recompute_energy_into_state(ige_state* s) {
    const auto m = s->model->view();
    auto st = s->view();
    check(ige::launch_energy(m, st, s->energy.p, s->stream), "energy");
}

}  // namespace

extern "C" {

const char* ige_last_error(void) { return g_last_error.c_str(); }

const char* ige_version(void) { return "ising_engine_amd 0.1 gfx950"; }

int32_t ige_context_create(int32_t device, ige_ctx** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        require(device >= 0 && device < count, "device index out of range");
        check(hipSetDevice(device), "hipSetDevice");
        check(ige::noalloc_prepare(device), "no-alloc workspace address");
        auto* c = new ige_ctx;
        c->device = device;
        *out = c;
    });
}

int32_t ige_context_destroy(ige_ctx* ctx) {
    delete ctx;
    return IGE_OK;
}

int32_t ige_model_from_pairs(ige_ctx* ctx, uint32_t n, int64_t offset, const int32_t* linear,
                             uint64_t npairs, const uint32_t* pi, const uint32_t* pj,
                             const int32_t* q, ige_model** out) {
    return guarded([&] {
        require(ctx && out && linear, "NULL argument");
        require(npairs == 0 || (pi && pj && q), "NULL pair arrays");
        set_device(ctx);
        auto m = std::make_unique<ige_model>();
        m->ctx = ctx;
        m->ir = ige::ProblemIR::from_pairs(n, offset, linear, npairs, pi, pj, q);
        finish_model(m.get());
        *out = m.release();
    });
}

int32_t ige_model_destroy(ige_model* model) {
    delete model;
    return IGE_OK;
}

int32_t ige_model_get_info(const ige_model* m, ige_model_info* info) {
    return guarded([&] {
        require(m && info, "NULL argument");
        std::memset(info, 0, sizeof(*info));
        info->n = m->ir.n;
        info->num_colors = m->co.num_colors;
        info->nnz = m->ir.nnz();
        info->max_degree = m->dmax;
        info->multispin_eligible = m->ms.eligible ? 1 : 0;
        info->coupling_magnitude = m->ms.w0;
        info->dense_scale = m->ds.eligible ? m->ds.scale : 0;
        info->max_local_field = m->max_local;
    });
}

int32_t ige_model_coloring(const ige_model* m, uint32_t* colors) {
    return guarded([&] {
        require(m && colors, "NULL argument");
        std::memcpy(colors, m->colors.data(), m->colors.size() * sizeof(uint32_t));
    });
}

int32_t ige_model_linear(const ige_model* m, int32_t* linear, int64_t* offset) {
    return guarded([&] {
        require(m && (linear || offset), "NULL argument");
        if (linear)
            for (uint32_t i = 0; i < m->ir.n; ++i) linear[i] = m->ir.linear[i];
        if (offset) *offset = m->ir.offset;
    });
}

int32_t ige_model_order(const ige_model* m, uint32_t* order, uint32_t* class_start) {
    return guarded([&] {
        require(m && order, "NULL argument");
        std::memcpy(order, m->co.order.data(), m->co.order.size() * sizeof(uint32_t));
        if (class_start)
            std::memcpy(class_start, m->co.class_start.data(),
                        m->co.class_start.size() * sizeof(uint32_t));
    });
}

int32_t ige_model_csr(const ige_model* m, uint32_t* row_ptr, uint32_t* col_idx, int32_t* w) {
    return guarded([&] {
        require(m && row_ptr && (m->ir.nnz() == 0 || (col_idx && w)), "NULL argument");
        std::memcpy(row_ptr, m->ir.row_ptr.data(), m->ir.row_ptr.size() * sizeof(uint32_t));
        std::memcpy(col_idx, m->ir.col_idx.data(), m->ir.nnz() * sizeof(uint32_t));
        std::memcpy(w, m->ir.weights.data(), m->ir.nnz() * sizeof(int32_t));
    });
}

int32_t ige_state_create(ige_ctx* ctx, const ige_model* model, uint32_t cells, uint64_t seed,
                         const uint64_t* init_words, uint32_t kernel, ige_state** out) {
    return guarded([&] {
        require(ctx && model && out, "NULL argument");
        require(cells >= 1 && cells <= 65536, "cells must be in [1, 65536]");
        require(kernel <= IGE_KERNEL_DENSE, "unknown kernel id");
        if (kernel == IGE_KERNEL_MULTISPIN && !model->ms.eligible)
            throw Failure(IGE_ERR_UNSUPPORTED,
                          "multispin kernel needs |q_ij| == w0, zero Ising field, degree <= 31");
        if (kernel == IGE_KERNEL_DENSE && !model->ds.eligible)
            throw Failure(IGE_ERR_UNSUPPORTED,
                          "dense kernel needs q_ij = scale * int8 and n <= 8192");
        set_device(ctx);
        auto s = std::make_unique<ige_state>();
        s->ctx = ctx;
        s->model = model;
        s->cells = cells;
        s->seed = seed;
        s->kernel = kernel;
        if (kernel == IGE_KERNEL_AUTO)
            s->kernel = model->ms.eligible                        ? IGE_KERNEL_MULTISPIN
                        : (model->ds.eligible && model->ds.dense) ? IGE_KERNEL_DENSE
                                                                  : IGE_KERNEL_GENERIC;
        check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
        if (s->kernel == IGE_KERNEL_MULTISPIN) ige::prepare_sweep_multispin(model->view());
        check(hipEventCreate(&s->ev0), "hipEventCreate");
        check(hipEventCreate(&s->ev1), "hipEventCreate");
        const uint32_t n = model->ir.n;
        s->spins.alloc(static_cast<size_t>(cells) * n);
        s->energy.alloc(static_cast<size_t>(cells) * 64);
        s->temp_of.alloc(static_cast<size_t>(cells) * 64);
        s->slot_of.alloc(static_cast<size_t>(cells) * 64);
        s->counters.alloc(1 + 2 * static_cast<size_t>(cells));
        s->beta.alloc(cells);
        if (s->kernel == IGE_KERNEL_GENERIC) {
            s->dtab = static_cast<uint32_t>(model->max_local);
            s->thr.alloc(static_cast<size_t>(cells) * (s->dtab + 1));
        } else if (s->kernel == IGE_KERNEL_DENSE) {
            model->ensure_dense();
        } else {
            s->lev_thr.alloc(static_cast<size_t>(cells) * (model->dmax + 1));
        }
        std::vector<unsigned long long> zeros(s->counters.count, 0);
        s->counters.upload(zeros.data(), zeros.size(), s->stream);
        auto st = s->view();
        check(ige::launch_identity_perm(st, s->stream), "perm");
        if (init_words) {
            std::vector<uint64_t> w(static_cast<size_t>(cells) * n);
            for (uint32_t c = 0; c < cells; ++c)
                for (uint32_t p = 0; p < n; ++p)
                    w[static_cast<size_t>(c) * n + p] =
                        init_words[static_cast<size_t>(c) * n + model->co.order[p]];
            s->spins.upload(w.data(), w.size(), s->stream);
            check(hipStreamSynchronize(s->stream), "init upload");
        } else {
            check(ige::launch_init_spins(model->view(), st, seed, s->stream), "init");
        }
        recompute_energy_into_state(s.get());
        check(hipStreamSynchronize(s->stream), "state create");
        *out = s.release();
    });
}

int32_t ige_state_destroy(ige_state* s) {
    if (s) (void)hipSetDevice(s->ctx->device);
    delete s;
    return IGE_OK;
}

int32_t ige_set_ladder(ige_state* s, const double* beta) {
    return guarded([&] {
        require(s && beta, "NULL argument");
        for (uint32_t t = 0; t < s->cells; ++t)
            require(std::isfinite(beta[t]) && beta[t] >= 0.0, "beta must be finite and >= 0");
        set_device(s->ctx);
        set_view(s, false);
        s->ladder.assign(beta, beta + s->cells);
        s->beta.upload(s->ladder.data(), s->cells, s->stream);
        s->ladder_set = true;
        refresh_thresholds(s, s->beta.p);
        check(hipStreamSynchronize(s->stream), "set ladder");
    });
}

int32_t ige_run(ige_state* s, uint32_t sweeps, uint32_t flags, const double* sched) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        require(sched != nullptr || s->ladder_set, "no ladder set and no beta schedule given");
        require((flags & ~(IGE_RUN_EXCHANGE | IGE_RUN_TRACK_BEST | IGE_RUN_HEATBATH |
                           IGE_RUN_ICM | IGE_RUN_EAGER)) == 0,
                "unknown flag");
        require(static_cast<uint64_t>(s->icm_round) + sweeps <= 0xFFFFFFFFull,
                "ICM counter would exceed 2^32-1");
        require(static_cast<uint64_t>(s->sweep) + sweeps <= 0xFFFFFFFFull &&
                    static_cast<uint64_t>(s->xround) + sweeps <= 0xFFFFFFFFull,
                "sweep or exchange counter would exceed 2^32-1");
        if (s->replica_view)
            require(!(flags & (IGE_RUN_EXCHANGE | IGE_RUN_ICM)) && sched == nullptr,
                    "a state in the replica view runs sweeps on its replica temperatures: no "
                    "schedule, no IGE_RUN_EXCHANGE (use ige_exchange_replicas), no IGE_RUN_ICM");
        set_device(s->ctx);
        StepPlan P;
        P.exch = (flags & IGE_RUN_EXCHANGE) && s->cells >= 2;
        P.track = (flags & IGE_RUN_TRACK_BEST) != 0;
        P.heat = (flags & IGE_RUN_HEATBATH) != 0;
        P.icm = (flags & IGE_RUN_ICM) != 0;
        P.sched = sched != nullptr;
        P.per_class = (flags & IGE_RUN_EAGER) || s->profile;
        P.hb_dtab = static_cast<uint32_t>(s->model->max_local);
        if (P.track) ensure_best(s);
        if (P.heat && !s->hb.p)
            s->hb.alloc(static_cast<size_t>(s->temps()) * (2ull * P.hb_dtab + 1));
        if (sched) {
            const size_t total = static_cast<size_t>(sweeps) * s->cells;
            for (size_t i = 0; i < total; ++i)
                require(std::isfinite(sched[i]) && sched[i] >= 0.0,
                        "beta must be finite and >= 0");
            if (s->sched_buf.count < total) s->sched_buf.alloc(total);
            s->sched_buf.upload(sched, total, s->stream);
        }
        const auto row = [&](uint32_t k) {
            return sched ? s->sched_buf.p + static_cast<size_t>(k) * s->cells : s->beta.p;
        };
        if (s->kernel == IGE_KERNEL_DENSE && !P.heat) {
            // Accept bounds need every beta of the call at or above their minimum.
            double bmin = std::numeric_limits<double>::infinity();
            if (sched) {
                for (size_t i = 0; i < static_cast<size_t>(sweeps) * s->cells; ++i)
                    bmin = std::min(bmin, sched[i]);
            } else {
                for (const double b : s->ladder) bmin = std::min(bmin, b);
            }
            P.dense_bounds = ige::dense_bounds_eligible(s->model->dense_view()) &&
                             bmin >= ige::kDenseBoundMinBeta;
        }
        uint32_t done = 0;
        // Small regular lattices without exchange: the whole call in one
        // resident launch (words in LDS, one workgroup per slot).
        const bool resident = !(flags & IGE_RUN_EAGER) && !s->profile &&
                              s->kernel == IGE_KERNEL_MULTISPIN && !P.exch && !P.track &&
                              !P.heat && !P.icm && sweeps >= 2 &&
                              s->model->ir.n <= ige::kResidentMaxSites &&
                              (s->model->regular_degree == 4 || s->model->regular_degree == 6);
        if (resident) {
            const auto m = s->model->view();
            const auto st = s->view();
            const uint32_t row = s->cells * (s->model->dmax + 1);
            const uint32_t* rows = s->lev_thr.p;
            uint32_t stride = 0;
            if (sched) {
                const size_t need = static_cast<size_t>(sweeps) * row;
                if (s->lev_rows.count < need) s->lev_rows.alloc(need);
                check(ige::launch_level_rows(m, s->cells, s->sched_buf.p, sweeps, s->lev_rows.p,
                                             s->stream),
                      "level rows");
                rows = s->lev_rows.p;
                stride = row;
            }
            check(ige::launch_resident_sweeps(m, st, s->model->cstart.p, s->model->co.num_colors,
                                              rows, stride, s->sweep, sweeps, s->seed, s->stream),
                  "resident sweeps");
            s->sweep += sweeps;
            s->resident_sweeps += sweeps;
            done = sweeps;
        }
        // One-slot multispin states of irregular degree without exchange,
        // incumbent, heat bath or ICM: every class of every sweep in one
        // launch, workgroups meeting at a device-wide barrier per class (small
        // regular lattices run resident above; with more slots the barriers'
        // cache invalidations cost more than the launches they save).
        const bool chain = !resident && !(flags & IGE_RUN_EAGER) && !s->profile && !P.exch &&
                           !P.track && !P.heat && !P.icm && !s->replica_view && sweeps >= 2 &&
                           s->cells == 1 && s->kernel == IGE_KERNEL_MULTISPIN &&
                           s->model->regular_degree == 0 && s->model->dmax < 32 &&
                           s->model->max_class <= ige::kChainMaxSites &&
                           ige::chain_ms_capacity_blocks(s->model->dmax) >= 1;
        if (chain) {
            // [0]: error flag, [1 + slot]: barrier counters, [1 + cells + slot]: XCDs
            if (s->chain_sync.count < 2 * s->cells + 1u) {
                s->chain_sync.alloc(2 * s->cells + 1u);
                const std::vector<uint32_t> zero(2 * s->cells + 1u, 0u);
                s->chain_sync.upload(zero.data(), zero.size(), s->stream);
            }
            const auto m = s->model->view();
            const uint32_t lrow = s->cells * (s->model->dmax + 1);
            const uint32_t* lev = s->lev_thr.p;
            uint32_t stride = 0;
            if (sched) {
                const size_t need = static_cast<size_t>(sweeps) * lrow;
                if (s->lev_rows.count < need) s->lev_rows.alloc(need);
                check(ige::launch_level_rows(m, s->cells, s->sched_buf.p, sweeps, s->lev_rows.p,
                                             s->stream),
                      "level rows");
                lev = s->lev_rows.p;
                stride = lrow;
            }
            if (s->cells == 1 && !ige::abi::chain_barriers_forced() &&
                ige::flow_ms_capacity_blocks(s->model->dmax) >= 1) {
                if (s->flow_units.count < 2u * m.n) s->flow_units.alloc(2u * m.n);
                check(ige::launch_sweep_flow_ms(m, s->view(), s->model->cstart.p,
                                                s->model->co.num_colors, s->model->max_class,
                                                lev, stride, s->sweep, sweeps, s->seed,
                                                s->flow_units.p, s->chain_sync.p,
                                                ige::abi::chain_spin_limit(kFlowSpinLimit),
                                                s->stream),
                      "chained multispin sweeps");
            } else {
                check(ige::launch_sweep_chain_ms(m, s->view(), s->model->cstart.p,
                                                 s->model->co.num_colors, s->model->max_class,
                                                 lev, stride, s->sweep, sweeps, s->seed,
                                                 s->chain_sync.p + 1, s->chain_sync.p,
                                                 ige::abi::chain_spin_limit(kChainSpinLimit),
                                                 s->stream),
                      "chained multispin sweeps");
            }
            s->sweep += sweeps;
            s->chain_sweeps += sweeps;
            done = sweeps;
        }
        // Graph replays next: kGraphSteps steps per launch.
        const bool graphable = ige::abi::graphs_enabled() && !s->profile && !P.icm &&
                               !(flags & IGE_RUN_EAGER) &&
                               (P.heat || s->kernel != IGE_KERNEL_DENSE);
        if (graphable && sweeps - done >= kGraphSteps) {
            const hipGraphExec_t g = sweep_graph(s, P);
            if (g) {
                // Step 0's heat-bath table on a fixed ladder (the graph omits it).
                if (P.heat && !sched)
                    check(ige::launch_heatbath_table(s->beta.p, s->hb.p, s->temps(), P.hb_dtab,
                                                     s->stream),
                          "heat-bath table");
                check(ige::launch_clock_set(s->clock.p, s->sweep, s->xround, 0, s->stream),
                      "clock");
                const uint32_t reps = sweeps / kGraphSteps;
                for (uint32_t r = 0; r < reps; ++r)
                    check(hipGraphLaunch(g, s->stream), "graph launch");
                done = reps * kGraphSteps;
                s->graph_sweeps += done;
                s->sweep += done;
                if (P.exch) s->xround += done;
            }
        }
        for (uint32_t k = done; k < sweeps; ++k) {
            // Threshold tables are indexed by temperature, so only a new beta
            // (a schedule row) needs them rebuilt; the heat-bath table is
            // built on the first step of a call.
            const bool refresh = P.heat ? (P.sched || k == 0) : P.sched;
            issue_step(s, P, row(k), s->sweep, s->xround, refresh, nullptr);
            s->sweep += 1;
            if (P.exch) s->xround += 1;
        }
        if (sched) {
            // Leave the ladder at the schedule's last row.
            if (sweeps > 0) {
                const double* last = sched + static_cast<size_t>(sweeps - 1) * s->cells;
                s->ladder.assign(last, last + s->cells);
                s->beta.upload(s->ladder.data(), s->cells, s->stream);
                s->ladder_set = true;
            }
            // the caller may free `sched` once we return
            check(hipStreamSynchronize(s->stream), "run");
            if (chain || (s->kernel == IGE_KERNEL_DENSE && s->dense_sync.p)) check_chain_error(s);
            if (sweeps > 0) refresh_thresholds(s, s->beta.p);
        }
    });
}

int32_t ige_set_replica_temperatures(ige_state* s, const double* temps, uint32_t len) {
    return guarded([&] {
        require(s && temps, "NULL argument");
        require(len >= 1, "need at least one temperature");
        for (uint32_t k = 0; k < len; ++k)
            require(std::isfinite(temps[k]) && temps[k] > 0.0,
                    "temperatures must be finite and > 0");
        require(s->pops == 1, "the replica view has one population");
        set_device(s->ctx);
        set_view(s, true);
        const uint32_t R = s->temps();
        s->ladder.resize(R);
        for (uint32_t r = 0; r < R; ++r) s->ladder[r] = 1.0 / temps[r % len];
        s->beta.upload(s->ladder.data(), R, s->stream);
        s->ladder_set = true;
        refresh_thresholds(s, s->beta.p);
        check(hipStreamSynchronize(s->stream), "set replica temperatures");
    });
}

int32_t ige_exchange_replicas(ige_state* s, uint32_t rounds) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        require(s->replica_view, "not in the replica view: call ige_set_replica_temperatures");
        require(static_cast<uint64_t>(s->xround) + rounds <= 0xFFFFFFFFull,
                "exchange counter would exceed 2^32-1");
        set_device(s->ctx);
        const auto st = s->view();
        for (uint32_t k = 0; k < rounds; ++k) {
            check(ige::launch_exchange_replicas(st, s->beta.p, k & 1u, s->xround, s->seed,
                                                s->stream),
                  "exchange replicas");
            s->xround += 1;
        }
    });
}

int32_t ige_exchange(ige_state* s, uint32_t rounds) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        require(!s->replica_view, "a state in the replica view exchanges with "
                                  "ige_exchange_replicas");
        require(s->ladder_set, "no ladder set");
        require(static_cast<uint64_t>(s->xround) + rounds <= 0xFFFFFFFFull,
                "exchange counter would exceed 2^32-1");
        if (s->cells < 2 || rounds == 0) return;
        set_device(s->ctx);
        auto st = s->view();
        for (uint32_t k = 0; k < rounds; ++k) {
            check(ige::launch_exchange(st, s->beta.p, s->xround, s->seed, s->stream), "exchange");
            s->xround += 1;
        }
    });
}

int32_t ige_synchronize(ige_state* s) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        set_device(s->ctx);
        check(hipStreamSynchronize(s->stream), "synchronize");
        check_chain_error(s);
    });
}

int32_t ige_read_spins(ige_state* s, uint64_t* words) {
    return guarded([&] {
        require(s && words, "NULL argument");
        set_device(s->ctx);
        const uint32_t n = s->model->ir.n;
        std::vector<uint64_t> w(static_cast<size_t>(s->cells) * n);
        s->spins.download(w.data(), w.size(), s->stream);
        check(hipStreamSynchronize(s->stream), "read spins");
        const auto& order = s->model->co.order;
        for (uint32_t c = 0; c < s->cells; ++c)
            for (uint32_t p = 0; p < n; ++p)
                words[static_cast<size_t>(c) * n + order[p]] = w[static_cast<size_t>(c) * n + p];
    });
}

int32_t ige_write_spins(ige_state* s, const uint64_t* w) {
    return guarded([&] {
        require(s && w, "NULL argument");
        set_device(s->ctx);
        const uint32_t n = s->model->ir.n;
        const auto& order = s->model->co.order;
        std::vector<uint64_t> pw(static_cast<size_t>(s->cells) * n);
        for (uint32_t c = 0; c < s->cells; ++c)
            for (uint32_t p = 0; p < n; ++p)
                pw[static_cast<size_t>(c) * n + p] = w[static_cast<size_t>(c) * n + order[p]];
        s->spins.upload(pw.data(), pw.size(), s->stream);
        recompute_energy_into_state(s);
        check(hipStreamSynchronize(s->stream), "write spins");
    });
}

int32_t ige_read_energies(ige_state* s, int64_t* e) {
    return guarded([&] {
        require(s && e, "NULL argument");
        set_device(s->ctx);
        s->energy.download(e, static_cast<size_t>(s->cells) * 64, s->stream);
        check(hipStreamSynchronize(s->stream), "read energies");
    });
}

int32_t ige_read_temp_of(ige_state* s, uint32_t* t) {
    return guarded([&] {
        require(s && t, "NULL argument");
        set_device(s->ctx);
        s->temp_of.download(t, static_cast<size_t>(s->cells) * 64, s->stream);
        check(hipStreamSynchronize(s->stream), "read temp_of");
    });
}

int32_t ige_read_fields(ige_state* s, int32_t* out) {
    return guarded([&] {
        require(s && out, "NULL argument");
        set_device(s->ctx);
        const size_t total = static_cast<size_t>(s->cells) * s->model->ir.n * 64;
        DevBuf<int32_t> f;
        f.alloc(total);
        if (s->kernel == IGE_KERNEL_DENSE)
            check(ige::launch_dense_fields(s->model->dense_view(), s->view(), f.p, s->stream),
                  "dense fields");
        else
            check(ige::launch_fields(s->model->view(), s->view(), f.p, s->stream), "fields");
        f.download(out, total, s->stream);
        check(hipStreamSynchronize(s->stream), "read fields");
    });
}

int32_t ige_recompute_energies(ige_state* s, int64_t* out) {
    return guarded([&] {
        require(s && out, "NULL argument");
        set_device(s->ctx);
        DevBuf<int64_t> e;
        e.alloc(static_cast<size_t>(s->cells) * 64);
        check(ige::launch_energy(s->model->view(), s->view(), e.p, s->stream), "energy");
        e.download(out, e.count, s->stream);
        check(hipStreamSynchronize(s->stream), "recompute energies");
    });
}

int32_t ige_read_stats(ige_state* s, ige_stats* stats, uint64_t* pair_attempts,
                       uint64_t* pair_accepts) {
    return guarded([&] {
        require(s && stats, "NULL argument");
        set_device(s->ctx);
        std::vector<unsigned long long> c(s->counters.count);
        s->counters.download(c.data(), c.size(), s->stream);
        check(hipStreamSynchronize(s->stream), "read stats");
        stats->sweeps = s->sweep;
        stats->exchanges = s->xround;
        stats->accepted = c[0];
        stats->proposed = static_cast<uint64_t>(s->sweep) * s->model->ir.n * 64ull * s->cells;
        stats->phase = s->xround & 1u;
        stats->kernel = s->kernel;
        stats->graph_sweeps = s->graph_sweeps;
        stats->graph_fallbacks = s->graph_fallbacks;
        stats->resident_sweeps = s->resident_sweeps;
        stats->chain_sweeps = s->chain_sweeps;
        for (uint32_t t = 0; t + 1 < s->temps(); ++t) {
            if (pair_attempts) pair_attempts[t] = c[1 + 2 * t];
            if (pair_accepts) pair_accepts[t] = c[2 + 2 * t];
        }
    });
}

int32_t ige_steepest_descent(ige_state* s, uint32_t max_steps, const uint8_t* clamped,
                             ige_descent_report* report) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        set_device(s->ctx);
        DevBuf<uint32_t> flips;
        flips.alloc(static_cast<size_t>(s->cells) * 64);
        std::vector<uint32_t> zeros(flips.count, 0);
        flips.upload(zeros.data(), zeros.size(), s->stream);
        const DescentResult r = descend(s->model, s->view(), clamped, max_steps, flips.p, s->stream);
        if (report) {
            report->steps = r.steps;
            report->converged = r.converged;
            report->flips = r.flips;
            report->proposed = (static_cast<uint64_t>(r.steps) + r.converged) * s->model->ir.n *
                               64ull * s->cells;
        }
    });
}

int32_t ige_steepest_descent_noalloc(ige_state* s, uint32_t max_steps, const uint8_t* clamped,
                                     ige_descent_report* report) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        set_device(s->ctx);
        const int dev = s->ctx->device;
        uint8_t* base = static_cast<uint8_t*>(ige::noalloc_region(dev));
        require(base != nullptr, "the no-alloc workspace of this device is not set up");
        const ige_model* model = s->model;
        const uint64_t n = model->ir.n, cells = s->cells;
        // dE from the matrix-core fields only when the model's dense matrix
        // already exists (building it allocates); the CSR scan gives the same
        // exact dE and so the same steps.
        const bool dense = model->ds.eligible && model->ds.dense && model->dense_ready() &&
                           cells * n * 64 * sizeof(int32_t) <= (1ull << 32);
        auto up = [](uint64_t b) { return (b + 255) / 256 * 256; };
        const uint64_t o_keys = 0;
        const uint64_t o_ctl = o_keys + up(cells * 64 * sizeof(unsigned long long));
        const uint64_t o_flips = o_ctl + up(ige::kDescentCtlWords * sizeof(unsigned long long));
        const uint64_t o_cl = o_flips + up(cells * 64 * sizeof(uint32_t));
        const uint64_t o_aux = o_cl + (clamped ? up(n) : 0);
        const uint64_t o_fields = o_aux + (clamped ? up(n * sizeof(uint32_t)) : 0);
        const uint64_t need = o_fields + (dense ? up(cells * n * 64 * sizeof(int32_t)) : 0);
        if (need > ige::kNoAllocScratchBytes)
            throw Failure(IGE_ERR_INVALID,
                          "state needs " + std::to_string(need) +
                              " bytes of descent workspace; the no-alloc limit is " +
                              std::to_string(ige::kNoAllocScratchBytes));
        std::lock_guard<std::mutex> lock(ige::noalloc_mutex(dev));
        ige::DescentWorkspace ws;
        ws.keys = reinterpret_cast<unsigned long long*>(base + o_keys);
        ws.ctl = reinterpret_cast<unsigned long long*>(base + o_ctl);
        ws.clamped = clamped ? base + o_cl : nullptr;
        ws.site_aux = clamped ? reinterpret_cast<uint32_t*>(base + o_aux) : nullptr;
        uint32_t* flips = reinterpret_cast<uint32_t*>(base + o_flips);
        check(hipMemsetAsync(flips, 0, cells * 64 * sizeof(uint32_t), s->stream), "clear flips");
        ige::DevDense dv;
        if (dense) dv = model->dense_view();
        const DescentResult r = ige::run_descent_in(
            model->view(), model->pos_of.p, dense ? &dv : nullptr,
            dense ? reinterpret_cast<int32_t*>(base + o_fields) : nullptr, s->view(), clamped,
            max_steps, flips, ws, s->stream);
        if (report) {
            report->steps = r.steps;
            report->converged = r.converged;
            report->flips = r.flips;
            report->proposed = (static_cast<uint64_t>(r.steps) + r.converged) * n * 64ull * cells;
        }
    });
}

int32_t ige_set_populations(ige_state* s, uint32_t pops) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        require(pops >= 1 && s->cells % pops == 0, "pops must divide the number of cells");
        require(pops == 1 || !s->replica_view, "the replica view has one population");
        s->pops = pops;
    });
}

// The acceptance-uniform tuning phase of UltimateSolver's PA-PT solve
// (src/solver/ultimate.rs:599-660, LadderMode::AcceptanceUniform): `iters`
// rounds of `steps` sweep + exchange steps on this (throwaway) state at
// T_t = temp_max (temp_min / temp_max)^alpha_t for every population, each
// followed by ige_pa_adapt_spacings on the per-pair swap acceptance summed over
// the populations.  alpha_0 stays, alpha_{T-1} is set to 1.
int32_t ige_pa_tune_ladder(ige_state* s, double temp_max, double temp_min, uint32_t iters,
                           uint32_t steps, double* alphas) {
    auto call = [](int32_t rc) {
        if (rc != IGE_OK) throw Failure(rc, ige::abi::g_last_error);
    };
    return guarded([&] {
        require(s && alphas, "NULL argument");
        require(!s->replica_view, "ladder tuning works on the slot ladder, not the replica view");
        require(std::isfinite(temp_max) && std::isfinite(temp_min) && temp_max > 0.0 &&
                    temp_min > 0.0,
                "temperatures must be finite and > 0");
        const uint32_t T = s->cells / s->pops;
        if (T < 3 || iters == 0 || steps == 0) return;
        std::vector<double> beta(s->cells), temps(T), spacings(T - 1), rates(T - 1),
            adapted(T - 1);
        std::vector<uint64_t> att0(s->cells), acc0(s->cells), att1(s->cells), acc1(s->cells);
        ige_stats st{};
        for (uint32_t it = 0; it < iters; ++it) {
            for (uint32_t t = 0; t < T; ++t) temps[t] = temp_max * std::pow(temp_min / temp_max, alphas[t]);
            for (uint32_t p = 0; p < s->pops; ++p)
                for (uint32_t t = 0; t < T; ++t) beta[p * T + t] = 1.0 / temps[t];
            call(ige_set_ladder(s, beta.data()));
            call(ige_read_stats(s, &st, att0.data(), acc0.data()));
            call(ige_run(s, steps, IGE_RUN_EXCHANGE, nullptr));
            call(ige_read_stats(s, &st, att1.data(), acc1.data()));
            for (uint32_t t = 0; t + 1 < T; ++t) {
                uint64_t na = 0, nc = 0;
                for (uint32_t p = 0; p < s->pops; ++p) {
                    na += att1[p * T + t] - att0[p * T + t];
                    nc += acc1[p * T + t] - acc0[p * T + t];
                }
                rates[t] = na == 0 ? 0.0 : static_cast<double>(nc) / static_cast<double>(na);
                spacings[t] = alphas[t + 1] - alphas[t];
            }
            call(ige_pa_adapt_spacings(spacings.data(), rates.data(), T - 1, adapted.data()));
            double acc = 0.0;
            for (uint32_t t = 0; t + 1 < T; ++t) {
                acc += adapted[t];
                alphas[t + 1] = acc;
            }
            alphas[T - 1] = 1.0;
        }
    });
}

int32_t ige_resample(ige_state* s, const uint32_t* parents) {
    return guarded([&] {
        require(s && parents, "NULL argument");
        require(!s->replica_view, "resampling works on the slot ladder, not the replica view");
        const uint32_t members = s->pops * 64;
        for (uint32_t m = 0; m < members; ++m)
            require(parents[m] < members, "parent index out of range");
        set_device(s->ctx);
        const size_t words = static_cast<size_t>(s->cells) * s->model->ir.n;
        DevBuf<uint64_t> old_spins;
        DevBuf<int64_t> old_energy;
        DevBuf<uint32_t> par;
        old_spins.alloc(words);
        old_energy.alloc(static_cast<size_t>(s->cells) * 64);
        par.alloc(members);
        check(hipMemcpyAsync(old_spins.p, s->spins.p, words * sizeof(uint64_t),
                             hipMemcpyDeviceToDevice, s->stream),
              "snapshot");
        check(hipMemcpyAsync(old_energy.p, s->energy.p, old_energy.count * sizeof(int64_t),
                             hipMemcpyDeviceToDevice, s->stream),
              "snapshot");
        par.upload(parents, members, s->stream);
        check(ige::launch_resample(s->view(), s->model->ir.n, old_spins.p, old_energy.p, par.p,
                                   s->stream),
              "resample");
        check(hipStreamSynchronize(s->stream), "resample");  // the snapshot is freed here
    });
}

int32_t ige_icm(ige_state* s, uint32_t rounds, uint64_t* flipped) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        require(static_cast<uint64_t>(s->icm_round) + rounds <= 0xFFFFFFFFull,
                "ICM counter would exceed 2^32-1");
        require(!s->replica_view, "cluster moves pair lanes of one temperature: not in the "
                                  "replica view");
        set_device(s->ctx);
        DevBuf<unsigned long long> count;
        if (flipped) {
            count.alloc(1);
            const unsigned long long zero = 0;
            count.upload(&zero, 1, s->stream);
        }
        icm_rounds(s, rounds, flipped ? count.p : nullptr);
        if (flipped) {
            unsigned long long c = 0;
            count.download(&c, 1, s->stream);
            check(hipStreamSynchronize(s->stream), "icm count");
            *flipped = c;
        }
    });
}

int32_t ige_update_best(ige_state* s) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        set_device(s->ctx);
        ensure_best(s);
        track_best(s);
    });
}

int32_t ige_read_best(ige_state* s, uint8_t* x, int64_t* energy, uint32_t* sweep) {
    return guarded([&] {
        require(s && energy, "NULL argument");
        require(s->best.p != nullptr, "no incumbent yet: call ige_update_best or run with "
                                      "IGE_RUN_TRACK_BEST");
        set_device(s->ctx);
        ige::BestRecord b{};
        s->best.download(&b, 1, s->stream);
        if (x) s->best_x.download(x, s->model->ir.n, s->stream);
        check(hipStreamSynchronize(s->stream), "read best");
        require(b.valid != 0, "no incumbent yet");
        *energy = b.energy;
        if (sweep) *sweep = b.sweep;
    });
}

int32_t ige_model_polish(const ige_model* model, uint32_t count, uint8_t* x,
                         const uint8_t* clamped, int64_t* energy, uint32_t* flips) {
    return guarded([&] {
        require(model && x && energy, "NULL argument");
        require(count >= 1 && count <= 65536u * 64u, "count must be in [1, 2^22]");
        set_device(model->ctx);
        const uint32_t n = model->ir.n;
        const uint32_t cells = (count + 63) / 64;
        const auto& order = model->co.order;
        // Lanes past `count` in the last cell repeat the cell's first state: they
        // follow its path and add no descent steps.
        std::vector<uint64_t> w(static_cast<size_t>(cells) * n, 0);
        for (uint32_t i = 0; i < cells * 64u; ++i) {
            const uint32_t src = i < count ? i : (i / 64) * 64;
            const uint8_t* xi = x + static_cast<size_t>(src) * n;
            uint64_t* wc = w.data() + static_cast<size_t>(i / 64) * n;
            const uint64_t bit = 1ull << (i % 64);
            for (uint32_t p = 0; p < n; ++p) {
                require(xi[order[p]] <= 1, "x must hold 0/1 bytes");
                if (xi[order[p]]) wc[p] |= bit;
            }
        }
        hipStream_t stream = nullptr;
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
        } guard{stream};
        DevBuf<uint64_t> spins;
        DevBuf<int64_t> e;
        DevBuf<uint32_t> fl;
        spins.alloc(w.size());
        e.alloc(static_cast<size_t>(cells) * 64);
        fl.alloc(static_cast<size_t>(cells) * 64);
        spins.upload(w.data(), w.size(), stream);
        std::vector<uint32_t> zeros(fl.count, 0);
        fl.upload(zeros.data(), zeros.size(), stream);
        ige::DevState st;
        st.cells = cells;
        st.spins = spins.p;
        st.energy = e.p;
        check(ige::launch_energy(model->view(), st, e.p, stream), "energy");
        (void)descend(model, st, clamped, IGE_STEPS_UNLIMITED, fl.p, stream);
        std::vector<int64_t> eh(e.count);
        std::vector<uint32_t> fh(fl.count);
        spins.download(w.data(), w.size(), stream);
        e.download(eh.data(), eh.size(), stream);
        fl.download(fh.data(), fh.size(), stream);
        check(hipStreamSynchronize(stream), "polish");
        for (uint32_t i = 0; i < count; ++i) {
            uint8_t* xi = x + static_cast<size_t>(i) * n;
            const uint64_t* wc = w.data() + static_cast<size_t>(i / 64) * n;
            for (uint32_t p = 0; p < n; ++p) xi[order[p]] = (wc[p] >> (i % 64)) & 1u;
            energy[i] = eh[i];
            if (flips) flips[i] = fh[i];
        }
    });
}

int32_t ige_profile_enable(ige_state* s, int32_t on) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        s->profile = on != 0;
        s->prof_ms = s->prof_other_ms = 0.0;
        s->prof_launches = s->prof_other_launches = 0;
    });
}

int32_t ige_profile_read(ige_state* s, double* ms_total, uint64_t* launches, double* ms_other,
                         uint64_t* other_launches) {
    return guarded([&] {
        require(s != nullptr, "NULL state");
        if (ms_total) *ms_total = s->prof_ms;
        if (launches) *launches = s->prof_launches;
        if (ms_other) *ms_other = s->prof_other_ms;
        if (other_launches) *other_launches = s->prof_other_launches;
    });
}

}  // extern "C"
