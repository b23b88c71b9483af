// batch.hip -- thousands of independent Ising/QUBO instances of uneven size
// in one persistent launch over a packed CSR arena.
//
// Each instance is one cell of 64 replicas annealed over its own geometric
// beta schedule.  A workgroup takes the next instance from a device work
// queue (instances ordered largest first), runs all of its sweeps with
// workgroup barriers between colour classes, and returns for the next one:
// no grid-wide synchronisation is ever needed.  The per-sweep Metropolis
// thresholds are built in LDS by the workgroup itself.
//
// The uniforms are keyed by (instance seed, position, slot 0, sweep), so an
// instance's trajectory is identical to a single state (ige_state_create +
// ige_run with the schedule of ige_anneal_schedule), whichever workgroup,
// co-resident instance or launch mode (persistent or plain) runs it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ising_gpu.h"
#include "abi_util.hpp"
#include "ige_math.h"
#include "kernels.hpp"
#include "problem_ir.hpp"
#include "batch_internal.hpp"

namespace ige {

// thr[d] = metropolis_threshold(beta, d) for d in [1, dtab], thr[0] = 2^32 - 1.
static __global__ void thr_value_kernel(double beta, uint32_t* thr, uint32_t dtab) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d <= dtab) thr[d] = d == 0 ? 0xFFFFFFFFu : metropolis_threshold(beta, (int64_t)d);
}

}  // namespace ige

using ige::abi::check;
using ige::abi::DevBuf;
using ige::abi::Failure;
using ige::abi::guarded;
using ige::abi::require;

struct ige_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t count = 0;
    std::vector<ige::LoweredInstance> inst;
    std::vector<ige::InstDesc> desc;
    std::vector<uint32_t> queue;  // largest work first
    uint64_t total_sites = 0, total_attempts = 0;
    uint32_t max_dtab = 0;
    uint32_t tuned_waves = 0;
    uint32_t max_gang = 1, gang_sites = ige::kGangSites, spin_limit = 1u << 24;
    std::vector<uint32_t> gang;
    std::vector<uint32_t> tickets;
    int num_cus = 0;
    int wall_khz = 0;  // device wall-clock rate (wall_clock64 ticks per ms)
    DevBuf<ige::InstDesc> d_desc;
    DevBuf<uint32_t> d_queue, d_rp, d_ci, d_cs, d_order, d_progress, d_zero_temp, d_thr,
        d_list, d_bars, d_err;
    DevBuf<const uint32_t*> d_orders;
    DevBuf<int32_t> d_wt, d_lin;
    DevBuf<uint64_t> d_spins, d_ticks;
    DevBuf<int64_t> d_energy, d_recompute;
    DevBuf<unsigned long long> d_accepted;
    DevBuf<unsigned int> d_next;
    // The persistent anneal as one captured graph (progress fill, counter
    // resets, persistent kernel), kept for the (waves, k0, k1) it was
    // captured with; dropped when a new plan moves the tickets.
    hipGraphExec_t graph = nullptr;
    uint32_t graph_key[3] = {0, 0, 0};
    // HIP events on the batch's stream around the last persistent run.
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    void drop_graph() {
        if (!graph) return;
        (void)hipStreamSynchronize(stream);  // its launches have ended
        (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    ~ige_batch() {
        drop_graph();
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

namespace {

void batch_reset(ige_batch* b) {
    hipLaunchKernelGGL(ige::instance_init_kernel, dim3(b->count), dim3(256), 0, b->stream,
                       b->d_desc.p, b->d_list.p, b->d_orders.p);
    check(hipGetLastError(), "batch init");
    hipLaunchKernelGGL(ige::instance_energy_kernel, dim3(b->count), dim3(256), 0, b->stream,
                       b->d_desc.p, b->d_list.p, (int64_t*)nullptr);
    check(hipGetLastError(), "batch energy");
    check(hipStreamSynchronize(b->stream), "batch reset");
}

ige::PersistentLaunch launch_args(ige_batch* b, uint32_t round) {
    ige::PersistentLaunch L;
    L.insts = b->d_desc.p;
    L.queue = b->d_queue.p;
    L.tickets = static_cast<uint32_t>(b->tickets.size());
    L.next = b->d_next.p;
    L.bars = b->d_bars.p;
    L.nbars = b->count;
    L.err = b->d_err.p;
    L.round = round;
    L.spin_limit = b->spin_limit;
    L.num_cus = b->num_cus;
    return L;
}

// Gang sizes, threshold modes, scratch tables, descriptors and tickets.
void plan(ige_batch* b) {
    b->gang.resize(b->count);
    for (uint32_t i = 0; i < b->count; ++i) {
        const auto& L = b->inst[i];
        b->gang[i] = ige::choose_gang(L.n, b->max_gang, b->gang_sites);
    }
    for (uint32_t i = 0; i < b->count; ++i) {
        const auto& L = b->inst[i];
        ige::InstDesc& D = b->desc[i];
        D.gang = b->gang[i];
        D.split = 1;
        D.bar = b->d_bars.p + i;
        D.ticks = b->d_ticks.p + i;
        D.thr_mode = ige::choose_thr_mode(L.dtab);
    }
    // largest work per workgroup first; a gang's tickets stay consecutive
    std::stable_sort(b->queue.begin(), b->queue.end(), [&](uint32_t a, uint32_t c) {
        return b->inst[a].work() / b->gang[a] > b->inst[c].work() / b->gang[c];
    });
    ige::build_tickets(b->queue, b->gang, b->tickets);
    b->d_desc.upload(b->desc.data(), b->count, b->stream);
    b->d_queue.alloc(b->tickets.size());
    b->d_queue.upload(b->tickets.data(), b->tickets.size(), b->stream);
    check(hipStreamSynchronize(b->stream), "plan");
}

void check_error_flag(ige_batch* b) {
    uint32_t e = 0;
    check(hipMemcpyAsync(&e, b->d_err.p, sizeof(e), hipMemcpyDeviceToHost, b->stream),
          "error flag");
    check(hipStreamSynchronize(b->stream), "error flag");
    if (e) {
        const uint32_t z = 0;
        check(hipMemcpyAsync(b->d_err.p, &z, sizeof(z), hipMemcpyHostToDevice, b->stream),
              "error flag");
        check(hipStreamSynchronize(b->stream), "error flag");
        throw Failure(IGE_ERR_HIP, "a gang barrier wait exceeded its spin limit; results invalid");
    }
}

void set_progress(ige_batch* b, uint32_t k0) {
    hipLaunchKernelGGL(ige::fill_u32_kernel, dim3((b->count + 255) / 256), dim3(256), 0,
                       b->stream, b->d_progress.p, k0, b->count);
    check(hipGetLastError(), "progress");
}

// The same sweeps as the persistent kernel, issued the way one would without
// batching: per instance, per sweep, one threshold launch and one sweep
// launch per colour class (the single-state kernels), on one stream.
void run_plain(ige_batch* b, uint32_t k0, uint32_t k1) {
    for (uint32_t id = 0; id < b->count; ++id) {
        const ige::LoweredInstance& L = b->inst[id];
        const ige::InstDesc& D = b->desc[id];
        ige::DevModel m;
        m.n = L.n;
        m.rp = D.rp;
        m.ci = D.ci;
        m.wt = D.wt;
        m.lin = D.lin;
        ige::DevState s;
        s.cells = 1;
        s.spins = D.spins;
        s.energy = D.energy;
        s.temp_of = b->d_zero_temp.p;
        s.accepted = D.accepted;
        for (uint32_t k = k0; k < std::min(k1, L.sweeps); ++k) {
            const double beta = ige::anneal_beta(L.beta_hi, L.c, L.sweeps, k);
            hipLaunchKernelGGL(ige::thr_value_kernel, dim3((L.dtab + 256) / 256), dim3(256), 0,
                               b->stream, beta, b->d_thr.p, L.dtab);
            check(hipGetLastError(), "thr");
            for (uint32_t c = 0; c < L.ncol; ++c)
                check(ige::launch_sweep_generic(m, s, L.cs[c], L.cs[c + 1], k, L.seed, b->d_thr.p,
                                                L.dtab, b->stream),
                      "plain sweep");
        }
    }
}

}  // namespace

extern "C" {

int32_t ige_anneal_schedule(double beta_lo, double beta_hi, uint32_t sweeps, double* out) {
    return guarded([&] {
        require(out != nullptr && sweeps >= 1, "bad arguments");
        ige::check_schedule(beta_lo, beta_hi, "schedule");
        const double c = sweeps > 1 ? (std::log(beta_hi) - std::log(beta_lo)) / (sweeps - 1) : 0.0;
        for (uint32_t k = 0; k < sweeps; ++k) out[k] = ige::anneal_beta(beta_hi, c, sweeps, k);
    });
}

int32_t ige_ultimate_ladder(double temp_max, double temp_min, uint32_t num_temps, double* out) {
    return guarded([&] {
        require(out != nullptr && num_temps >= 1, "bad arguments");
        require(std::isfinite(temp_max) && std::isfinite(temp_min) && temp_max > 0.0 &&
                    temp_min > 0.0,
                "temperatures must be finite and > 0");
        const std::vector<double> b = ige::ultimate_ladder(temp_max, temp_min, num_temps);
        std::copy(b.begin(), b.end(), out);
    });
}

int32_t ige_batch_create(ige_ctx* ctx, uint32_t count, const ige_instance_desc* desc,
                         ige_batch** out) {
    return guarded([&] {
        require(ctx && desc && out && count >= 1, "bad arguments");
        require(count < (1u << (32 - ige::kGangBits)), "too many instances in one batch");
        for (uint32_t i = 0; i < count; ++i)
            if (desc[i].num_temps != 0)
                throw Failure(IGE_ERR_UNSUPPORTED, "a batch runs anneals (num_temps = 0)");
        auto b = std::make_unique<ige_batch>();
        b->device = ige::abi::context_device(ctx);
        check(hipSetDevice(b->device), "hipSetDevice");
        check(hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount,
                                    b->device),
              "device attribute");
        check(hipDeviceGetAttribute(&b->wall_khz, hipDeviceAttributeWallClockRate, b->device),
              "device attribute");
        check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
        check(hipEventCreate(&b->ev0), "event");
        check(hipEventCreate(&b->ev1), "event");
        b->count = count;
        b->inst.reserve(count);
        uint64_t n_tot = 0, e_tot = 0, c_tot = 0;
        for (uint32_t i = 0; i < count; ++i) {
            b->inst.push_back(ige::lower_instance(desc[i], "instance " + std::to_string(i)));
            const auto& L = b->inst.back();
            n_tot += L.n;
            e_tot += L.ci.size();
            c_tot += L.cs.size();
            b->max_dtab = std::max(b->max_dtab, L.dtab);
            b->total_sites += L.n;
            b->total_attempts += static_cast<uint64_t>(L.n) * 64ull * L.sweeps;
        }
        require(e_tot < 0xFFFFFFFFull && n_tot + count < 0xFFFFFFFFull, "batch too large");
        // Packed arena: one allocation per array, instances back to back.
        b->d_rp.alloc(n_tot + count);
        b->d_ci.alloc(std::max<uint64_t>(e_tot, 1));
        b->d_wt.alloc(std::max<uint64_t>(e_tot, 1));
        b->d_lin.alloc(n_tot);
        b->d_cs.alloc(c_tot);
        b->d_order.alloc(n_tot);
        b->d_spins.alloc(n_tot);
        b->d_energy.alloc(static_cast<size_t>(count) * 64);
        b->d_accepted.alloc(count);
        b->d_progress.alloc(count);
        b->d_bars.alloc(count);
        b->d_ticks.alloc(count);
        b->d_err.alloc(1);
        b->d_thr.alloc(static_cast<size_t>(b->max_dtab) + 1);
        std::vector<uint32_t> rp, ci, cs, order;
        std::vector<int32_t> wt, lin;
        std::vector<const uint32_t*> orders(count);
        rp.reserve(n_tot + count);
        ci.reserve(e_tot);
        wt.reserve(e_tot);
        b->desc.resize(count);
        for (uint32_t i = 0; i < count; ++i) {
            const auto& L = b->inst[i];
            ige::InstDesc& D = b->desc[i];
            const uint32_t e_base = static_cast<uint32_t>(ci.size());
            D.rp = b->d_rp.p + rp.size();
            D.ci = b->d_ci.p;
            D.wt = b->d_wt.p;
            D.lin = b->d_lin.p + lin.size();
            D.cs = b->d_cs.p + cs.size();
            D.spins = b->d_spins.p + lin.size();
            D.energy = b->d_energy.p + static_cast<size_t>(i) * 64;
            D.accepted = b->d_accepted.p + i;
            D.progress = b->d_progress.p + i;
            orders[i] = b->d_order.p + lin.size();
            D.n = L.n;
            D.ncol = L.ncol;
            D.sweeps = L.sweeps;
            D.dtab = L.dtab;
            D.key0 = static_cast<uint32_t>(L.seed);
            D.key1 = static_cast<uint32_t>(L.seed >> 32);
            D.round = 0;
            D.beta_hi = L.beta_hi;
            D.c = L.c;
            D.offset = L.offset;
            D.cells = 1;
            D.pt = 0;
            D.ladder = nullptr;
            D.temp_of = nullptr;
            D.slot_of = nullptr;
            D.round_ticks = 0;
            D.stop = nullptr;
            D.swept = nullptr;
            for (uint32_t r : L.rp) rp.push_back(e_base + r);
            ci.insert(ci.end(), L.ci.begin(), L.ci.end());
            wt.insert(wt.end(), L.wt.begin(), L.wt.end());
            lin.insert(lin.end(), L.lin.begin(), L.lin.end());
            cs.insert(cs.end(), L.cs.begin(), L.cs.end());
            order.insert(order.end(), L.order.begin(), L.order.end());
        }
        b->queue.resize(count);
        std::iota(b->queue.begin(), b->queue.end(), 0u);
        std::vector<uint32_t> list(count);
        std::iota(list.begin(), list.end(), 0u);
        b->d_rp.upload(rp.data(), rp.size(), b->stream);
        b->d_ci.upload(ci.data(), ci.size(), b->stream);
        b->d_wt.upload(wt.data(), wt.size(), b->stream);
        b->d_lin.upload(lin.data(), lin.size(), b->stream);
        b->d_cs.upload(cs.data(), cs.size(), b->stream);
        b->d_order.upload(order.data(), order.size(), b->stream);
        b->d_desc.alloc(count);
        b->d_list.alloc(count);
        b->d_list.upload(list.data(), count, b->stream);
        b->d_orders.alloc(count);
        b->d_orders.upload(orders.data(), count, b->stream);
        b->d_next.alloc(1);
        std::vector<uint32_t> zt(64, 0);
        b->d_zero_temp.alloc(64);
        b->d_zero_temp.upload(zt.data(), 64, b->stream);
        const uint32_t zero = 0;
        b->d_err.upload(&zero, 1, b->stream);
        check(hipMemsetAsync(b->d_ticks.p, 0, sizeof(uint64_t) * count, b->stream), "ticks");
        check(hipStreamSynchronize(b->stream), "batch upload");
        plan(b.get());
        batch_reset(b.get());
        *out = b.release();
    });
}

int32_t ige_batch_destroy(ige_batch* b) {
    if (b) (void)hipSetDevice(b->device);
    delete b;
    return IGE_OK;
}

int32_t ige_batch_reset(ige_batch* b) {
    return guarded([&] {
        require(b != nullptr, "NULL batch");
        check(hipSetDevice(b->device), "hipSetDevice");
        batch_reset(b);
    });
}

int32_t ige_batch_run(ige_batch* b, uint32_t mode, uint32_t waves, uint32_t k0, uint32_t k1) {
    return guarded([&] {
        require(b != nullptr, "NULL batch");
        require(k0 <= k1, "need k0 <= k1");
        check(hipSetDevice(b->device), "hipSetDevice");
        if (mode == IGE_BATCH_PLAIN) {
            run_plain(b, k0, k1);
            set_progress(b, k1);
        } else {
            require(mode == IGE_BATCH_PERSISTENT, "unknown batch mode");
            const uint32_t w = waves ? waves : (b->tuned_waves ? b->tuned_waves : 8);
            require(w == 1 || w == 2 || w == 4 || w == 8 || w == 16,
                    "waves per instance must be 1, 2, 4, 8 or 16");
            const uint32_t key[3] = {w, k0, k1};
            check(hipEventRecord(b->ev0, b->stream), "event");
            b->timed = false;
            struct EndEvent {  // records the closing event on every exit path
                ige_batch* b;
                ~EndEvent() { b->timed = hipEventRecord(b->ev1, b->stream) == hipSuccess; }
            } end_event{b};
            if (!ige::abi::graphs_enabled()) {
                set_progress(b, k0);
                ige::launch_persistent<true>(w, launch_args(b, k1 - k0), b->stream);
                return;
            }
            if (!b->graph || !std::equal(key, key + 3, b->graph_key)) {
                b->drop_graph();
                // Relaxed: the launch sizes itself with an occupancy query.  A
                // capture invalidated from outside (another thread synchronising
                // the device meanwhile) launched nothing: launch directly then.
                hipGraph_t g = nullptr;
                bool recorded = hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed) ==
                                hipSuccess;
                if (recorded) {
                    try {
                        set_progress(b, k0);
                        ige::launch_persistent<true>(w, launch_args(b, k1 - k0), b->stream);
                    } catch (const std::exception&) {
                        recorded = false;
                    }
                    recorded = hipStreamEndCapture(b->stream, &g) == hipSuccess && g && recorded;
                }
                hipGraphExec_t exec = nullptr;
                if (recorded && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess)
                    exec = nullptr;
                if (g) (void)hipGraphDestroy(g);
                if (!exec) {
                    (void)hipGetLastError();
                    set_progress(b, k0);
                    ige::launch_persistent<true>(w, launch_args(b, k1 - k0), b->stream);
                    return;
                }
                b->graph = exec;
                std::copy(key, key + 3, b->graph_key);
            }
            check(hipGraphLaunch(b->graph, b->stream), "graph launch");
        }
    });
}

int32_t ige_batch_autotune(ige_batch* b, uint32_t probe_sweeps, uint32_t* waves_out,
                           double* ms_out) {
    return guarded([&] {
        require(b != nullptr && probe_sweeps >= 1, "bad arguments");
        check(hipSetDevice(b->device), "hipSetDevice");
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "event");
        check(hipEventCreate(&e1), "event");
        const uint32_t cand[5] = {1, 2, 4, 8, 16};
        float best = 0.f;
        uint32_t best_w = 8;
        for (int i = 0; i < 5; ++i) {
            set_progress(b, 0);
            ige::launch_persistent<true>(cand[i], launch_args(b, probe_sweeps), b->stream);  // warm
            set_progress(b, 0);
            check(hipEventRecord(e0, b->stream), "event");
            ige::launch_persistent<true>(cand[i], launch_args(b, probe_sweeps), b->stream);
            check(hipEventRecord(e1, b->stream), "event");
            check(hipEventSynchronize(e1), "event");
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, e0, e1), "event");
            if (ms_out) ms_out[i] = ms;
            if (i == 0 || ms < best) {
                best = ms;
                best_w = cand[i];
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->tuned_waves = best_w;
        if (waves_out) *waves_out = best_w;
        batch_reset(b);  // the probes advanced the state; start over
    });
}

int32_t ige_batch_synchronize(ige_batch* b) {
    return guarded([&] {
        require(b != nullptr, "NULL batch");
        check(hipStreamSynchronize(b->stream), "batch synchronize");
        check_error_flag(b);
    });
}

int32_t ige_batch_configure(ige_batch* b, uint32_t max_gang, uint32_t gang_sites,
                            uint32_t spin_limit) {
    return guarded([&] {
        require(b != nullptr, "NULL batch");
        require(max_gang >= 1 && max_gang <= ige::kMaxGang, "max_gang must be in [1, 128]");
        check(hipSetDevice(b->device), "hipSetDevice");
        check(hipStreamSynchronize(b->stream), "configure");
        b->max_gang = max_gang;
        b->gang_sites = gang_sites ? gang_sites : ige::kGangSites;
        b->spin_limit = spin_limit;
        b->drop_graph();
        plan(b);
    });
}

int32_t ige_batch_info(const ige_batch* b, uint64_t* total_sites, uint64_t* total_attempts) {
    return guarded([&] {
        require(b != nullptr, "NULL batch");
        if (total_sites) *total_sites = b->total_sites;
        if (total_attempts) *total_attempts = b->total_attempts;
    });
}

int32_t ige_batch_energies(ige_batch* b, int64_t* tracked, int64_t* recomputed) {
    return guarded([&] {
        require(b != nullptr && (tracked || recomputed), "bad arguments");
        check(hipSetDevice(b->device), "hipSetDevice");
        check(hipStreamSynchronize(b->stream), "batch energies");
        check_error_flag(b);
        const size_t cnt = static_cast<size_t>(b->count) * 64;
        if (tracked) b->d_energy.download(tracked, cnt, b->stream);
        if (recomputed) {
            // Instance i's energies are the 64 entries from i * 64 in both
            // arrays, and d_list is the identity, so out[list index] = out[i].
            if (!b->d_recompute.p) b->d_recompute.alloc(cnt);
            hipLaunchKernelGGL(ige::instance_energy_kernel, dim3(b->count), dim3(256), 0,
                               b->stream, b->d_desc.p, b->d_list.p, b->d_recompute.p);
            check(hipGetLastError(), "batch recompute");
            b->d_recompute.download(recomputed, cnt, b->stream);
        }
        check(hipStreamSynchronize(b->stream), "batch energies");
    });
}

int32_t ige_batch_timing(ige_batch* b, double* launch_ms, double* ms) {
    return guarded([&] {
        require(b != nullptr && (launch_ms || ms), "bad arguments");
        check(hipSetDevice(b->device), "hipSetDevice");
        check(hipStreamSynchronize(b->stream), "batch timing");
        if (launch_ms) {
            require(b->timed, "no persistent run timed yet");
            float f = 0.f;
            check(hipEventElapsedTime(&f, b->ev0, b->ev1), "event");
            *launch_ms = f;
        }
        if (!ms) return;
        std::vector<uint64_t> t(b->count);
        b->d_ticks.download(t.data(), b->count, b->stream);
        check(hipStreamSynchronize(b->stream), "batch timing");
        const double per_ms = b->wall_khz > 0 ? static_cast<double>(b->wall_khz) : 1.0e5;
        for (uint32_t i = 0; i < b->count; ++i) ms[i] = static_cast<double>(t[i]) / per_ms;
    });
}

int32_t ige_batch_read(ige_batch* b, uint32_t index, uint64_t* words, int64_t* energies,
                       uint64_t* accepted) {
    return guarded([&] {
        require(b != nullptr && index < b->count, "bad batch index");
        check(hipSetDevice(b->device), "hipSetDevice");
        const ige::LoweredInstance& L = b->inst[index];
        const ige::InstDesc& D = b->desc[index];
        check(hipStreamSynchronize(b->stream), "batch read");
        check_error_flag(b);
        // Copies on the batch's own stream (never the null stream, which
        // would order against other streams of the process).
        std::vector<uint64_t> w(words ? L.n : 0);
        unsigned long long a = 0;
        if (words)
            check(hipMemcpyAsync(w.data(), D.spins, L.n * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                 b->stream),
                  "read words");
        if (energies)
            check(hipMemcpyAsync(energies, D.energy, 64 * sizeof(int64_t), hipMemcpyDeviceToHost,
                                 b->stream),
                  "read energies");
        if (accepted)
            check(hipMemcpyAsync(&a, D.accepted, sizeof(a), hipMemcpyDeviceToHost, b->stream),
                  "read accepted");
        check(hipStreamSynchronize(b->stream), "batch read");
        if (words)
            for (uint32_t p = 0; p < L.n; ++p) words[L.order[p]] = w[p];
        if (accepted) *accepted = a;
    });
}

}  // extern "C"
