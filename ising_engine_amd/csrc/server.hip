// server.hip -- continuous-batching serving loop for solve requests.
//
// Requests (an integer QUBO, a seed, a sweep budget and an anneal range) are
// lowered to position space on the submitting thread and queued.  A
// background loop works in rounds:
//   1. admit queued requests, first come first served, while their HBM
//      footprint fits the arena (one device allocation made up front);
//   2. run one persistent launch in which every live request advances: a
//      parallel-tempering solve sweeps until round_ms of device time has
//      passed (checked at sweep boundaries), an anneal by its own number of
//      sweeps sized from its measured rate, so that a round takes about the
//      target round time whatever the request's size;
//   3. retire finished requests: copy the final replicas, their energies and
//      the counters back, form the solution (best seen at the incumbent
//      updates, or the best final replica; 1-opt polished on request) and
//      free their arena space.
// Requests that arrive while others run are admitted at the next round, so
// their work overlaps on the device with the requests already there.  The
// uniforms are keyed by (request seed, position, sweep), so a request's
// result is the same alone and in any mix.
//
// Request validation mirrors the reference server's limits
// (src/bin/server_api.rs:15-17, 70-99): 1 <= n <= 50000, sweeps <= 10000,
// pair indices in range.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_internal.hpp"
#include "descent_host.hpp"

using ige::abi::DevBuf;
using ige::abi::guarded;

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint64_t kAlign = 256;
constexpr uint32_t kMaxVars = 50000;     // server_api.rs:15
constexpr uint32_t kMaxSweeps = 10000;   // server_api.rs:16
constexpr uint32_t kServerGangSites = 1024;  // fewest sites a gang member gets

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over byte offsets of the device arena.
class ArenaAllocator {
   public:
    explicit ArenaAllocator(uint64_t cap) : cap_(cap) {
        if (cap) free_[0] = cap;
    }
    bool alloc(uint64_t bytes, uint64_t* off) {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[*off + bytes] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = align_up(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t used() const { return used_; }
    uint64_t capacity() const { return cap_; }

   private:
    uint64_t cap_;
    uint64_t used_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

// Byte layout of one request inside the arena; each region 256-B aligned.
// The model regions (rp .. ladder) come first and are uploaded together.
struct Layout {
    uint64_t rp, ci, wt, lin, cs, order, pos_of, ladder, spins, energy, temp_of, slot_of,
        accepted, progress, stop, best, best_x, total;
    explicit Layout(const ige::LoweredInstance& L) {
        uint64_t o = 0;
        auto take = [&](uint64_t bytes) {
            const uint64_t at = o;
            o += align_up(bytes);
            return at;
        };
        rp = take((L.n + 1ull) * 4);
        ci = take(std::max<uint64_t>(L.ci.size(), 1) * 4);
        wt = take(std::max<uint64_t>(L.wt.size(), 1) * 4);
        lin = take(L.n * 4ull);
        cs = take(L.cs.size() * 4ull);
        order = take(L.n * 4ull);
        pos_of = take(L.n * 4ull);
        ladder = take(L.cells * 8ull);
        spins = take(static_cast<uint64_t>(L.cells) * L.n * 8ull);
        energy = take(L.cells * 64ull * 8);
        temp_of = take(L.cells * 64ull * 4);
        slot_of = take(L.cells * 64ull * 4);
        accepted = take(8);
        progress = take(4);
        stop = take(4);
        best = take(sizeof(ige::BestRecord));
        best_x = take(L.n);
        total = o;
    }
};

struct Request {
    uint64_t id = 0;
    ige::LoweredInstance L;
    Clock::time_point t_submit, t_admit, t_done;
    int32_t status = IGE_REQ_QUEUED;
    uint64_t off = 0, bytes = 0;
    ige::InstDesc D{};
    uint32_t gang = 1;      // workgroups in the current round
    uint32_t gang_cap = 1;  // most workgroups it may get (a PT solve: owning whole cells)
    uint32_t split = 1;     // PT: workgroups per cell in the current round
    uint32_t split_cap = 1; // PT: most workgroups per cell
    uint32_t progress = 0;
    double attempts_per_tick = 0.0;  // measured on the device, 0 until its first round
    std::vector<uint8_t> staging;  // host image of the model regions (kept until uploaded)
    std::vector<uint64_t> words;   // result, site order
    std::vector<int64_t> energies;
    uint64_t accepted = 0;
    std::vector<uint8_t> solution;  // UltimateSolver-style solution, site order
    int64_t solution_energy = 0;
    uint32_t solution_sweep = 0, polish_flips = 0;
};

}  // namespace

struct ige_server {
    int device = 0;
    int num_cus = 0;
    int per_cu = 1;  // persistent workgroups resident per CU (occupancy of `waves`)
    hipStream_t stream = nullptr;
    uint32_t waves = 8;
    double round_ms = 10.0;
    uint32_t max_live = 4096;
    // Gangs per round from the load (plan_gangs); gang_sites bounds how finely a
    // request's positions are split.
    uint32_t max_gang = 16, gang_sites = kServerGangSites, spin_limit = 1u << 24;
    ArenaAllocator arena{0};
    uint8_t* d_arena = nullptr;
    DevBuf<ige::InstDesc> d_desc;
    DevBuf<uint32_t> d_queue, d_list, d_bars, d_err;
    DevBuf<uint64_t> d_ticks;
    std::vector<uint64_t> h_ticks;
    DevBuf<uint32_t> d_swept;  // sweeps each live request did in the last round
    std::vector<uint32_t> h_swept;
    double ticks_per_ms = 1.0e5;  // device wall clock (hipDeviceAttributeWallClockRate)
    DevBuf<const uint32_t*> d_orders;
    DevBuf<unsigned int> d_next;
    // polish scratch: one cell of words, its 64 energies and flip counts
    DevBuf<uint64_t> d_pol_words;
    DevBuf<int64_t> d_pol_energy;
    DevBuf<uint32_t> d_pol_flips;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<Request*> live;
    uint64_t next_id = 1;
    bool stop = false;
    std::string fatal;
    std::thread loop;

    // statistics
    uint64_t rounds = 0, completed = 0, attempts_done = 0, max_live_seen = 0;
    double busy_ms = 0.0;
    double attempts_per_ms_per_block = 2.0e4;  // refined from measured rounds

    ~ige_server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (loop.joinable()) loop.join();
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (d_arena) (void)hipFree(d_arena);
    }
};

namespace {

void upload_request(ige_server* s, Request& r) {
    const ige::LoweredInstance& L = r.L;
    const Layout lay(L);
    r.staging.assign(lay.spins, 0);  // model regions only: rp .. ladder
    auto put = [&](uint64_t at, const void* src, uint64_t bytes) {
        if (bytes) std::memcpy(r.staging.data() + at, src, bytes);
    };
    put(lay.rp, L.rp.data(), L.rp.size() * 4);
    put(lay.ci, L.ci.data(), L.ci.size() * 4);
    put(lay.wt, L.wt.data(), L.wt.size() * 4);
    put(lay.lin, L.lin.data(), L.lin.size() * 4);
    put(lay.cs, L.cs.data(), L.cs.size() * 4);
    put(lay.order, L.order.data(), L.order.size() * 4);
    put(lay.pos_of, L.pos_of.data(), L.pos_of.size() * 4);
    put(lay.ladder, L.ladder.data(), L.ladder.size() * 8);
    uint8_t* base = s->d_arena + r.off;
    check(hipMemcpyAsync(base, r.staging.data(), r.staging.size(), hipMemcpyHostToDevice,
                         s->stream),
          "request upload");
    ige::InstDesc& D = r.D;
    D.rp = reinterpret_cast<const uint32_t*>(base + lay.rp);
    D.ci = reinterpret_cast<const uint32_t*>(base + lay.ci);
    D.wt = reinterpret_cast<const int32_t*>(base + lay.wt);
    D.lin = reinterpret_cast<const int32_t*>(base + lay.lin);
    D.cs = reinterpret_cast<const uint32_t*>(base + lay.cs);
    D.spins = reinterpret_cast<uint64_t*>(base + lay.spins);
    D.energy = reinterpret_cast<int64_t*>(base + lay.energy);
    D.accepted = reinterpret_cast<unsigned long long*>(base + lay.accepted);
    D.progress = reinterpret_cast<uint32_t*>(base + lay.progress);
    D.best = reinterpret_cast<ige::BestRecord*>(base + lay.best);
    D.best_x = base + lay.best_x;
    D.track_every = L.track_every;
    D.thr_mode = ige::choose_thr_mode(L.dtab, L.pt ? L.cells : 1u);
    D.cells = L.cells;
    D.pt = L.pt;
    D.ladder = reinterpret_cast<const double*>(base + lay.ladder);
    D.temp_of = reinterpret_cast<uint32_t*>(base + lay.temp_of);
    D.slot_of = reinterpret_cast<uint32_t*>(base + lay.slot_of);
    D.stop = reinterpret_cast<uint32_t*>(base + lay.stop);
    D.round_ticks = 0;   // set per round
    D.swept = nullptr;   // set per round
    D.n = L.n;
    D.ncol = L.ncol;
    D.sweeps = L.sweeps;
    D.dtab = L.dtab;
    D.key0 = static_cast<uint32_t>(L.seed);
    D.key1 = static_cast<uint32_t>(L.seed >> 32);
    D.round = 1;
    D.gang = r.gang;
    D.split = r.split;
    D.bar = nullptr;    // set per round
    D.ticks = nullptr;  // set per round
    D.beta_hi = L.beta_hi;
    D.c = L.c;
    D.offset = L.offset;
}

// The smallest PT gang whose members own at most as many cells as g members
// would: ceil(C / ceil(C / g)).  Members then own equal counts up to one.
uint32_t even_gang(uint32_t cells, uint32_t g) {
    const uint32_t per = (cells + g - 1) / g;
    return (cells + per - 1) / per;
}

// Workgroups per live request for the next round.  Every request first gets
// its base gang: one workgroup per temperature cell for a PT solve (at most
// gang_cap, members then owning several cells), one for an anneal.
// - The base gangs fit in the resident workgroups (per_cu on each CU): each
//   request gets its base gang, and CUs left without any workgroup are shared
//   out by site-cells (n x cells) among the requests that can use more
//   members: an anneal up to gang_cap members, a PT solve whose members own
//   one cell each cells x s members with s <= split_cap splitting each cell's
//   positions.  Splits thus buy latency only with CUs that would otherwise
//   idle: a split gang pays a gang barrier per colour class, and splits into
//   CUs already holding a workgroup lowered both capacity and latency on the
//   serving workload (measured: 18.4 -> 15.6 req/s with up to two per CU).
// - They do not fit: each request gets its share of the resident workgroups by
//   site-cells, at least one, at most its base gang, a PT gang rounded down to
//   the smallest gang with the same cells per member; the workgroups left over
//   then go, one gang step at a time, to the requests with the most sites per
//   member.  A round's PT items all run to the same deadline, so the round
//   lasts one deadline only while every ticket has a resident workgroup.
// A request's trajectory does not depend on its gang.
void plan_gangs(ige_server* s) {
    const uint64_t cus = static_cast<uint64_t>(std::max(1, s->num_cus));
    const uint64_t slots = cus * static_cast<uint64_t>(std::max(1, s->per_cu));
    uint64_t base_sum = 0;
    double work = 0.0, grow_work = 0.0;
    auto base_of = [](const Request* r) { return r->L.pt ? r->gang_cap : 1u; };
    auto can_grow = [](const Request* r) {
        return r->L.pt ? r->split_cap > 1 && r->gang_cap == r->L.cells : r->gang_cap > 1;
    };
    for (Request* r : s->live) {
        const double w = static_cast<double>(r->L.n) * r->L.cells;
        base_sum += base_of(r);
        work += w;
        if (can_grow(r)) grow_work += w;
    }
    if (base_sum > slots) {
        uint64_t used = 0;
        for (Request* r : s->live) {
            const double w = static_cast<double>(r->L.n) * r->L.cells;
            const double share = std::floor(w * static_cast<double>(slots) / work);
            uint32_t g = std::min(base_of(r), static_cast<uint32_t>(std::max(1.0, share)));
            if (r->L.pt) g = even_gang(r->L.cells, g);
            r->gang = g;
            r->split = 1;
            used += g;
        }
        if (used < slots) {
            // sites per member, largest first
            std::vector<std::pair<double, Request*>> load;
            load.reserve(s->live.size());
            for (Request* r : s->live)
                if (r->gang < base_of(r))
                    load.push_back({static_cast<double>(r->L.n) *
                                        ((r->L.cells + r->gang - 1) / r->gang),
                                    r});
            std::sort(load.begin(), load.end(),
                      [](const auto& a, const auto& b) { return a.first > b.first; });
            uint64_t left = slots - used;
            for (auto& [l, r] : load) {
                (void)l;
                const uint32_t per = (r->L.cells + r->gang - 1) / r->gang;
                const uint32_t next = std::min(base_of(r), even_gang(r->L.cells, (r->L.cells + per - 2) / (per - 1)));
                if (next > r->gang && next - r->gang <= left) {
                    left -= next - r->gang;
                    r->gang = next;
                }
            }
        }
    } else {
        for (Request* r : s->live) {
            const double w = static_cast<double>(r->L.n) * r->L.cells;
            r->split = 1;
            r->gang = base_of(r);
            if (base_sum < cus && can_grow(r) && grow_work > 0.0) {
                const uint32_t extra = static_cast<uint32_t>(
                    std::floor(static_cast<double>(cus - base_sum) * w / grow_work));
                if (!r->L.pt) {
                    r->gang = std::min(r->gang_cap, 1 + extra);
                } else {
                    const uint32_t sp = std::min(r->split_cap, 1 + extra / r->L.cells);
                    if (sp > 1) {
                        r->split = sp;
                        r->gang = r->L.cells * sp;
                    }
                }
            }
        }
    }
    for (Request* r : s->live) {
        r->D.gang = r->gang;
        r->D.split = r->split;
    }
}

// 1-opt polish of one state (bytes, position order) on the request's own CSR
// in the arena.  Every lane of a scratch cell holds the state, so all lanes
// follow the same deterministic path and the run ends when lane 0's does.
void polish_state(ige_server* s, Request& r, std::vector<uint8_t>& x, int64_t& energy) {
    const ige::LoweredInstance& L = r.L;
    const Layout lay(L);
    uint8_t* base = s->d_arena + r.off;
    if (s->d_pol_words.count < L.n) s->d_pol_words.alloc(std::max<size_t>(L.n, 1024));
    if (!s->d_pol_energy.p) {
        s->d_pol_energy.alloc(64);
        s->d_pol_flips.alloc(64);
    }
    std::vector<uint64_t> w(L.n);
    for (uint32_t p = 0; p < L.n; ++p) w[p] = x[p] ? ~0ull : 0ull;
    s->d_pol_words.upload(w.data(), L.n, s->stream);
    std::vector<uint32_t> zeros(64, 0);
    s->d_pol_flips.upload(zeros.data(), 64, s->stream);
    ige::DevModel m;
    m.n = L.n;
    m.rp = r.D.rp;
    m.ci = r.D.ci;
    m.wt = r.D.wt;
    m.lin = r.D.lin;
    m.order = reinterpret_cast<const uint32_t*>(base + lay.order);
    m.offset = L.offset;
    ige::DevState st;
    st.cells = 1;
    st.spins = s->d_pol_words.p;
    st.energy = s->d_pol_energy.p;
    check(ige::launch_energy(m, st, st.energy, s->stream), "polish energy");
    (void)ige::run_descent(m, reinterpret_cast<const uint32_t*>(base + lay.pos_of), nullptr,
                           nullptr, st, nullptr, IGE_STEPS_UNLIMITED, s->d_pol_flips.p,
                           s->stream);
    int64_t e0 = 0;
    uint32_t f0 = 0;
    s->d_pol_words.download(w.data(), L.n, s->stream);
    check(hipMemcpyAsync(&e0, st.energy, 8, hipMemcpyDeviceToHost, s->stream), "polish energy");
    check(hipMemcpyAsync(&f0, s->d_pol_flips.p, 4, hipMemcpyDeviceToHost, s->stream), "flips");
    check(hipStreamSynchronize(s->stream), "polish");
    for (uint32_t p = 0; p < L.n; ++p) x[p] = static_cast<uint8_t>(w[p] & 1u);
    energy = e0;
    r.polish_flips = f0;
}

void retire(ige_server* s, Request& r) {
    const ige::LoweredInstance& L = r.L;
    const uint32_t C = L.cells;
    std::vector<uint64_t> w(static_cast<size_t>(C) * L.n);
    std::vector<uint32_t> tof(C * 64, 0);
    r.energies.resize(C * 64);
    check(hipMemcpyAsync(w.data(), r.D.spins, w.size() * 8, hipMemcpyDeviceToHost, s->stream),
          "result words");
    check(hipMemcpyAsync(r.energies.data(), r.D.energy, C * 64 * 8, hipMemcpyDeviceToHost,
                         s->stream),
          "result energies");
    if (L.pt)
        check(hipMemcpyAsync(tof.data(), r.D.temp_of, C * 64 * 4, hipMemcpyDeviceToHost,
                             s->stream),
              "result temperatures");
    unsigned long long acc = 0;
    check(hipMemcpyAsync(&acc, r.D.accepted, 8, hipMemcpyDeviceToHost, s->stream), "accepted");
    ige::BestRecord best{};
    std::vector<uint8_t> x(L.n);
    if (L.track_every) {
        check(hipMemcpyAsync(&best, r.D.best, sizeof(best), hipMemcpyDeviceToHost, s->stream),
              "best record");
        check(hipMemcpyAsync(x.data(), r.D.best_x, L.n, hipMemcpyDeviceToHost, s->stream),
              "best state");
    }
    check(hipStreamSynchronize(s->stream), "retire");
    r.words.assign(w.size(), 0);
    for (uint32_t c = 0; c < C; ++c)
        for (uint32_t p = 0; p < L.n; ++p)
            r.words[static_cast<size_t>(c) * L.n + L.order[p]] = w[static_cast<size_t>(c) * L.n + p];
    r.accepted = acc;
    // The solution (ultimate.rs:339-373): the best final replica, first in
    // (temperature index, lane) order on an energy tie, replaced by the
    // incumbent when that is strictly lower.
    uint32_t q = 0;
    uint64_t key_q = ~0ull;
    for (uint32_t i = 0; i < C * 64; ++i) {
        const uint64_t key = static_cast<uint64_t>(L.pt ? tof[i] : 0u) * 64 + i % 64;
        if (key_q == ~0ull || r.energies[i] < r.energies[q] ||
            (r.energies[i] == r.energies[q] && key < key_q)) {
            q = i;
            key_q = key;
        }
    }
    int64_t e = r.energies[q];
    if (L.track_every) {
        if (!best.valid) throw Failure(IGE_ERR_HIP, "incumbent record missing after the solve");
    }
    if (L.track_every && best.energy < e) {
        e = best.energy;
        r.solution_sweep = best.sweep;
    } else {
        const uint64_t* wc = w.data() + static_cast<size_t>(q / 64) * L.n;
        for (uint32_t p = 0; p < L.n; ++p) x[p] = static_cast<uint8_t>((wc[p] >> (q % 64)) & 1u);
        r.solution_sweep = L.sweeps;
    }
    if (L.polish) polish_state(s, r, x, e);
    r.solution.assign(L.n, 0);
    for (uint32_t p = 0; p < L.n; ++p) r.solution[L.order[p]] = x[p];
    r.solution_energy = e;
    r.staging.clear();
    r.staging.shrink_to_fit();
}

void serve_loop(ige_server* s) {
    (void)hipSetDevice(s->device);
    std::vector<ige::InstDesc> descs;
    std::vector<uint32_t> queue, list;
    std::vector<const uint32_t*> orders;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    try {
        for (;;) {
            std::vector<Request*> admitted;
            {
                std::unique_lock<std::mutex> lk(s->mu);
                s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty() || !s->live.empty(); });
                if (s->stop) break;
                // 1. admission: FIFO while the head fits (no overtaking, no starvation)
                while (!s->pending.empty() && s->live.size() < s->max_live) {
                    Request& r = *s->reqs[s->pending.front()];
                    const Layout lay(r.L);
                    uint64_t off = 0;
                    if (!s->arena.alloc(lay.total, &off)) break;
                    r.off = off;
                    r.bytes = lay.total;
                    r.status = IGE_REQ_RUNNING;
                    r.t_admit = Clock::now();
                    s->pending.pop_front();
                    s->live.push_back(&r);
                    admitted.push_back(&r);
                }
                s->max_live_seen = std::max<uint64_t>(s->max_live_seen, s->live.size());
            }
            for (Request* r : admitted) upload_request(s, *r);
            // 2. one round over every live request
            plan_gangs(s);
            const size_t L = s->live.size();
            descs.resize(L);
            queue.resize(L);
            uint64_t round_attempts = 0;
            // Sweeps per request: about round_ms of its own measured rate (its
            // member 0's device time), or of the per-workgroup estimate before
            // its first round, so that no request stretches the round.
            const double per_block = std::max(1.0, s->attempts_per_ms_per_block) * s->round_ms;
            // A PT solve instead sweeps until round_ms of device time has passed
            // (at sweep boundaries), up to its remaining sweeps: every workgroup
            // stays busy until the round's end.
            for (size_t i = 0; i < L; ++i) {
                Request& r = *s->live[i];
                const uint64_t per_sweep = static_cast<uint64_t>(r.L.n) * 64 * r.L.cells;
                uint32_t k = r.L.sweeps - r.progress;
                if (r.L.pt) {
                    r.D.round_ticks =
                        std::max<uint64_t>(1, static_cast<uint64_t>(s->round_ms * s->ticks_per_ms));
                } else {
                    const double budget = r.attempts_per_tick > 0.0
                                              ? r.attempts_per_tick * s->ticks_per_ms * s->round_ms
                                              : per_block * r.gang;
                    k = std::min<uint32_t>(k, static_cast<uint32_t>(
                                                  std::max(1.0, budget / per_sweep)));
                }
                r.D.round = k;
                descs[i] = r.D;
                if (!r.L.pt) round_attempts += per_sweep * k;
            }
            std::vector<std::pair<uint64_t, uint32_t>> order_by_work(L);
            std::vector<uint32_t> gangs(L), ord(L);
            for (size_t i = 0; i < L; ++i) {
                gangs[i] = s->live[i]->gang;
                order_by_work[i] = {static_cast<uint64_t>(s->live[i]->L.n) * s->live[i]->L.cells *
                                        descs[i].round / gangs[i],
                                    static_cast<uint32_t>(i)};
            }
            std::sort(order_by_work.begin(), order_by_work.end(),
                      [](const auto& a, const auto& b) { return a.first > b.first; });
            for (size_t i = 0; i < L; ++i) ord[i] = order_by_work[i].second;
            ige::build_tickets(ord, gangs, queue);
            if (s->d_desc.count < L) {
                s->d_desc.alloc(std::max<size_t>(L, 2 * s->d_desc.count));
                s->d_list.alloc(s->d_desc.count);
                s->d_orders.alloc(s->d_desc.count);
                s->d_bars.alloc(s->d_desc.count);
                s->d_ticks.alloc(s->d_desc.count);
                s->d_swept.alloc(s->d_desc.count);
            }
            if (s->d_queue.count < queue.size()) s->d_queue.alloc(std::max<size_t>(queue.size(), 2 * s->d_queue.count));
            for (size_t i = 0; i < L; ++i) {
                descs[i].bar = s->d_bars.p + i;
                descs[i].ticks = s->d_ticks.p + i;
                descs[i].swept = s->d_swept.p + i;
            }
            s->d_desc.upload(descs.data(), L, s->stream);
            s->d_queue.upload(queue.data(), queue.size(), s->stream);
            if (!admitted.empty()) {
                list.clear();
                orders.assign(L, nullptr);
                for (size_t i = 0; i < L; ++i)
                    if (s->live[i]->status == IGE_REQ_RUNNING && s->live[i]->progress == 0 &&
                        std::find(admitted.begin(), admitted.end(), s->live[i]) != admitted.end()) {
                        list.push_back(static_cast<uint32_t>(i));
                        const Layout lay(s->live[i]->L);
                        orders[i] = reinterpret_cast<const uint32_t*>(s->d_arena + s->live[i]->off +
                                                                      lay.order);
                    }
                s->d_list.upload(list.data(), list.size(), s->stream);
                s->d_orders.upload(orders.data(), L, s->stream);
                hipLaunchKernelGGL(ige::instance_init_kernel, dim3(list.size()), dim3(256), 0,
                                   s->stream, s->d_desc.p, s->d_list.p, s->d_orders.p);
                hipLaunchKernelGGL(ige::instance_energy_kernel, dim3(list.size()), dim3(256), 0,
                                   s->stream, s->d_desc.p, s->d_list.p, (int64_t*)nullptr);
                check(hipGetLastError(), "request init");
            }
            check(hipEventRecord(e0, s->stream), "event");
            ige::PersistentLaunch pl;
            pl.insts = s->d_desc.p;
            pl.queue = s->d_queue.p;
            pl.tickets = static_cast<uint32_t>(queue.size());
            pl.next = s->d_next.p;
            pl.bars = s->d_bars.p;
            pl.nbars = static_cast<uint32_t>(L);
            pl.err = s->d_err.p;
            pl.round = 1;
            pl.spin_limit = s->spin_limit;
            pl.num_cus = s->num_cus;
            ige::launch_persistent<false>(s->waves, pl, s->stream);
            check(hipEventRecord(e1, s->stream), "event");
            check(hipEventSynchronize(e1), "round");
            {
                uint32_t e = 0;
                check(hipMemcpyAsync(&e, s->d_err.p, sizeof(e), hipMemcpyDeviceToHost, s->stream),
                      "error flag");
                check(hipStreamSynchronize(s->stream), "error flag");
                if (e) throw Failure(IGE_ERR_HIP, "a gang barrier wait exceeded its spin limit");
            }
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, e0, e1), "event");
            // per-request device time of this round -> its rate for the next one
            s->h_ticks.resize(L);
            s->h_swept.resize(L);
            check(hipMemcpyAsync(s->h_ticks.data(), s->d_ticks.p, L * sizeof(uint64_t),
                                 hipMemcpyDeviceToHost, s->stream),
                  "round ticks");
            check(hipMemcpyAsync(s->h_swept.data(), s->d_swept.p, L * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s->stream),
                  "round sweeps");
            check(hipStreamSynchronize(s->stream), "round ticks");
            for (size_t i = 0; i < L; ++i) {
                Request& r = *s->live[i];
                if (r.L.pt) {
                    r.D.round = s->h_swept[i];  // what the deadline left it
                    round_attempts += static_cast<uint64_t>(r.L.n) * 64 * r.L.cells * r.D.round;
                }
                const double done = static_cast<double>(r.L.n) * 64.0 * r.L.cells * r.D.round;
                if (s->h_ticks[i] > 0 && r.D.round > 0) {
                    const double rate = done / static_cast<double>(s->h_ticks[i]);
                    r.attempts_per_tick = r.attempts_per_tick > 0.0
                                              ? 0.5 * r.attempts_per_tick + 0.5 * rate
                                              : rate;
                }
            }
            // refine the per-workgroup rate used to size the next round
            {
                const double blocks = std::max(
                    1.0, std::min<double>(queue.size(), std::max(1, s->per_cu) * s->num_cus));
                if (ms > 0.05)
                    s->attempts_per_ms_per_block =
                        0.5 * s->attempts_per_ms_per_block + 0.5 * (round_attempts / ms / blocks);
            }
            // 3. retire
            std::vector<Request*> done;
            for (Request* r : s->live) {
                r->progress += r->D.round;
                if (r->progress >= r->L.sweeps) done.push_back(r);
            }
            for (Request* r : done) retire(s, *r);
            {
                std::lock_guard<std::mutex> g(s->mu);
                const auto now = Clock::now();
                for (Request* r : done) {
                    r->status = IGE_REQ_DONE;
                    r->t_done = now;
                    s->arena.release(r->off, r->bytes);
                    s->completed++;
                }
                s->live.erase(std::remove_if(s->live.begin(), s->live.end(),
                                             [](Request* r) { return r->status == IGE_REQ_DONE; }),
                              s->live.end());
                s->rounds++;
                s->busy_ms += ms;
                s->attempts_done += round_attempts;
            }
            if (!done.empty()) s->cv_done.notify_all();
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> g(s->mu);
        s->fatal = e.what();
        for (Request* r : s->live) r->status = IGE_REQ_FAILED;
        for (uint64_t id : s->pending) s->reqs[id]->status = IGE_REQ_FAILED;
        s->live.clear();
        s->pending.clear();
        s->cv_done.notify_all();
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
}

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

void fill_info(const Request& r, ige_request_info* info) {
    std::memset(info, 0, sizeof(*info));
    info->id = r.id;
    info->status = r.status;
    info->n = r.L.n;
    info->cells = r.L.cells;
    if (r.status == IGE_REQ_DONE) {
        uint32_t best = 0;
        for (uint32_t i = 1; i < r.energies.size(); ++i)
            if (r.energies[i] < r.energies[best]) best = i;
        info->best_replica = best;
        info->best_energy = r.energies[best];
        info->accepted = r.accepted;
        info->latency_ms = ms_between(r.t_submit, r.t_done);
        info->queue_ms = ms_between(r.t_submit, r.t_admit);
        info->solution_energy = r.solution_energy;
        info->solution_sweep = r.solution_sweep;
        info->polish_flips = r.polish_flips;
    }
}

}  // namespace

extern "C" {

int32_t ige_server_create(ige_ctx* ctx, uint64_t hbm_bytes, double round_ms, uint32_t waves,
                          ige_server** out) {
    return guarded([&] {
        require(ctx && out, "NULL argument");
        require(hbm_bytes >= (1u << 20), "arena must be at least 1 MiB");
        require(round_ms > 0.0, "round_ms must be > 0");
        require(waves == 0 || waves == 1 || waves == 2 || waves == 4 || waves == 8 || waves == 16,
                "waves must be 0, 1, 2, 4, 8 or 16");
        auto s = std::make_unique<ige_server>();
        s->device = ige::abi::context_device(ctx);
        check(hipSetDevice(s->device), "hipSetDevice");
        check(hipDeviceGetAttribute(&s->num_cus, hipDeviceAttributeMultiprocessorCount,
                                    s->device),
              "device attribute");
        check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "stream");
        check(hipMalloc(&s->d_arena, hbm_bytes), "arena");
        s->arena = ArenaAllocator(hbm_bytes);
        s->waves = waves ? waves : 8;
        s->round_ms = round_ms;
        s->per_cu = ige::persistent_per_cu<false>(s->waves);
        s->d_next.alloc(1);
        s->d_desc.alloc(256);
        s->d_queue.alloc(256);
        s->d_list.alloc(256);
        s->d_orders.alloc(256);
        s->d_bars.alloc(256);
        s->d_ticks.alloc(256);
        s->d_swept.alloc(256);
        {
            int khz = 0;
            if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device) ==
                    hipSuccess &&
                khz > 0)
                s->ticks_per_ms = khz;
        }
        s->d_err.alloc(1);
        const uint32_t zero = 0;
        s->d_err.upload(&zero, 1, s->stream);
        check(hipStreamSynchronize(s->stream), "server init");
        ige_server* raw = s.get();
        s->loop = std::thread(serve_loop, raw);
        *out = s.release();
    });
}

int32_t ige_server_configure(ige_server* s, uint32_t max_gang, uint32_t gang_sites,
                             uint32_t spin_limit) {
    return guarded([&] {
        require(s != nullptr, "NULL server");
        require(max_gang >= 1 && max_gang <= ige::kMaxGang, "max_gang must be in [1, 128]");
        std::lock_guard<std::mutex> g(s->mu);
        require(s->reqs.empty(), "configure before submitting requests");
        s->max_gang = max_gang;
        s->gang_sites = gang_sites ? gang_sites : kServerGangSites;
        s->spin_limit = spin_limit;
    });
}

int32_t ige_server_destroy(ige_server* s) {
    delete s;
    return IGE_OK;
}

int32_t ige_server_submit(ige_server* s, const ige_instance_desc* d, uint64_t* id) {
    return guarded([&] {
        require(s && d && id, "NULL argument");
        require(d->n >= 1 && d->n <= kMaxVars, "num_vars must be in [1, 50000]");
        if (d->num_temps == 0) {
            require(d->sweeps >= 1 && d->sweeps <= kMaxSweeps, "sweeps must be in [1, 10000]");
        } else {
            // server_api.rs:16-17: sweeps (per exchange) and exchanges up to 10000 each
            require(d->track_every >= 1 && d->track_every <= kMaxSweeps,
                    "sweeps per exchange must be in [1, 10000]");
            require(d->sweeps % d->track_every == 0 && d->sweeps / d->track_every <= kMaxSweeps,
                    "exchanges must be in [1, 10000]");
        }
        auto r = std::make_unique<Request>();
        r->t_submit = Clock::now();
        r->L = ige::lower_instance(*d, "request");
        // A PT solve's gang members own whole temperature cells (at most one
        // each), or split every cell's positions among up to max_gang and
        // n / gang_sites of them, within kMaxGang in all.
        if (r->L.pt) {
            r->gang_cap = std::min(r->L.cells, s->max_gang);
            r->split_cap = std::max(1u, std::min(ige::choose_gang(r->L.n, s->max_gang, s->gang_sites),
                                                 ige::kMaxGang / r->L.cells));
        } else {
            r->gang_cap = ige::choose_gang(r->L.n, s->max_gang, s->gang_sites);
        }
        r->gang = 1;
        const Layout lay(r->L);
        std::lock_guard<std::mutex> g(s->mu);
        if (!s->fatal.empty()) throw Failure(IGE_ERR_HIP, "server stopped: " + s->fatal);
        if (lay.total > s->arena.capacity())
            throw Failure(IGE_ERR_NOMEM, "request footprint exceeds the server's HBM arena");
        r->id = s->next_id++;
        *id = r->id;
        s->pending.push_back(r->id);
        s->reqs[r->id] = std::move(r);
        s->cv_work.notify_one();
    });
}

int32_t ige_server_poll(ige_server* s, uint64_t id, int32_t release, ige_request_info* info,
                        uint64_t* words, int64_t* energies) {
    return guarded([&] {
        require(s && info, "NULL argument");
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->reqs.find(id);
        require(it != s->reqs.end(), "unknown request id");
        const Request& r = *it->second;
        fill_info(r, info);
        if (r.status == IGE_REQ_DONE) {
            if (words) std::memcpy(words, r.words.data(), r.words.size() * 8);
            if (energies) std::memcpy(energies, r.energies.data(), r.energies.size() * 8);
            if (release) s->reqs.erase(it);
        }
    });
}

int32_t ige_server_solution(ige_server* s, uint64_t id, uint8_t* x, int64_t* energy) {
    return guarded([&] {
        require(s && energy, "NULL argument");
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->reqs.find(id);
        require(it != s->reqs.end(), "unknown request id");
        const Request& r = *it->second;
        require(r.status == IGE_REQ_DONE, "request not done");
        if (x) std::memcpy(x, r.solution.data(), r.solution.size());
        *energy = r.solution_energy;
    });
}

int32_t ige_server_wait(ige_server* s, uint64_t id, double timeout_ms, ige_request_info* info) {
    return guarded([&] {
        require(s && info, "NULL argument");
        std::unique_lock<std::mutex> lk(s->mu);
        auto done = [&] {
            auto it = s->reqs.find(id);
            return it == s->reqs.end() || it->second->status == IGE_REQ_DONE ||
                   it->second->status == IGE_REQ_FAILED;
        };
        s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done);
        auto it = s->reqs.find(id);
        require(it != s->reqs.end(), "unknown request id");
        fill_info(*it->second, info);
    });
}

int32_t ige_server_stats(ige_server* s, ige_server_counters* st) {
    return guarded([&] {
        require(s && st, "NULL argument");
        std::lock_guard<std::mutex> g(s->mu);
        std::memset(st, 0, sizeof(*st));
        st->queued = s->pending.size();
        st->live = s->live.size();
        st->completed = s->completed;
        st->rounds = s->rounds;
        st->attempts = s->attempts_done;
        st->arena_used = s->arena.used();
        st->arena_capacity = s->arena.capacity();
        st->max_live = s->max_live_seen;
        st->busy_ms = s->busy_ms;
    });
}

}  // extern "C"
