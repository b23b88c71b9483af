// engine_v2.hpp -- C++ host layer over the C ABI, shaped like the reference's
// operator registry so the GPU sweep drops in as registered operators.
//
// Mirrors (paths relative to the reference checkout):
//   Nature, Budget, Report, CostEstimate, InstanceShape, Operator
//                                     src/engine_v2/operator.rs:17-95
//   Capability(Set), Constraints, Observable(Set), Complexity, Guarantees,
//   OperatorDescriptor, CapabilityQuery src/engine_v2/capability.rs:1-244
//   OperatorRegistry                  src/engine_v2/registry.rs:25-121
//   RuntimeView                       src/engine_v2/runtime.rs:27-33
//   SpinState (as GpuBitSlice)        src/engine_v2/state.rs:70-149
//
// Differences, all forced by the device state:
//   - apply() takes the GpuBitSlice itself: SpinState exposes no weights
//     (state.rs:99-102), so a device operator needs the device handle;
//   - the rng argument is accepted for signature parity and not consumed:
//     uniforms are keyed by the state's seed and its sweep/exchange counters,
//     which is what makes a resumed run bit-exact;
//   - temperatures are per cell of 64 replicas (view.temperatures[cell]),
//     where the reference indexes temps[rep % len].
// Errors from the C ABI are thrown as ige::v2::Error (the reference panics).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ising_gpu.h"

namespace ige::v2 {

struct Error : std::runtime_error {
    int32_t code;
    Error(int32_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int32_t rc) {
    if (rc != IGE_OK) throw Error(rc, ige_last_error());
}

// ---- operator.rs ---------------------------------------------------------
enum class Nature { TrajectoryPreserving, BehaviorChanging };

struct Budget {
    uint32_t sweeps = 1;
};

struct Report {
    uint64_t accepted = 0;
    uint64_t proposed = 0;
    std::optional<double> best_energy;
    double work = 0.0;
    std::vector<std::pair<std::string, double>> aux;
    double acceptance_rate() const {
        return proposed == 0 ? 0.0 : static_cast<double>(accepted) / static_cast<double>(proposed);
    }
};

struct CostEstimate {
    double work_per_sweep = 0.0;
};

struct InstanceShape {
    size_t n = 0;
    size_t num_pairs = 0;
    size_t num_replicas = 0;
};

// ---- capability.rs -------------------------------------------------------
enum class Capability : uint32_t {
    Exploration = 0,
    Exploitation = 1,
    BarrierCrossing = 2,
    ExactInference = 3,
    Approximate = 4,
    Warmstart = 5,
};

struct CapabilitySet {
    uint32_t bits = 0;
    static constexpr CapabilitySet empty() { return {}; }
    constexpr CapabilitySet with(Capability c) const {
        return CapabilitySet{bits | (1u << static_cast<uint32_t>(c))};
    }
    bool contains(Capability c) const { return bits & (1u << static_cast<uint32_t>(c)); }
    bool contains_all(CapabilitySet o) const { return (bits & o.bits) == o.bits; }
    bool intersects(CapabilitySet o) const { return (bits & o.bits) != 0; }
    bool operator==(const CapabilitySet& o) const { return bits == o.bits; }
};

enum class Backend { SparseBitSlice, DenseByte, GpuBitSlice };

struct Constraints {
    bool needs_integer = false;
    bool needs_float = false;
    bool supports_sparse = true;
    bool supports_dense = true;
    bool needs_replicas = false;
    bool needs_temperature = false;
    bool supports_gpu = false;  // new: the operator runs on the GpuBitSlice backend
    bool supports_backend(Backend b) const {
        switch (b) {
            case Backend::SparseBitSlice: return supports_sparse && !supports_gpu;
            case Backend::DenseByte: return supports_dense && !supports_gpu;
            case Backend::GpuBitSlice: return supports_gpu;
        }
        return false;
    }
};

enum class Observable : uint32_t {
    Acceptance = 0,
    EnergyDelta = 1,
    Entropy = 2,
    Diversity = 3,
    TemperatureDrift = 4,
};

struct ObservableSet {
    uint32_t bits = 0;
    constexpr ObservableSet with(Observable o) const {
        return ObservableSet{bits | (1u << static_cast<uint32_t>(o))};
    }
    bool contains(Observable o) const { return bits & (1u << static_cast<uint32_t>(o)); }
};

enum class Complexity { Constant, LinearVars, LinearEdges, Superlinear };

struct Guarantees {
    bool monotone = false;
    bool equilibrates = false;
    bool deterministic = false;
    bool energy_preserving = false;
};

struct OperatorDescriptor {
    std::string name;
    Nature nature = Nature::TrajectoryPreserving;
    CapabilitySet capabilities;
    Constraints constraints;
    ObservableSet observables;
    Guarantees guarantees;
    Complexity complexity = Complexity::LinearEdges;
};

struct CapabilityQuery {
    CapabilitySet required;
    CapabilitySet any_of;
    std::optional<Backend> backend;
    std::optional<bool> integral;
    std::optional<bool> multi_replica;
    std::optional<Nature> nature;

    // capability.rs:214-244, conjunctive.
    bool matches(const OperatorDescriptor& d) const {
        if (!d.capabilities.contains_all(required)) return false;
        if (!(any_of == CapabilitySet::empty()) && !d.capabilities.intersects(any_of))
            return false;
        if (backend && !d.constraints.supports_backend(*backend)) return false;
        if (integral == std::optional<bool>(false) && d.constraints.needs_integer) return false;
        if (integral == std::optional<bool>(true) && d.constraints.needs_float) return false;
        if (multi_replica == std::optional<bool>(false) && d.constraints.needs_replicas)
            return false;
        if (nature && d.nature != *nature) return false;
        return true;
    }
};

// ---- runtime.rs ----------------------------------------------------------
struct RuntimeView {
    uint64_t iteration = 0;
    std::vector<double> temperatures;  // one per cell
    size_t num_replicas = 0;
    double recent_acceptance = 0.0;
    double remaining_ms = 1e300;
};

// ---- the device state (a SpinState whose replicas live in HBM) ----------
class GpuBitSlice {
   public:
    // A model built by ProblemIR::from_pairs semantics (ir.rs:30-69).
    GpuBitSlice(ige_ctx* ctx, const ige_model* model, uint32_t cells, uint64_t seed,
                const uint64_t* init_words = nullptr, uint32_t kernel = IGE_KERNEL_AUTO)
        : model_(model), cells_(cells) {
        ige_model_info info{};
        check(ige_model_get_info(model, &info));
        n_ = info.n;
        num_pairs_ = info.nnz / 2;
        check(ige_state_create(ctx, model, cells, seed, init_words, kernel, &state_));
    }
    ~GpuBitSlice() {
        if (state_) ige_state_destroy(state_);
    }
    GpuBitSlice(const GpuBitSlice&) = delete;
    GpuBitSlice& operator=(const GpuBitSlice&) = delete;

    size_t num_vars() const { return n_; }
    size_t num_replicas() const { return 64ull * cells_; }
    size_t num_pairs() const { return num_pairs_; }
    uint32_t cells() const { return cells_; }
    ige_state* device() { return state_; }

    // state.rs:84-85: tracked energy per replica (replica = cell * 64 + lane).
    void energies_into(double* out) {
        std::vector<int64_t> e(num_replicas());
        check(ige_read_energies(state_, e.data()));
        for (size_t r = 0; r < e.size(); ++r) out[r] = static_cast<double>(e[r]);
    }
    std::vector<uint64_t> words() {
        std::vector<uint64_t> w(static_cast<size_t>(cells_) * n_);
        check(ige_read_spins(state_, w.data()));
        return w;
    }
    bool spin(size_t site, size_t replica) {
        const auto w = words();
        return (w[(replica / 64) * n_ + site] >> (replica % 64)) & 1u;
    }
    void extract_into(size_t replica, uint8_t* out) {
        const auto w = words();
        for (size_t i = 0; i < n_; ++i) out[i] = (w[(replica / 64) * n_ + i] >> (replica % 64)) & 1u;
    }
    double overlap(size_t a, size_t b) {
        const auto w = words();
        int64_t acc = 0;
        for (size_t i = 0; i < n_; ++i) {
            const bool sa = (w[(a / 64) * n_ + i] >> (a % 64)) & 1u;
            const bool sb = (w[(b / 64) * n_ + i] >> (b % 64)) & 1u;
            acc += sa == sb ? 1 : -1;
        }
        return n_ == 0 ? 1.0 : static_cast<double>(acc) / static_cast<double>(n_);
    }
    std::vector<uint32_t> coloring() const {
        std::vector<uint32_t> c(n_);
        check(ige_model_coloring(model_, c.data()));
        return c;
    }
    // state.rs:55-67: FNV-1a over the spin planes, site-major, little endian,
    // in the reference's word order spins[site * w + word] (word = cell here).
    uint64_t digest() {
        const auto w = words();
        uint64_t h = 0xcbf29ce484222325ull;
        for (size_t i = 0; i < n_; ++i)
            for (uint32_t c = 0; c < cells_; ++c) {
                const uint64_t v = w[static_cast<size_t>(c) * n_ + i];
                for (int b = 0; b < 8; ++b) {
                    h ^= (v >> (8 * b)) & 0xFFu;
                    h *= 0x100000001b3ull;
                }
            }
        return h;
    }
    // sparse_bitslice.rs:317-346: max |tracked - recomputed| energy (fields
    // are never ledgered here, so only energies can drift).
    double audit() {
        std::vector<int64_t> e(num_replicas()), f(num_replicas());
        check(ige_read_energies(state_, e.data()));
        check(ige_recompute_energies(state_, f.data()));
        int64_t d = 0;
        for (size_t r = 0; r < e.size(); ++r) d = std::max<int64_t>(d, std::llabs(e[r] - f[r]));
        return static_cast<double>(d);
    }
    ige_stats stats(std::vector<uint64_t>* attempts = nullptr,
                    std::vector<uint64_t>* accepts = nullptr) {
        ige_stats s{};
        std::vector<uint64_t> a(cells_ > 1 ? cells_ - 1 : 1), c(a.size());
        check(ige_read_stats(state_, &s, a.data(), c.data()));
        if (attempts) *attempts = a;
        if (accepts) *accepts = c;
        return s;
    }
    // Upload a ladder when it differs from the one on the device.
    void ensure_ladder(const std::vector<double>& temperatures) {
        if (temperatures.size() != cells_)
            throw Error(IGE_ERR_INVALID, "one temperature per cell of 64 replicas is required");
        std::vector<double> beta(cells_);
        for (uint32_t t = 0; t < cells_; ++t)
            beta[t] = temperatures[t] > 0.0 ? 1.0 / temperatures[t] : 1e300;  // T = 0: descent
        if (beta != beta_) {
            check(ige_set_ladder(state_, beta.data()));
            beta_ = beta;
        }
    }

   private:
    const ige_model* model_;
    uint32_t cells_;
    size_t n_ = 0;
    size_t num_pairs_ = 0;
    ige_state* state_ = nullptr;
    std::vector<double> beta_;
};

// ---- operator.rs: the uniform contract ----------------------------------
class Operator {
   public:
    virtual ~Operator() = default;
    virtual OperatorDescriptor descriptor() const = 0;
    virtual CostEstimate cost_model(InstanceShape shape) const = 0;
    virtual Report apply(GpuBitSlice& state, const RuntimeView& view, uint64_t& rng,
                         Budget budget) = 0;
    std::string name() const { return descriptor().name; }
    Nature nature() const { return descriptor().nature; }
};

// metropolis_sweep.rs:95-146 on the device: chromatic Metropolis sweeps.
class GpuMetropolisSweep : public Operator {
   public:
    static constexpr const char* NAME = "gpu_metropolis_sweep";
    OperatorDescriptor descriptor() const override {
        OperatorDescriptor d;
        d.name = NAME;
        d.capabilities = CapabilitySet::empty()
                             .with(Capability::Exploration)
                             .with(Capability::Exploitation);
        d.constraints = Constraints{true, false, true, false, false, true, true};
        d.observables = ObservableSet{}.with(Observable::Acceptance).with(Observable::EnergyDelta);
        d.guarantees.equilibrates = true;
        d.guarantees.deterministic = true;
        d.complexity = Complexity::LinearEdges;
        return d;
    }
    CostEstimate cost_model(InstanceShape s) const override {
        return {static_cast<double>((s.n + 2 * s.num_pairs) * std::max<size_t>(s.num_replicas, 1))};
    }
    Report apply(GpuBitSlice& state, const RuntimeView& view, uint64_t&, Budget budget) override {
        state.ensure_ladder(view.temperatures);
        const ige_stats before = state.stats();
        check(ige_run(state.device(), budget.sweeps, 0u, nullptr));
        const ige_stats after = state.stats();
        Report r;
        r.accepted = after.accepted - before.accepted;
        r.proposed = after.proposed - before.proposed;
        r.work = cost_model({state.num_vars(), state.num_pairs(), state.num_replicas()})
                     .work_per_sweep * budget.sweeps;
        return r;
    }
};

// replica_exchange.rs:61-109 with the DEO schedule of engine.rs:713-818.
class GpuReplicaExchange : public Operator {
   public:
    static constexpr const char* NAME = "gpu_replica_exchange";
    OperatorDescriptor descriptor() const override {
        OperatorDescriptor d;
        d.name = NAME;
        d.capabilities = CapabilitySet::empty().with(Capability::BarrierCrossing);
        d.constraints = Constraints{true, false, true, false, true, true, true};
        d.observables = ObservableSet{}.with(Observable::Acceptance);
        d.guarantees.deterministic = true;
        d.complexity = Complexity::LinearVars;
        return d;
    }
    CostEstimate cost_model(InstanceShape s) const override {
        return {static_cast<double>(s.n * std::max<size_t>(s.num_replicas, 1))};
    }
    Report apply(GpuBitSlice& state, const RuntimeView& view, uint64_t&, Budget budget) override {
        state.ensure_ladder(view.temperatures);
        std::vector<uint64_t> a0, c0, a1, c1;
        state.stats(&a0, &c0);
        check(ige_exchange(state.device(), budget.sweeps));
        state.stats(&a1, &c1);
        Report r;
        for (size_t t = 0; t + 1 < state.cells(); ++t) {
            r.proposed += a1[t] - a0[t];
            r.accepted += c1[t] - c0[t];
        }
        r.work = static_cast<double>(budget.sweeps) * state.num_vars() * state.num_replicas();
        return r;
    }
};

// engine::step (engine.rs:573-819): sweep + DEO exchange per budget sweep, fused
// on the device (one stream, no host round trip between sweeps).
class GpuMetropolisPt : public Operator {
   public:
    static constexpr const char* NAME = "gpu_metropolis_pt";
    OperatorDescriptor descriptor() const override {
        OperatorDescriptor d;
        d.name = NAME;
        d.capabilities = CapabilitySet::empty()
                             .with(Capability::Exploration)
                             .with(Capability::Exploitation)
                             .with(Capability::BarrierCrossing);
        d.constraints = Constraints{true, false, true, false, true, true, true};
        d.observables = ObservableSet{}.with(Observable::Acceptance).with(Observable::EnergyDelta);
        d.guarantees.equilibrates = true;
        d.guarantees.deterministic = true;
        d.complexity = Complexity::LinearEdges;
        return d;
    }
    CostEstimate cost_model(InstanceShape s) const override {
        return {static_cast<double>((s.n + 2 * s.num_pairs) * std::max<size_t>(s.num_replicas, 1))};
    }
    Report apply(GpuBitSlice& state, const RuntimeView& view, uint64_t&, Budget budget) override {
        state.ensure_ladder(view.temperatures);
        const ige_stats before = state.stats();
        check(ige_run(state.device(), budget.sweeps, IGE_RUN_EXCHANGE, nullptr));
        const ige_stats after = state.stats();
        Report r;
        r.accepted = after.accepted - before.accepted;
        r.proposed = after.proposed - before.proposed;
        r.work = cost_model({state.num_vars(), state.num_pairs(), state.num_replicas()})
                     .work_per_sweep * budget.sweeps;
        return r;
    }
};

// gibbs_color_sweep.rs:119-174 on the device: heat-bath colour sweeps, new
// spin 1 with probability 1 / (1 + exp(h / T)), one temperature per cell.
class GpuGibbsColorSweep : public Operator {
   public:
    static constexpr const char* NAME = "gpu_gibbs_color_sweep";
    OperatorDescriptor descriptor() const override {
        OperatorDescriptor d;
        d.name = NAME;
        d.capabilities = CapabilitySet::empty()
                             .with(Capability::Exploration)
                             .with(Capability::Exploitation);
        d.constraints = Constraints{true, false, true, false, false, true, true};
        d.observables = ObservableSet{}.with(Observable::Acceptance).with(Observable::EnergyDelta);
        d.guarantees.equilibrates = true;
        d.guarantees.deterministic = true;
        d.complexity = Complexity::LinearEdges;
        return d;
    }
    CostEstimate cost_model(InstanceShape s) const override {
        return {static_cast<double>((s.n + 2 * s.num_pairs) * std::max<size_t>(s.num_replicas, 1))};
    }
    Report apply(GpuBitSlice& state, const RuntimeView& view, uint64_t&, Budget budget) override {
        state.ensure_ladder(view.temperatures);
        const ige_stats before = state.stats();
        check(ige_run(state.device(), budget.sweeps, IGE_RUN_HEATBATH, nullptr));
        const ige_stats after = state.stats();
        Report r;
        r.accepted = after.accepted - before.accepted;
        r.proposed = after.proposed - before.proposed;
        r.work = static_cast<double>(state.num_vars()) * state.num_replicas() * budget.sweeps;
        return r;
    }
};

// steepest_descent.rs:80-124 on the device: per budget sweep, every replica
// flips its most improving site (lowest site on a tie); stops early once no
// replica can improve.  Monotone and deterministic; needs no temperature.
class GpuSteepestDescent : public Operator {
   public:
    static constexpr const char* NAME = "gpu_steepest_descent";
    OperatorDescriptor descriptor() const override {
        OperatorDescriptor d;
        d.name = NAME;
        d.capabilities = CapabilitySet::empty().with(Capability::Exploitation);
        d.constraints = Constraints{true, false, true, false, false, false, true};
        d.observables = ObservableSet{}.with(Observable::Acceptance).with(Observable::EnergyDelta);
        d.guarantees.monotone = true;
        d.guarantees.deterministic = true;
        d.complexity = Complexity::LinearEdges;
        return d;
    }
    CostEstimate cost_model(InstanceShape s) const override {
        return {static_cast<double>(s.n * std::max<size_t>(s.num_replicas, 1))};
    }
    Report apply(GpuBitSlice& state, const RuntimeView&, uint64_t&, Budget budget) override {
        ige_descent_report rep{};
        check(ige_steepest_descent(state.device(), budget.sweeps, nullptr, &rep));
        Report r;
        r.accepted = rep.flips;
        r.proposed = rep.proposed;
        r.work = static_cast<double>(budget.sweeps) * state.num_vars() * state.num_replicas();
        return r;
    }
};

// ---- registry.rs ---------------------------------------------------------
class OperatorRegistry {
   public:
    using Factory = std::function<std::unique_ptr<Operator>()>;

    // The GPU operator library.  Selection is by capability (find), not name.
    static OperatorRegistry standard_gpu() {
        OperatorRegistry reg;
        reg.register_factory([] { return std::make_unique<GpuMetropolisSweep>(); });
        reg.register_factory([] { return std::make_unique<GpuReplicaExchange>(); });
        reg.register_factory([] { return std::make_unique<GpuMetropolisPt>(); });
        reg.register_factory([] { return std::make_unique<GpuSteepestDescent>(); });
        reg.register_factory([] { return std::make_unique<GpuGibbsColorSweep>(); });
        return reg;
    }
    // registry.rs:70-82: the descriptor is read once, up front.
    void register_factory(Factory f) {
        OperatorDescriptor d = f()->descriptor();
        const std::string key = d.name;
        entries_[key] = Entry{std::move(d), std::move(f)};
    }
    std::unique_ptr<Operator> lookup(const std::string& name) const {
        auto it = entries_.find(name);
        return it == entries_.end() ? nullptr : it->second.factory();
    }
    const OperatorDescriptor* metadata(const std::string& name) const {
        auto it = entries_.find(name);
        return it == entries_.end() ? nullptr : &it->second.descriptor;
    }
    bool contains(const std::string& name) const { return entries_.count(name) != 0; }
    size_t len() const { return entries_.size(); }
    bool is_empty() const { return entries_.empty(); }
    std::vector<std::string> names() const {
        std::vector<std::string> v;
        for (const auto& e : entries_) v.push_back(e.first);
        return v;
    }
    // registry.rs:112-121: matching names, sorted for deterministic selection.
    std::vector<std::string> find(const CapabilityQuery& q) const {
        std::vector<std::string> out;
        for (const auto& e : entries_)
            if (q.matches(e.second.descriptor)) out.push_back(e.first);
        std::sort(out.begin(), out.end());
        return out;
    }

   private:
    struct Entry {
        OperatorDescriptor descriptor;
        Factory factory;
    };
    std::map<std::string, Entry> entries_;
};

}  // namespace ige::v2
