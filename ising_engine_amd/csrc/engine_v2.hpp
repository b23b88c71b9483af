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
//   - replicas come in cells of 64, and temperatures must be finite and > 0
//     for the operators that read temps[rep % len] (the reference's T = 0
//     pure descent is GpuSteepestDescent here).
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
    // MetropolisSweep, ReplicaExchange and GibbsColorSweep read
    // temps[rep % len] (one per replica, repeating); GpuMetropolisPt, the
    // engine::step form, takes one per cell of 64 replicas.
    std::vector<double> temperatures;
    size_t num_replicas = 0;
    double recent_acceptance = 0.0;
    double remaining_ms = 1e300;
};

// ---- state.rs: the replica mask ------------------------------------------
// Bit r set: replica r flips (state.rs:11-53).
class ReplicaMask {
   public:
    explicit ReplicaMask(size_t num_replicas = 0)
        : words_((num_replicas + 63) / 64, 0), len_(num_replicas) {}
    size_t len() const { return len_; }
    bool is_empty() const {
        return std::all_of(words_.begin(), words_.end(), [](uint64_t w) { return w == 0; });
    }
    void set(size_t r) { words_[r / 64] |= 1ull << (r % 64); }
    bool get(size_t r) const { return (words_[r / 64] >> (r % 64)) & 1u; }
    void clear() { std::fill(words_.begin(), words_.end(), 0ull); }
    const std::vector<uint64_t>& words() const { return words_; }
    uint32_t count() const {
        uint32_t c = 0;
        for (uint64_t w : words_) c += static_cast<uint32_t>(__builtin_popcountll(w));
        return c;
    }

   private:
    std::vector<uint64_t> words_;
    size_t len_;
};

// ---- the device state (a SpinState whose replicas live in HBM) ----------
//
// Replica numbering.  Replica rho is the configuration at temperature index
// rho / 64, lane rho % 64 (the reference's engine::step keeps temperature t in
// cell t), or, once the state is in the replica view
// (ensure_replica_temperatures), the configuration whose replica index is
// rho.  Either way the device's temperature permutation says where replica rho
// sits; on a state that never exchanged it is slot rho / 64, lane rho % 64.
//
// The per-site SpinState methods (state.rs:70-149) work on a host mirror of
// the words, tracked energies and permutation, downloaded once after device
// work and reused; delta_e_into and apply_flips add a host field ledger
// (n x replicas int64, built on first use).  Host flips are written back
// (ige_write_spins) before the next device call.  Operators bracket their
// device work with begin_device() / end_device().
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
        rp_.resize(n_ + 1);
        ci_.resize(info.nnz);
        w_.resize(info.nnz);
        lin_.resize(n_);
        check(ige_model_csr(model, rp_.data(), ci_.data(), w_.data()));
        check(ige_model_linear(model, lin_.data(), nullptr));
    }
    // SparseBitSlice::new (sparse_bitslice.rs:42-81): num_replicas copies of
    // init (n bytes, 0/1).  The device holds replicas in cells of 64, so any
    // other count is refused (the reference's backend choice then falls back
    // to a CPU state, evolution.rs:566-582).
    static std::unique_ptr<GpuBitSlice> with_replicas(ige_ctx* ctx, const ige_model* model,
                                                      size_t num_replicas, const uint8_t* init,
                                                      uint64_t seed = 0,
                                                      uint32_t kernel = IGE_KERNEL_AUTO) {
        if (num_replicas == 0 || num_replicas % 64 != 0 || num_replicas / 64 > 65536)
            throw Error(IGE_ERR_UNSUPPORTED,
                        "GpuBitSlice holds replicas in cells of 64: need a positive multiple "
                        "of 64");
        ige_model_info info{};
        check(ige_model_get_info(model, &info));
        const uint32_t cells = static_cast<uint32_t>(num_replicas / 64);
        std::vector<uint64_t> w(static_cast<size_t>(cells) * info.n);
        for (uint32_t c = 0; c < cells; ++c)
            for (uint32_t i = 0; i < info.n; ++i) w[static_cast<size_t>(c) * info.n + i] =
                (init[i] & 1u) ? ~0ull : 0ull;
        return std::make_unique<GpuBitSlice>(ctx, model, cells, seed, w.data(), kernel);
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
    bool replica_view() const { return replica_view_; }

    // The device handle, for device operators: host flips are written back
    // first, and the mirror is dropped afterwards (end_device).
    ige_state* device() { return state_; }
    void begin_device() {
        if (host_dirty_) {
            check(ige_write_spins(state_, words_.data()));
            host_dirty_ = false;
        }
    }
    void end_device() {
        mirror_ = false;
        fields_ = false;
    }

    // state.rs:77-79: dE of flipping `site` in every replica, from the ledger.
    void delta_e_into(size_t site, double* out) {
        ensure_fields();
        const size_t R = num_replicas();
        for (size_t rho = 0; rho < R; ++rho) {
            const size_t q = loc_[rho];
            const int64_t h = field_[site * R + q];
            out[rho] = static_cast<double>(bit(site, q) ? -h : h);
        }
    }
    // state.rs:80-81: flip `site` in the masked replicas, keeping energies and
    // fields exact (the reference's sparse ledger update, sparse_bitslice.rs:145-235).
    void apply_flips(size_t site, const ReplicaMask& mask) {
        ensure_fields();
        const size_t R = num_replicas();
        for (size_t rho = 0; rho < R; ++rho) {
            if (!mask.get(rho)) continue;
            const size_t q = loc_[rho];
            const bool x = bit(site, q);
            const int64_t h = field_[site * R + q];
            energy_[q] += x ? -h : h;
            words_[(q / 64) * n_ + site] ^= 1ull << (q % 64);
            for (uint32_t e = rp_[site]; e < rp_[site + 1]; ++e)
                field_[static_cast<size_t>(ci_[e]) * R + q] += x ? -w_[e] : w_[e];
            host_dirty_ = true;
        }
    }
    // state.rs:83-84: tracked energy per replica.
    void energies_into(double* out) {
        ensure_mirror();
        for (size_t rho = 0; rho < num_replicas(); ++rho)
            out[rho] = static_cast<double>(energy_[loc_[rho]]);
    }
    bool spin(size_t site, size_t replica) {
        ensure_mirror();
        return bit(site, loc_[replica]);
    }
    void extract_into(size_t replica, uint8_t* out) {
        ensure_mirror();
        const size_t q = loc_[replica];
        for (size_t i = 0; i < n_; ++i) out[i] = bit(i, q) ? 1u : 0u;
    }
    double overlap(size_t a, size_t b) {
        ensure_mirror();
        const size_t qa = loc_[a], qb = loc_[b];
        int64_t acc = 0;
        for (size_t i = 0; i < n_; ++i) acc += bit(i, qa) == bit(i, qb) ? 1 : -1;
        return n_ == 0 ? 1.0 : static_cast<double>(acc) / static_cast<double>(n_);
    }
    std::vector<uint32_t> coloring() const {
        std::vector<uint32_t> c(n_);
        check(ige_model_coloring(model_, c.data()));
        return c;
    }
    // state.rs:99-102: the sites sharing an edge with `site` (topology only).
    std::pair<const uint32_t*, size_t> neighbors(size_t site) const {
        return {ci_.data() + rp_[site], rp_[site + 1] - rp_[site]};
    }
    // state.rs:108-124 / :126-140: the defaults, through apply_flips.
    void copy_replica(size_t from, size_t to) {
        if (from == to) return;
        ReplicaMask mask(num_replicas());
        for (size_t site = 0; site < n_; ++site)
            if (spin(site, from) != spin(site, to)) {
                mask.clear();
                mask.set(to);
                apply_flips(site, mask);
            }
    }
    void swap_replicas(size_t a, size_t b) {
        if (a == b) return;
        ReplicaMask mask(num_replicas());
        for (size_t site = 0; site < n_; ++site)
            if (spin(site, a) != spin(site, b)) {
                mask.clear();
                mask.set(a);
                mask.set(b);
                apply_flips(site, mask);
            }
    }
    // state.rs:55-67: FNV-1a over the spin planes, site-major, little endian,
    // in the reference's order spins[site * w + word], replica word * 64 + bit.
    uint64_t digest() {
        ensure_mirror();
        uint64_t h = 0xcbf29ce484222325ull;
        for (size_t i = 0; i < n_; ++i)
            for (uint32_t c = 0; c < cells_; ++c) {
                uint64_t v = 0;
                for (uint32_t b = 0; b < 64; ++b)
                    v |= static_cast<uint64_t>(bit(i, loc_[static_cast<size_t>(c) * 64 + b])) << b;
                for (int k = 0; k < 8; ++k) {
                    h ^= (v >> (8 * k)) & 0xFFu;
                    h *= 0x100000001b3ull;
                }
            }
        return h;
    }
    // sparse_bitslice.rs:317-346: max |ledger - recomputed| over the tracked
    // energies (device, after writing host flips back) and, when the host
    // field ledger is in use, over the fields.
    double audit() {
        begin_device();
        std::vector<int64_t> e(num_replicas()), f(num_replicas());
        check(ige_read_energies(state_, e.data()));
        check(ige_recompute_energies(state_, f.data()));
        int64_t d = 0;
        for (size_t r = 0; r < e.size(); ++r) d = std::max<int64_t>(d, std::llabs(e[r] - f[r]));
        if (mirror_)
            for (size_t r = 0; r < e.size(); ++r)
                d = std::max<int64_t>(d, std::llabs(energy_[r] - f[r]));
        if (fields_) {
            const size_t R = num_replicas();
            std::vector<int32_t> fd(static_cast<size_t>(cells_) * n_ * 64);
            check(ige_read_fields(state_, fd.data()));
            for (uint32_t c = 0; c < cells_; ++c)
                for (size_t i = 0; i < n_; ++i)
                    for (uint32_t l = 0; l < 64; ++l) {
                        const size_t q = static_cast<size_t>(c) * 64 + l;
                        d = std::max<int64_t>(
                            d, std::llabs(field_[i * R + q] - fd[(c * n_ + i) * 64 + l]));
                    }
        }
        return static_cast<double>(d);
    }
    ige_stats stats(std::vector<uint64_t>* attempts = nullptr,
                    std::vector<uint64_t>* accepts = nullptr) {
        ige_stats s{};
        const size_t pairs = (replica_view_ ? num_replicas() : cells_) - 1;
        std::vector<uint64_t> a(std::max<size_t>(pairs, 1)), c(a.size());
        check(ige_read_stats(state_, &s, a.data(), c.data()));
        a.resize(pairs);
        c.resize(pairs);
        if (attempts) *attempts = a;
        if (accepts) *accepts = c;
        return s;
    }
    // The slot-ladder view (engine::step): one temperature per cell, T = 0 a
    // pure descent.  Uploaded only when it differs from the device's.
    void ensure_ladder(const std::vector<double>& temperatures) {
        if (temperatures.size() != cells_)
            throw Error(IGE_ERR_INVALID, "one temperature per cell of 64 replicas is required");
        std::vector<double> beta(cells_);
        for (uint32_t t = 0; t < cells_; ++t)
            beta[t] = temperatures[t] > 0.0 ? 1.0 / temperatures[t] : 1e300;  // T = 0: descent
        if (replica_view_ || beta != beta_) {
            begin_device();
            check(ige_set_ladder(state_, beta.data()));
            beta_ = beta;
            temps_.clear();
            replica_view_ = false;
            end_device();
        }
    }
    // The replica view (metropolis_sweep.rs:115): replica rho at
    // T = temperatures[rho % len], every T finite and > 0.
    void ensure_replica_temperatures(const std::vector<double>& temperatures) {
        if (temperatures.empty())
            throw Error(IGE_ERR_INVALID, "at least one temperature is required");
        if (!replica_view_ || temperatures != temps_) {
            begin_device();
            check(ige_set_replica_temperatures(state_, temperatures.data(),
                                               static_cast<uint32_t>(temperatures.size())));
            temps_ = temperatures;
            beta_.clear();
            replica_view_ = true;
            end_device();
        }
    }

   private:
    bool bit(size_t site, size_t q) const {
        return (words_[(q / 64) * n_ + site] >> (q % 64)) & 1u;
    }
    void ensure_mirror() {
        if (mirror_) return;
        const size_t R = num_replicas();
        words_.resize(static_cast<size_t>(cells_) * n_);
        energy_.resize(R);
        std::vector<uint32_t> tof(R);
        check(ige_read_spins(state_, words_.data()));
        check(ige_read_energies(state_, energy_.data()));
        check(ige_read_temp_of(state_, tof.data()));
        loc_.assign(R, 0);
        for (size_t q = 0; q < R; ++q)
            loc_[replica_view_ ? tof[q] : static_cast<size_t>(tof[q]) * 64 + q % 64] = q;
        mirror_ = true;
        host_dirty_ = false;
    }
    void ensure_fields() {
        ensure_mirror();
        if (fields_) return;
        const size_t R = num_replicas();
        field_.assign(n_ * R, 0);
        for (size_t i = 0; i < n_; ++i)
            for (size_t q = 0; q < R; ++q) {
                int64_t h = lin_[i];
                for (uint32_t e = rp_[i]; e < rp_[i + 1]; ++e)
                    if (bit(ci_[e], q)) h += w_[e];
                field_[i * R + q] = h;
            }
        fields_ = true;
    }

    const ige_model* model_;
    uint32_t cells_;
    size_t n_ = 0;
    size_t num_pairs_ = 0;
    ige_state* state_ = nullptr;
    std::vector<double> beta_, temps_;
    bool replica_view_ = false;
    std::vector<uint32_t> rp_, ci_;
    std::vector<int32_t> w_, lin_;
    // host mirror
    bool mirror_ = false, fields_ = false, host_dirty_ = false;
    std::vector<uint64_t> words_;   // [cells][n], site order (ige_read_spins)
    std::vector<int64_t> energy_;   // per position slot * 64 + lane
    std::vector<size_t> loc_;       // replica -> position
    std::vector<int64_t> field_;    // [site][position]
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
        state.ensure_replica_temperatures(view.temperatures);  // T = temps[rep % len]
        state.begin_device();
        const ige_stats before = state.stats();
        check(ige_run(state.device(), budget.sweeps, 0u, nullptr));
        const ige_stats after = state.stats();
        state.end_device();
        Report r;
        r.accepted = after.accepted - before.accepted;
        r.proposed = after.proposed - before.proposed;
        r.work = cost_model({state.num_vars(), state.num_pairs(), state.num_replicas()})
                     .work_per_sweep * budget.sweeps;
        return r;
    }
};

// replica_exchange.rs:61-109: adjacent replica indices, start = round & 1 of
// each apply, on temps[rep % len].
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
        if (state.num_replicas() < 2 || view.temperatures.empty()) return Report{};
        state.ensure_replica_temperatures(view.temperatures);
        state.begin_device();
        std::vector<uint64_t> a0, c0, a1, c1;
        state.stats(&a0, &c0);
        check(ige_exchange_replicas(state.device(), budget.sweeps));
        state.stats(&a1, &c1);
        state.end_device();
        Report r;
        for (size_t t = 0; t < a1.size(); ++t) {
            r.proposed += a1[t] - a0[t];
            r.accepted += c1[t] - c0[t];
        }
        r.work = static_cast<double>(budget.sweeps) * state.num_vars() * state.num_replicas();
        return r;
    }
};

// engine::step (engine.rs:573-819): sweep + DEO exchange per budget sweep, fused
// on the device (one stream, no host round trip between sweeps); one
// temperature per cell of 64 replicas, as engine::step's temps[t].
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
        state.begin_device();
        const ige_stats before = state.stats();
        check(ige_run(state.device(), budget.sweeps, IGE_RUN_EXCHANGE, nullptr));
        const ige_stats after = state.stats();
        state.end_device();
        Report r;
        r.accepted = after.accepted - before.accepted;
        r.proposed = after.proposed - before.proposed;
        r.work = cost_model({state.num_vars(), state.num_pairs(), state.num_replicas()})
                     .work_per_sweep * budget.sweeps;
        return r;
    }
};

// gibbs_color_sweep.rs:119-174 on the device: heat-bath colour sweeps, new
// spin 1 with probability 1 / (1 + exp(h / T)), T = temps[rep % len].
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
        state.ensure_replica_temperatures(view.temperatures);
        state.begin_device();
        const ige_stats before = state.stats();
        check(ige_run(state.device(), budget.sweeps, IGE_RUN_HEATBATH, nullptr));
        const ige_stats after = state.stats();
        state.end_device();
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
        state.begin_device();
        check(ige_steepest_descent(state.device(), budget.sweeps, nullptr, &rep));
        state.end_device();
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
