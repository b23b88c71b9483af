// test_engine_v2.cpp -- operator-registry tests in the shape of the
// reference's own (src/engine_v2/registry.rs:124-207,
// metropolis_sweep.rs:174-214, replica_exchange.rs:138-196), run against the
// GPU operators.  `--cpu` runs only the parts that need no device.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../ising_engine_amd/csrc/engine_v2.hpp"

using namespace ige::v2;

static int failures = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

// A mock operator, as registry.rs:136-168 registers one.
struct Mock : Operator {
    OperatorDescriptor descriptor() const override {
        OperatorDescriptor d;
        d.name = "mock";
        d.capabilities = CapabilitySet::empty().with(Capability::Warmstart);
        d.constraints = Constraints{};
        return d;
    }
    CostEstimate cost_model(InstanceShape s) const override { return {double(s.n)}; }
    Report apply(GpuBitSlice&, const RuntimeView&, uint64_t&, Budget) override { return {}; }
};

static void registry_tests() {
    OperatorRegistry reg = OperatorRegistry::standard_gpu();
    CHECK(reg.len() == 5);
    CHECK(reg.contains("gpu_metropolis_sweep"));
    CHECK(reg.contains("gpu_gibbs_color_sweep"));
    CHECK(reg.contains("gpu_steepest_descent"));
    CHECK(reg.contains("gpu_replica_exchange"));
    CHECK(reg.contains("gpu_metropolis_pt"));
    CHECK(reg.lookup("nope") == nullptr);
    auto op = reg.lookup("gpu_metropolis_pt");
    CHECK(op && op->name() == "gpu_metropolis_pt");
    CHECK(op->nature() == Nature::TrajectoryPreserving);
    reg.register_factory([] { return std::make_unique<Mock>(); });
    CHECK(reg.len() == 6);
    const OperatorDescriptor* md = reg.metadata("mock");
    CHECK(md && md->capabilities.contains(Capability::Warmstart));

    CapabilityQuery q;
    q.required = CapabilitySet::empty().with(Capability::BarrierCrossing);
    auto found = reg.find(q);
    CHECK((found == std::vector<std::string>{"gpu_metropolis_pt", "gpu_replica_exchange"}));

    CapabilityQuery g;
    g.backend = Backend::GpuBitSlice;
    found = reg.find(g);
    CHECK(found.size() == 5 && std::is_sorted(found.begin(), found.end()));
    CapabilityQuery cpu;
    cpu.backend = Backend::SparseBitSlice;
    CHECK((reg.find(cpu) == std::vector<std::string>{"mock"}));

    CapabilityQuery floaty;  // a float instance excludes needs_integer operators
    floaty.integral = false;
    CHECK((reg.find(floaty) == std::vector<std::string>{"mock"}));

    CapabilityQuery single;  // one replica: exchange needs several
    single.multi_replica = false;
    single.backend = Backend::GpuBitSlice;
    CHECK((reg.find(single) == std::vector<std::string>{"gpu_gibbs_color_sweep",
                                                        "gpu_metropolis_sweep",
                                                        "gpu_steepest_descent"}));
    const OperatorDescriptor* sd = reg.metadata("gpu_steepest_descent");
    CHECK(sd && sd->guarantees.monotone && sd->guarantees.deterministic &&
          !sd->constraints.needs_temperature);

    CapabilityQuery any;
    any.any_of = CapabilitySet::empty().with(Capability::Exploitation).with(Capability::Warmstart);
    CHECK(reg.find(any).size() == 5);
}

// 3D +-J L=4 periodic, Ising lowered to QUBO (pair -4J, linear 2 sum J).
static ige_model* lattice(ige_ctx* ctx, uint32_t L, uint64_t seed) {
    const uint32_t n = L * L * L;
    std::vector<uint32_t> pi, pj;
    std::vector<int32_t> q, lin(n, 0);
    uint64_t s = seed;
    auto next = [&] {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (s >> 33) & 1u;
    };
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t x = i % L, y = (i / L) % L, z = i / (L * L);
        const uint32_t nb[3] = {(x + 1) % L + y * L + z * L * L, x + ((y + 1) % L) * L + z * L * L,
                                x + y * L + ((z + 1) % L) * L * L};
        for (uint32_t j : nb) {
            const int J = next() ? 1 : -1;
            pi.push_back(std::min(i, j));
            pj.push_back(std::max(i, j));
            q.push_back(-4 * J);
            lin[i] += 2 * J;
            lin[j] += 2 * J;
        }
    }
    int64_t offset = 0;
    for (int32_t v : q) offset += v / 4;  // -sum J = sum(q)/4
    ige_model* m = nullptr;
    check(ige_model_from_pairs(ctx, n, offset, lin.data(), q.size(), pi.data(), pj.data(),
                               q.data(), &m));
    return m;
}

static void gpu_tests() {
    ige_ctx* ctx = nullptr;
    check(ige_context_create(0, &ctx));
    ige_model* m = lattice(ctx, 4, 9);
    OperatorRegistry reg = OperatorRegistry::standard_gpu();
    RuntimeView view;
    view.temperatures = {3.0, 1.5, 0.9, 0.5};
    uint64_t rng = 11;

    // Identical seeds -> identical trajectories (metropolis_sweep.rs:174-214).
    GpuBitSlice a(ctx, m, 4, 1234), b(ctx, m, 4, 1234);
    auto sweep = reg.lookup("gpu_metropolis_sweep");
    Report ra = sweep->apply(a, view, rng, Budget{30});
    Report rb = reg.lookup("gpu_metropolis_sweep")->apply(b, view, rng, Budget{30});
    CHECK(ra.accepted == rb.accepted && ra.proposed == rb.proposed);
    CHECK(ra.proposed == 30ull * 64 * 64 * 4);
    CHECK(ra.accepted > 0 && ra.accepted <= ra.proposed);
    std::vector<double> ea(a.num_replicas()), eb(b.num_replicas());
    a.energies_into(ea.data());
    b.energies_into(eb.data());
    CHECK(ea == eb);
    CHECK(a.digest() == b.digest());
    CHECK(a.audit() == 0.0);

    // Exchange permutes configurations between temperatures: the multiset of
    // energies per lane is invariant (replica_exchange.rs:138-196).
    std::vector<double> before(ea);
    auto exch = reg.lookup("gpu_replica_exchange");
    Report rx = exch->apply(a, view, rng, Budget{6});
    CHECK(rx.proposed == 64ull * (3 * 2 + 3 * 1));  // 3 even rounds x 2 pairs, 3 odd x 1
    CHECK(rx.accepted <= rx.proposed);
    std::vector<double> after(a.num_replicas());
    a.energies_into(after.data());
    CHECK(after == before);  // spins stay in their slots; only temperatures move
    std::vector<uint32_t> tof(a.num_replicas());
    check(ige_read_temp_of(a.device(), tof.data()));
    for (uint32_t r = 0; r < 64; ++r) {
        std::vector<uint32_t> col;
        for (uint32_t c = 0; c < 4; ++c) col.push_back(tof[c * 64 + r]);
        std::sort(col.begin(), col.end());
        CHECK((col == std::vector<uint32_t>{0, 1, 2, 3}));
    }

    // Fused PT operator and a T = 0 cell (pure descent) stay exact.
    auto pt = reg.lookup("gpu_metropolis_pt");
    RuntimeView cold = view;
    cold.temperatures = {3.0, 1.0, 0.4, 0.0};
    Report rp = pt->apply(a, cold, rng, Budget{40});
    CHECK(rp.proposed == 40ull * 64 * 64 * 4);
    CHECK(a.audit() == 0.0);
    ige_stats st = a.stats();
    CHECK(st.sweeps == 70 && st.exchanges == 46);

    // Steepest descent (steepest_descent.rs:80-124): monotone per replica, exact,
    // and once at a local minimum a step flips nothing and ends the call.
    std::vector<double> e0(a.num_replicas()), e1(a.num_replicas());
    a.energies_into(e0.data());
    auto sd = reg.lookup("gpu_steepest_descent");
    Report rd = sd->apply(a, view, rng, Budget{1000});
    a.energies_into(e1.data());
    for (size_t r = 0; r < e0.size(); ++r) CHECK(e1[r] <= e0[r]);
    CHECK(rd.accepted > 0 && a.audit() == 0.0);
    Report rz = sd->apply(a, view, rng, Budget{5});
    CHECK(rz.accepted == 0 && rz.proposed == 64ull * 64 * 4);

    // Heat-bath sweeps (gibbs_color_sweep.rs:119-174): exact bookkeeping, one proposal
    // per (site, replica) and sweep, and the sweep counter shared with Metropolis.
    auto gibbs = reg.lookup("gpu_gibbs_color_sweep");
    const ige_stats g0 = a.stats();
    Report rg = gibbs->apply(a, view, rng, Budget{12});
    CHECK(rg.proposed == 12ull * 64 * 64 * 4 && rg.accepted > 0 && rg.accepted < rg.proposed);
    CHECK(a.stats().sweeps == g0.sweeps + 12 && a.audit() == 0.0);

    // A wrong-sized ladder is an error, not a silent fallback.
    RuntimeView bad;
    bad.temperatures = {1.0};
    bool threw = false;
    try {
        sweep->apply(a, bad, rng, Budget{1});
    } catch (const Error& e) {
        threw = e.code == IGE_ERR_INVALID;
    }
    CHECK(threw);

    ige_model_destroy(m);
    ige_context_destroy(ctx);
}

int main(int argc, char** argv) {
    const bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
    registry_tests();
    if (!cpu_only) {
        try {
            gpu_tests();
        } catch (const Error& e) {
            std::fprintf(stderr, "FAIL: ige error %d: %s\n", e.code, e.what());
            ++failures;
        }
    }
    std::printf("%s: %d failure(s)\n", cpu_only ? "registry" : "registry+gpu", failures);
    return failures == 0 ? 0 : 1;
}
