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

    // Exchange permutes configurations between replicas: the multiset of
    // energies is invariant (replica_exchange.rs:138-196).  Pairs (k, k + 1)
    // from start 0, 1, 0, 1, 0, 1 over 256 replicas: 3 x 128 + 3 x 127.
    std::vector<double> before(ea);
    auto exch = reg.lookup("gpu_replica_exchange");
    Report rx = exch->apply(a, view, rng, Budget{6});
    CHECK(rx.proposed == 3ull * 128 + 3ull * 127);
    CHECK(rx.accepted > 0 && rx.accepted <= rx.proposed);
    std::vector<double> after(a.num_replicas());
    a.energies_into(after.data());
    std::vector<double> sb(before), sa(after);
    std::sort(sb.begin(), sb.end());
    std::sort(sa.begin(), sa.end());
    CHECK(sa == sb);
    CHECK(after != before);
    CHECK(a.audit() == 0.0);

    // Fused PT operator (one temperature per cell) and a T = 0 cell (pure
    // descent) stay exact; switching back to the cell ladder keeps the counters.
    auto pt = reg.lookup("gpu_metropolis_pt");
    RuntimeView cold = view;
    cold.temperatures = {3.0, 1.0, 0.4, 0.0};
    Report rp = pt->apply(a, cold, rng, Budget{40});
    CHECK(rp.proposed == 40ull * 64 * 64 * 4);
    CHECK(a.audit() == 0.0 && !a.replica_view());
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

    // A wrong-sized cell ladder and a T = 0 replica temperature are errors, not
    // silent fallbacks; a replica count off the cells of 64 is refused.
    RuntimeView bad;
    bad.temperatures = {1.0};
    bool threw = false;
    try {
        pt->apply(a, bad, rng, Budget{1});
    } catch (const Error& e) {
        threw = e.code == IGE_ERR_INVALID;
    }
    CHECK(threw);
    RuntimeView zero;
    zero.temperatures = {1.0, 0.0};
    threw = false;
    try {
        sweep->apply(a, zero, rng, Budget{1});
    } catch (const Error& e) {
        threw = e.code == IGE_ERR_INVALID;
    }
    CHECK(threw);
    threw = false;
    std::vector<uint8_t> init(64, 0);
    try {
        (void)GpuBitSlice::with_replicas(ctx, m, 100, init.data());
    } catch (const Error& e) {
        threw = e.code == IGE_ERR_UNSUPPORTED;
    }
    CHECK(threw);

    // The per-site SpinState API on the host mirror (state.rs:70-149).
    std::vector<double> de(a.num_replicas()), e_before(a.num_replicas()),
        e_after(a.num_replicas());
    a.energies_into(e_before.data());
    a.delta_e_into(5, de.data());
    ReplicaMask mask(a.num_replicas());
    mask.set(0);
    mask.set(77);
    mask.set(255);
    const bool s0 = a.spin(5, 0);
    a.apply_flips(5, mask);
    CHECK(a.spin(5, 0) != s0);
    a.energies_into(e_after.data());
    for (size_t r = 0; r < a.num_replicas(); ++r)
        CHECK(e_after[r] == e_before[r] + (mask.get(r) ? de[r] : 0.0));
    std::vector<double> de2(a.num_replicas());
    a.delta_e_into(5, de2.data());
    CHECK(de2[0] == -de[0] && de2[1] == de[1]);
    auto nb = a.neighbors(5);
    CHECK(nb.second == 6);  // 3D lattice
    a.swap_replicas(3, 200);
    a.copy_replica(9, 10);
    std::vector<uint8_t> x9(a.num_vars()), x10(a.num_vars());
    a.extract_into(9, x9.data());
    a.extract_into(10, x10.data());
    CHECK(x9 == x10 && a.overlap(9, 10) == 1.0);
    CHECK(a.audit() == 0.0);  // host ledger, written back, equals a device recompute
    sweep->apply(a, view, rng, Budget{3});
    CHECK(a.audit() == 0.0);

    ige_model_destroy(m);
    ige_context_destroy(ctx);
}

// The operator layer against the oracle (driven from tests/test_engine_v2_cpp.py):
// the frustrated seven-site instance of the reference's operator tests
// (metropolis_sweep.rs:156-172), num_replicas copies of init 1010010, a fixed
// sequence of operators and host per-site calls; writes each step's Report and
// then every replica's configuration, energy, the digest and the audit.
static int dump(const char* in_path, const char* out_path) {
    FILE* fi = std::fopen(in_path, "rb");
    if (!fi) return 2;
    uint64_t seed = 0;
    uint32_t nrep = 0, ntemps = 0;
    bool ok = std::fread(&seed, 8, 1, fi) == 1 && std::fread(&nrep, 4, 1, fi) == 1 &&
              std::fread(&ntemps, 4, 1, fi) == 1;
    std::vector<double> temps(ntemps);
    ok = ok && std::fread(temps.data(), 8, ntemps, fi) == ntemps;
    std::fclose(fi);
    if (!ok) return 2;
    ige_ctx* ctx = nullptr;
    check(ige_context_create(0, &ctx));
    const int32_t lin[7] = {1, -2, 3, 0, -1, 2, 1};
    const uint32_t pi[8] = {0, 1, 2, 0, 3, 4, 5, 2}, pj[8] = {1, 2, 3, 3, 4, 5, 6, 6};
    const int32_t q[8] = {-1, 2, -3, 1, -1, 2, -2, 1};
    ige_model* m = nullptr;
    check(ige_model_from_pairs(ctx, 7, -2, lin, 8, pi, pj, q, &m));
    const uint8_t init[7] = {1, 0, 1, 0, 0, 1, 0};
    auto st = GpuBitSlice::with_replicas(ctx, m, nrep, init, seed, IGE_KERNEL_GENERIC);
    OperatorRegistry reg = OperatorRegistry::standard_gpu();
    RuntimeView view;
    view.temperatures = temps;
    uint64_t rng = 0;
    std::vector<Report> reps;
    reps.push_back(reg.lookup("gpu_gibbs_color_sweep")->apply(*st, view, rng, Budget{5}));
    reps.push_back(reg.lookup("gpu_replica_exchange")->apply(*st, view, rng, Budget{6}));
    reps.push_back(reg.lookup("gpu_metropolis_sweep")->apply(*st, view, rng, Budget{30}));
    reps.push_back(reg.lookup("gpu_replica_exchange")->apply(*st, view, rng, Budget{3}));
    ReplicaMask mask(st->num_replicas());
    mask.set(0);
    mask.set(5);
    mask.set(nrep - 1);
    st->apply_flips(2, mask);
    st->swap_replicas(3, 4);
    st->copy_replica(10, 11);
    reps.push_back(reg.lookup("gpu_metropolis_sweep")->apply(*st, view, rng, Budget{4}));
    FILE* fo = std::fopen(out_path, "wb");
    if (!fo) return 2;
    for (const Report& r : reps) {
        std::fwrite(&r.accepted, 8, 1, fo);
        std::fwrite(&r.proposed, 8, 1, fo);
    }
    std::vector<uint8_t> x(7);
    for (size_t r = 0; r < st->num_replicas(); ++r) {
        st->extract_into(r, x.data());
        std::fwrite(x.data(), 1, 7, fo);
    }
    std::vector<double> e(st->num_replicas());
    st->energies_into(e.data());
    std::fwrite(e.data(), 8, e.size(), fo);
    const uint64_t dg = st->digest();
    std::fwrite(&dg, 8, 1, fo);
    const double au = st->audit();
    std::fwrite(&au, 8, 1, fo);
    std::fclose(fo);
    st.reset();
    ige_model_destroy(m);
    ige_context_destroy(ctx);
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4 && std::strcmp(argv[1], "--dump") == 0) {
        try {
            return dump(argv[2], argv[3]);
        } catch (const Error& e) {
            std::fprintf(stderr, "FAIL: ige error %d: %s\n", e.code, e.what());
            return 1;
        }
    }
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
