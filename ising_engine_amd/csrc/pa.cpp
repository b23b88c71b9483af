// pa.cpp -- population-annealing weights and resampling plans (host side),
// restating src/solver/population_annealing.rs:45-128 and :229-252.
// The uniforms that the reference draws from its generator come here from
// Philox4x32-10 keyed by (seed), counter (event, k, 0, tag 7), so a plan is a
// pure function of its inputs.
#include <cmath>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>

#include "abi_util.hpp"
#include "ige_math.h"

using ige::abi::guarded;
using ige::abi::require;

namespace {

constexpr uint32_t kTagResample = 7u;

double max_of(const double* v, uint32_t n) {
    double m = -std::numeric_limits<double>::infinity();
    for (uint32_t i = 0; i < n; ++i) m = std::fmax(m, v[i]);
    return m;
}

}  // namespace

extern "C" {

// normalized_weights + effective_sample_size (population_annealing.rs:45-60).
int32_t ige_pa_normalize(const double* log_w, uint32_t count, double* w_norm, double* ess) {
    return guarded([&] {
        require(log_w && w_norm && count >= 1, "NULL argument or empty population");
        const double m = max_of(log_w, count);
        double z = 0.0;
        for (uint32_t i = 0; i < count; ++i) {
            w_norm[i] = std::exp(log_w[i] - m);
            z += w_norm[i];
        }
        double s2 = 0.0;
        for (uint32_t i = 0; i < count; ++i) {
            w_norm[i] /= z;
            s2 += w_norm[i] * w_norm[i];
        }
        if (ess) *ess = 1.0 / s2;
    });
}

// free_energy_increment (population_annealing.rs:62-75): logsumexp(log_w) - ln R.
int32_t ige_pa_free_energy_increment(const double* log_w, uint32_t count, double* out) {
    return guarded([&] {
        require(log_w && out && count >= 1, "NULL argument or empty population");
        const double m = max_of(log_w, count);
        if (!std::isfinite(m)) {
            *out = 0.0;
            return;
        }
        double s = 0.0;
        for (uint32_t i = 0; i < count; ++i) s += std::exp(log_w[i] - m);
        *out = m + std::log(s) - std::log(static_cast<double>(count));
    });
}

// systematic_resample (population_annealing.rs:103-124): positions (j + u) / R
// against the cumulative weights; unplaced positions go to the last parent.
int32_t ige_pa_systematic_resample(const double* w_norm, uint32_t count, double u,
                                   uint32_t* counts) {
    return guarded([&] {
        require(w_norm && counts && count >= 1, "NULL argument or empty population");
        require(u >= 0.0 && u < 1.0, "u must be in [0, 1)");
        std::memset(counts, 0, sizeof(uint32_t) * count);
        double cum = 0.0;
        uint32_t j = 0;
        for (uint32_t i = 0; i < count; ++i) {
            cum += w_norm[i];
            while (j < count && ((static_cast<double>(j) + u) / count) < cum) {
                counts[i] += 1;
                ++j;
            }
        }
        if (j < count) counts[count - 1] += count - j;
    });
}

// apply_resample's parent list (population_annealing.rs:237-246): parent i
// repeated counts[i] times, then a Fisher-Yates shuffle from the back with
// j = floor(u_k (k + 1) / 2^32), u_k = Philox(seed; event, k, 0, tag 7)[0].
int32_t ige_pa_parents(const uint32_t* counts, uint32_t count, uint64_t seed, uint32_t event,
                       uint32_t* parents) {
    return guarded([&] {
        require(counts && parents && count >= 1, "NULL argument or empty population");
        uint64_t total = 0;
        for (uint32_t i = 0; i < count; ++i) total += counts[i];
        require(total == count, "counts must sum to the population size");
        uint32_t k = 0;
        for (uint32_t i = 0; i < count; ++i)
            for (uint32_t c = 0; c < counts[i]; ++c) parents[k++] = i;
        for (uint32_t q = count - 1; q >= 1; --q) {
            const ige::U32x4 o = ige::philox4x32_10(event, q, 0u, kTagResample << 24,
                                                    static_cast<uint32_t>(seed),
                                                    static_cast<uint32_t>(seed >> 32));
            const uint32_t j = static_cast<uint32_t>((static_cast<uint64_t>(o.v[0]) * (q + 1)) >> 32);
            const uint32_t tmp = parents[q];
            parents[q] = parents[j];
            parents[j] = tmp;
        }
    });
}

// family_entropy (population_annealing.rs:77-101): rho = R sum_f (n_f / R)^2.
int32_t ige_pa_family_entropy(const uint32_t* family, uint32_t count, double* rho,
                              uint32_t* families) {
    return guarded([&] {
        require(family && rho && count >= 1, "NULL argument or empty population");
        std::unordered_map<uint32_t, uint32_t> c;
        for (uint32_t i = 0; i < count; ++i) c[family[i]] += 1;
        double s = 0.0;
        for (const auto& kv : c) {
            const double f = static_cast<double>(kv.second) / count;
            s += f * f;
        }
        *rho = s * count;
        if (families) *families = static_cast<uint32_t>(c.size());
    });
}

// adapt_spacings (population_annealing.rs:137-148): d'_t = d_t sqrt(max(A_t,
// 0.02)), rescaled to the old total span; sums taken in index order.
int32_t ige_pa_adapt_spacings(const double* spacings, const double* acceptance, uint32_t count,
                              double* out) {
    return guarded([&] {
        require(spacings && acceptance && out && count >= 1, "NULL argument or no spacings");
        constexpr double kFloor = 0.02;
        std::vector<double> raw(count);
        double total_old = 0.0, total_new = 0.0;
        for (uint32_t t = 0; t < count; ++t) {
            raw[t] = spacings[t] * std::sqrt(std::fmax(acceptance[t], kFloor));
            total_old += spacings[t];
        }
        for (uint32_t t = 0; t < count; ++t) total_new += raw[t];
        for (uint32_t t = 0; t < count; ++t) out[t] = raw[t] * total_old / total_new;
    });
}

}  // extern "C"
