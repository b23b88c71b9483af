// pa.hip -- population-annealing resampling on the MSC words
// (apply_resample, src/solver/population_annealing.rs:229-286).
//
// A member is lane r of population p across that population's temperature
// levels t < T (temperature index p*T + t).  After resampling, the replica at
// temperature p_d*T + t, lane r_d holds what member parents[p_d*64 + r_d] =
// (p_s, r_s) held at p_s*T + t, lane r_s: its spins and its tracked energy.
// Configurations stay where the temperature permutation (temp_of / slot_of)
// puts them, so the kernel gathers, for every destination word, one bit per
// lane from that lane's source word (lane = replica, one ballot per word).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "sweep_generic.h"

namespace ige {

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWavesPerBlock = kBlock / 64;

__global__ __launch_bounds__(kBlock) void resample_kernel(
    const uint64_t* __restrict__ old_spins, uint64_t* __restrict__ spins,
    const int64_t* __restrict__ old_energy, int64_t* __restrict__ energy,
    const uint32_t* __restrict__ temp_of, const uint32_t* __restrict__ slot_of,
    const uint32_t* __restrict__ parents, uint32_t n, uint32_t cells, uint32_t pop_temps,
    uint32_t chunk, uint32_t waves_per_slot) {
    const uint32_t r = threadIdx.x & 63u;
    const uint32_t w = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
    if (w >= waves_per_slot * cells) return;
    const uint32_t d = w / waves_per_slot;
    const uint32_t p0 = (w % waves_per_slot) * chunk;
    const uint32_t p1 = min(n, p0 + chunk);
    const uint32_t g = temp_of[d * 64 + r];
    const uint32_t pop = g / pop_temps, t = g % pop_temps;
    const uint32_t m = parents[pop * 64 + r];
    const uint32_t rs = m % 64;
    const uint32_t src = slot_of[((m / 64) * pop_temps + t) * 64 + rs];
    if (p0 == 0) energy[d * 64 + r] = old_energy[src * 64 + rs];
    const uint64_t* S = old_spins + (uint64_t)src * n;
    uint64_t* D = spins + (uint64_t)d * n;
    for (uint32_t p = p0; p < p1; ++p) {
        const uint64_t b = wave_ballot((S[p] >> rs) & 1u);
        if (r == 0) D[p] = b;
    }
}

uint32_t ceil_div(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

}  // namespace

hipError_t launch_resample(const DevState& s, uint32_t n, const uint64_t* old_spins,
                           const int64_t* old_energy, const uint32_t* parents,
                           hipStream_t stream) {
    const uint32_t T = s.pop_temps ? s.pop_temps : s.cells;
    if (T == 0 || s.cells % T) return hipErrorInvalidValue;
    // about 16K waves, at least 64 positions each
    uint64_t chunk = ((uint64_t)n * s.cells + 16383) / 16384;
    if (chunk < 64) chunk = 64;
    const uint32_t waves_per_slot = ceil_div(n, chunk);
    hipLaunchKernelGGL(resample_kernel, dim3(ceil_div((uint64_t)waves_per_slot * s.cells,
                                                      kWavesPerBlock)),
                       dim3(kBlock), 0, stream, old_spins, s.spins, old_energy, s.energy,
                       s.temp_of, s.slot_of, parents, n, s.cells, T, (uint32_t)chunk,
                       waves_per_slot);
    return hipGetLastError();
}

}  // namespace ige
