// descent_host.hpp -- host driver of the steepest-descent kernels, shared by
// the state API (ige_steepest_descent, ige_model_polish) and the serving loop.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "abi_util.hpp"
#include "kernels.hpp"

namespace ige {

struct DescentResult {
    uint32_t steps = 0;      // steps in which at least one replica flipped
    uint32_t converged = 0;  // 1 if a step found nothing to flip
    uint64_t flips = 0;
};

// Steepest descent of every replica of `st` (spins and tracked energies in
// place); flips[cells*64] receives one per flip.  m: the model in position
// space, pos_of its site -> position map (device), host_order its position ->
// site map.  dense / dense_fields: NULL, or the matrix-core view and a
// [cells][n][64] buffer, to take dE from the matrix-core fields.  clamped:
// NULL or n host bytes in site order.  Steps are launched in growing batches;
// the host reads the control words between batches and stops once a step has
// flipped nothing or max_steps steps ran.
inline DescentResult run_descent(const DevModel& m, const uint32_t* pos_of,
                                 const uint32_t* host_order, const DevDense* dense,
                                 int32_t* dense_fields, const DevState& st,
                                 const uint8_t* clamped, uint32_t max_steps, uint32_t* flips,
                                 hipStream_t stream) {
    using abi::check;
    using abi::DevBuf;
    const uint32_t n = m.n;
    abi::require(n < 0x80000000u, "descent needs n < 2^31");
    DevBuf<unsigned long long> keys, ctl;
    keys.alloc(static_cast<size_t>(st.cells) * 64);
    ctl.alloc(kDescentCtlWords);
    DevBuf<uint8_t> cl;
    DevBuf<uint32_t> aux;  // position -> site | clamped << 31
    if (clamped) {
        cl.alloc(n);
        cl.upload(clamped, n, stream);
        std::vector<uint32_t> a(n);
        for (uint32_t p = 0; p < n; ++p)
            a[p] = host_order[p] | (clamped[host_order[p]] ? 0x80000000u : 0u);
        aux.alloc(n);
        aux.upload(a.data(), n, stream);
        check(hipStreamSynchronize(stream), "descent upload");
    }
    check(launch_descent_reset(keys.p, st.cells, ctl.p, stream), "descent reset");
    unsigned long long c[kDescentCtlWords] = {};
    uint64_t launched = 0;
    uint32_t batch = 1;
    while (launched < max_steps) {
        const uint32_t k = static_cast<uint32_t>(std::min<uint64_t>(batch, max_steps - launched));
        for (uint32_t i = 0; i < k; ++i) {
            if (dense)
                check(launch_dense_fields(*dense, st, dense_fields, stream), "descent fields");
            check(launch_descent_step(m, st, dense ? dense_fields : nullptr, pos_of, cl.p,
                                      clamped ? aux.p : m.order, keys.p, flips, ctl.p, stream),
                  "descent step");
        }
        launched += k;
        ctl.download(c, kDescentCtlWords, stream);
        check(hipStreamSynchronize(stream), "descent");
        if (c[kDescentDone]) break;
        batch = std::min<uint32_t>(batch * 2, 64);
    }
    DescentResult r;
    r.steps = static_cast<uint32_t>(c[kDescentSteps]);
    r.converged = c[kDescentDone] ? 1u : 0u;
    r.flips = c[kDescentFlips];
    return r;
}

}  // namespace ige
