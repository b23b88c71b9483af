// descent_host.hpp -- host driver of the steepest-descent kernels, shared by
// the state API (ige_steepest_descent, ige_steepest_descent_noalloc,
// ige_model_polish) and the serving loop.
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

// Device scratch of one descent: the per-replica argmin keys, the control
// words and, when sites are clamped, the clamp bytes and the scan's
// position -> site map.
struct DescentWorkspace {
    unsigned long long* keys = nullptr;  // cells * 64
    unsigned long long* ctl = nullptr;   // kDescentCtlWords
    uint8_t* clamped = nullptr;          // n (used when clamped)
    uint32_t* site_aux = nullptr;        // n (used when clamped)
};

// Steepest descent of every replica of `st` (spins and tracked energies in
// place); flips[cells*64] receives one per flip.  m: the model in position
// space, pos_of its site -> position map (device).  dense / dense_fields:
// NULL, or the matrix-core view and a [cells][n][64] buffer, to take dE from
// the matrix-core fields.  clamped: NULL or n host bytes in site order.
// Steps are launched in growing batches; the host reads the control words
// between batches and stops once a step has flipped nothing or max_steps
// steps ran.  Every launch has completed when this returns.
inline DescentResult run_descent_in(const DevModel& m, const uint32_t* pos_of,
                                    const DevDense* dense, int32_t* dense_fields,
                                    const DevState& st, const uint8_t* clamped,
                                    uint32_t max_steps, uint32_t* flips,
                                    const DescentWorkspace& ws, hipStream_t stream) {
    using abi::check;
    const uint32_t n = m.n;
    abi::require(n < 0x80000000u, "descent needs n < 2^31");
    if (clamped) {
        check(hipMemcpyAsync(ws.clamped, clamped, n, hipMemcpyHostToDevice, stream),
              "descent upload");
        check(launch_descent_aux(m.order, ws.clamped, n, ws.site_aux, stream), "descent aux");
    }
    check(launch_descent_reset(ws.keys, st.cells, ws.ctl, stream), "descent reset");
    unsigned long long c[kDescentCtlWords] = {};
    uint64_t launched = 0;
    uint32_t batch = 1;
    while (launched < max_steps) {
        const uint32_t k = static_cast<uint32_t>(std::min<uint64_t>(batch, max_steps - launched));
        for (uint32_t i = 0; i < k; ++i) {
            if (dense)
                check(launch_dense_fields(*dense, st, dense_fields, stream), "descent fields");
            check(launch_descent_step(m, st, dense ? dense_fields : nullptr, pos_of,
                                      clamped ? ws.clamped : nullptr,
                                      clamped ? ws.site_aux : m.order, ws.keys, flips, ws.ctl,
                                      stream),
                  "descent step");
        }
        launched += k;
        check(hipMemcpyAsync(c, ws.ctl, sizeof(c), hipMemcpyDeviceToHost, stream), "descent ctl");
        check(hipStreamSynchronize(stream), "descent");
        if (c[kDescentDone]) break;
        batch = std::min<uint32_t>(batch * 2, 64);
    }
    if (launched == 0) check(hipStreamSynchronize(stream), "descent");
    DescentResult r;
    r.steps = static_cast<uint32_t>(c[kDescentSteps]);
    r.converged = c[kDescentDone] ? 1u : 0u;
    r.flips = c[kDescentFlips];
    return r;
}

// run_descent_in with a workspace allocated for this call.
inline DescentResult run_descent(const DevModel& m, const uint32_t* pos_of,
                                 const DevDense* dense, int32_t* dense_fields,
                                 const DevState& st, const uint8_t* clamped, uint32_t max_steps,
                                 uint32_t* flips, hipStream_t stream) {
    abi::DevBuf<unsigned long long> keys, ctl;
    abi::DevBuf<uint8_t> cl;
    abi::DevBuf<uint32_t> aux;
    keys.alloc(static_cast<size_t>(st.cells) * 64);
    ctl.alloc(kDescentCtlWords);
    if (clamped) {
        cl.alloc(m.n);
        aux.alloc(m.n);
    }
    DescentWorkspace ws;
    ws.keys = keys.p;
    ws.ctl = ctl.p;
    ws.clamped = cl.p;
    ws.site_aux = aux.p;
    return run_descent_in(m, pos_of, dense, dense_fields, st, clamped, max_steps, flips, ws,
                          stream);
}

}  // namespace ige
