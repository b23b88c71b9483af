// abi_util.hpp -- shared plumbing of the C ABI implementation: error
// classes, the exception guard every entry point runs under, and an owning
// device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <stdexcept>
#include <string>

#include "../../include/ising_gpu.h"
#include "problem_ir.hpp"

struct ige_ctx {
    int device = 0;
};

namespace ige::abi {

inline thread_local std::string g_last_error;

struct Failure : std::runtime_error {
    int32_t code;
    Failure(int32_t c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw Failure(e == hipErrorOutOfMemory ? IGE_ERR_NOMEM : IGE_ERR_HIP,
                      std::string(what) + ": " + hipGetErrorString(e));
}

inline void require(bool ok, const std::string& msg) {
    if (!ok) throw Failure(IGE_ERR_INVALID, msg);
}

// Runs f, mapping every exception to a status code and a message.
template <class F>
int32_t guarded(F&& f) {
    try {
        f();
        return IGE_OK;
    } catch (const Failure& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const ige::RangeError& e) {
        g_last_error = e.what();
        return IGE_ERR_RANGE;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return IGE_ERR_INVALID;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return IGE_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return IGE_ERR_HIP;
    }
}

// Owning device buffer.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t count = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t n) {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        count = n;
        if (n) check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    }
    void upload(const T* src, size_t n, hipStream_t st) {
        if (n) check(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, st), "upload");
    }
    void download(T* dst, size_t n, hipStream_t st) const {
        if (n) check(hipMemcpyAsync(dst, p, n * sizeof(T), hipMemcpyDeviceToHost, st), "download");
    }
};

inline int context_device(const ige_ctx* ctx) { return ctx->device; }

// Sweep loops and batch anneals replay captured hipGraphs unless the
// environment sets IGE_NO_GRAPH to a non-empty value other than "0" (for tools
// that cannot follow graph launches); results are the same bits either way.
inline bool graphs_enabled() {
    static const bool on = [] {
        const char* v = std::getenv("IGE_NO_GRAPH");
        return !(v && *v && !(v[0] == '0' && v[1] == '\0'));
    }();
    return on;
}

// IGE_DENSE_HELPERS=0 keeps a dense sweep's fields in its own workgroup (the
// helpers' sums are the same integers, so the result is the same).
inline bool dense_helpers_enabled() {
    static const bool on = [] {
        const char* v = std::getenv("IGE_DENSE_HELPERS");
        return !(v && v[0] == '0' && v[1] == '\0');
    }();
    return on;
}

// IGE_CHAIN_BARRIERS=1 runs a one-slot chained run with a gang barrier after
// every colour class instead of per-site waits (the same bits); read at each
// call.
inline bool chain_barriers_forced() {
    const char* v = std::getenv("IGE_CHAIN_BARRIERS");
    return v && v[0] == '1' && v[1] == '\0';
}

// Polls a chained run's waits may make before they give up and raise the
// state's error flag: `fallback`, or IGE_CHAIN_SPIN_LIMIT when set (a test
// sets 0 so that waits expire at once); read at each call.
inline uint32_t chain_spin_limit(uint32_t fallback) {
    const char* v = std::getenv("IGE_CHAIN_SPIN_LIMIT");
    if (!v || !*v) return fallback;
    char* end = nullptr;
    const unsigned long x = std::strtoul(v, &end, 10);
    return (end && *end == '\0' && x <= 0xFFFFFFFFul) ? (uint32_t)x : fallback;
}

}  // namespace ige::abi
