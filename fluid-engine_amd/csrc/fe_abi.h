// fe_abi.h — host-side helpers of the C ABI: status codes with a per-thread message, HIP
// error mapping and the device guard every launching entry point takes.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "fe_internal.h"

namespace fe {

// Records msg as the calling thread's last error and returns code.
fe_status fail(fe_status code, const std::string& msg);

inline fe_status hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? FE_ERR_OUT_OF_MEMORY : FE_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

// Makes the context's device current for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// Drop a captured frame graph (its kernel arguments are stale after any change).
void drop_graph(fe_ctx* c);

// Warm-start hooks of the Jacobi solves (fe_learn_api.cpp): before the first sweep
// (features, prediction) and after the last (samples, the pressures kept for the next solve).
fe_status learn_before_solve(fe_ctx* c, int kind);
fe_status learn_after_solve(fe_ctx* c, int kind);
// Forget the per-particle pressure history (after uploads or out-of-sequence sorts).
void learn_invalidate(fe_ctx* c);
void learn_free(fe_ctx* c);
// Whether a network is loaded for solve `kind` / for either solve.
bool learn_ready(const fe_ctx* c, int kind);

}  // namespace fe

#define FE_HIP(call, what)                                    \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return ::fe::hip_fail(e_, what); \
    } while (0)

#define FE_CHECK_CTX(c)                                                         \
    do {                                                                        \
        if (!(c)) return ::fe::fail(FE_ERR_INVALID_ARGUMENT, "null fe_ctx");    \
    } while (0)
