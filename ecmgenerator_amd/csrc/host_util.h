// host_util.h -- error reporting and device buffers shared by the host
// runtime translation units (sim_api.cpp, mlp_api.cpp).
#pragma once

#include "../../include/ecm_sim.h"

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ecm {

// Records the message returned by ecm_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        if (hipMalloc((void**)&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return fail(ECM_E_HIP, "hipMalloc of %zu bytes failed", count * sizeof(T));
        }
        n = count;
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace ecm

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return ecm::fail(ECM_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)
