// Host-side helpers of the device set-up stages (pattern_dev.hip,
// setup_dev.hip): HIP error checks that return PG_E_HIP with a message, a
// pool of device buffers freed together, device prefix sums and downloads.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "pgfem_asm.h"

#define DEV_OK(x)                                                            \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) {                                              \
            err = std::string(#x) + ": " + hipGetErrorString(e_);            \
            return PG_E_HIP;                                                 \
        }                                                                    \
    } while (0)

namespace pgh {

// Device buffers freed together at the end of a build.
struct Pool {
    std::vector<void*> p;
    ~Pool() {
        for (void* q : p) (void)hipFree(q);
    }
    template <typename T>
    hipError_t get(T** out, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(q);
        *out = static_cast<T*>(q);
        return e;
    }
};

inline unsigned blocks(int64_t n, int t = 256) { return unsigned((n + t - 1) / t); }

// exclusive prefix of cnt[0 .. n) into ptr[0 .. n] (ptr[0] = 0), on the device
template <typename TI, typename TO>
int prefix(Pool& pool, const TI* cnt, TO* ptr, int64_t n, std::string& err) {
    DEV_OK(hipMemset(ptr, 0, sizeof(TO)));
    if (n == 0) return PG_OK;
    size_t tb = 0;
    DEV_OK(hipcub::DeviceScan::InclusiveSum(nullptr, tb, cnt, ptr + 1, int(n)));
    char* tmp = nullptr;
    DEV_OK(pool.get(&tmp, tb));
    DEV_OK(hipcub::DeviceScan::InclusiveSum(tmp, tb, cnt, ptr + 1, int(n)));
    return PG_OK;
}

template <typename T>
int down(std::vector<T>& v, const T* src, int64_t n, std::string& err) {
    v.resize(size_t(n));
    if (n) DEV_OK(hipMemcpy(v.data(), src, size_t(n) * sizeof(T), hipMemcpyDeviceToHost));
    return PG_OK;
}

}  // namespace pgh
