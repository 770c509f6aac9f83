// api_common.hpp -- error reporting and launch helpers shared by the C-ABI
// translation units of libbasinhip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/basinhip.h"

namespace bhapi {

// message of the last failed call on this thread (bh_last_error)
extern thread_local char g_err[512];

inline int32_t fail(int32_t code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define BH_HIP_CHECK(expr)                                                              \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return bhapi::fail(-6, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

// resident lanes that fill the device for a kernel, from its occupancy
template <class K>
int32_t device_lanes(K kernel, int block, int64_t *lanes) {
    int dev = 0, cus = 0, per_cu = 0;
    BH_HIP_CHECK(hipGetDevice(&dev));
    BH_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
    if (per_cu < 1) per_cu = 1;
    *lanes = (int64_t)cus * per_cu * block;
    return 0;
}

}  // namespace bhapi
