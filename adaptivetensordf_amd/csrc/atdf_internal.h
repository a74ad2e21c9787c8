// atdf_internal.h -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/atdf.h"

namespace atdf {

void set_error(const char* fmt, ...);

inline hipStream_t to_hip(atdf_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Returns ATDF_ERR_HIP with a message if the last launch failed.
inline atdf_status check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return ATDF_ERR_HIP;
    }
    return ATDF_OK;
}

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace atdf

#define ATDF_REQUIRE(cond, ...)                 \
    do {                                        \
        if (!(cond)) {                          \
            ::atdf::set_error(__VA_ARGS__);     \
            return ATDF_ERR_INVALID_ARGUMENT;   \
        }                                       \
    } while (0)
