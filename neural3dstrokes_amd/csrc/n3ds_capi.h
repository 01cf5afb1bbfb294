// Internal helpers shared by the C-ABI entry points: error reporting and launch checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "n3ds.h"  // public declarations: every definition below must match them

namespace n3ds {

void set_error(const char* fmt, ...);

}  // namespace n3ds

#define N3DS_EXPORT extern "C" __attribute__((visibility("default")))

// status codes returned by every entry point
#define N3DS_OK 0
#define N3DS_EINVAL 1
#define N3DS_EHIP 2

#define N3DS_REQUIRE(cond, ...)                  \
    do {                                         \
        if (!(cond)) {                           \
            n3ds::set_error(__VA_ARGS__);        \
            return N3DS_EINVAL;                  \
        }                                        \
    } while (0)

#define N3DS_HIP_CHECK(expr)                                                              \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            n3ds::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return N3DS_EHIP;                                                             \
        }                                                                                 \
    } while (0)

#define N3DS_LAUNCH_CHECK() N3DS_HIP_CHECK(hipGetLastError())
