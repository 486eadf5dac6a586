// Internal helpers shared by the libdsx HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/dsx.h"

namespace dsx {

// Record the calling thread's last error message and return `code`.
int set_error(int code, const char *fmt, ...);

// Check the most recent launch; returns DSX_OK or DSX_ELAUNCH with the HIP error text.
int check_launch(const char *what);

#if defined(__HIPCC__)
// device-side helpers (host-only translation units built by a plain C++
// compiler, e.g. the planner under sanitizers, skip them)
typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float ld_as_float(const float *p, int64_t i) { return p[i]; }
__device__ __forceinline__ float ld_as_float(const _Float16 *p, int64_t i) { return (float)p[i]; }

__device__ __forceinline__ void st_from_float(float *p, int64_t i, float v) { p[i] = v; }
__device__ __forceinline__ void st_from_float(_Float16 *p, int64_t i, float v) { p[i] = (_Float16)v; }
#endif

inline hipStream_t as_stream(dsx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace dsx

#define DSX_REQUIRE(cond, ...)                                   \
    do {                                                         \
        if (!(cond)) return ::dsx::set_error(DSX_EINVAL, __VA_ARGS__); \
    } while (0)
