// Shared host/device helpers for the gfx950 D3Q19 BGK kernels.
//
// Error policy (mirrors the reference, SRC/col_MC_gpu.hpp:34-47 and
// SRC/bcond_gpu.hpp:548-571): a launch error is reported on stderr right
// after the launch without synchronising; an asynchronous fault surfaces at
// the next device barrier, which aborts.  Entry points of the handle API
// additionally return an error code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace bgk {

// Stream used by the reference-compatible entry points.  The reference
// launches everything on the null stream so that boundary kernels and the
// collision serialise in issue order (SRC/bcond_gpu.hpp:26-30); we keep that
// default and let a host select another in-order stream.
hipStream_t& abi_stream();

inline void report_launch(const char* name)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::fprintf(stderr, "bgk3d_gpu: launch of %s failed: %s\n", name, hipGetErrorString(e));
        std::fflush(stderr);
    }
}

inline int ceil_div(long a, long b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace bgk

#define BGK_HIP_CHECK(call)                                                                        \
    do {                                                                                           \
        hipError_t _e = (call);                                                                    \
        if (_e != hipSuccess) {                                                                    \
            std::fprintf(stderr, "bgk3d_gpu: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(_e), \
                         __FILE__, __LINE__);                                                      \
            std::fflush(stderr);                                                                   \
            return static_cast<int>(_e);                                                           \
        }                                                                                          \
    } while (0)
