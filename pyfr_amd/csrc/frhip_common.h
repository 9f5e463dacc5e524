// Shared helpers for the frhip library (internal header).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/frhip.h"

namespace frhip {

void set_error(const std::string &msg);

// Return a nonzero status with a message if a HIP call failed.
#define FR_HIP(call)                                                        \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) {                                             \
            frhip::set_error(std::string(#call) + ": " +                    \
                             hipGetErrorString(_e));                        \
            return 1;                                                       \
        }                                                                   \
    } while (0)

#define FR_CHECK(cond, msg)                                                 \
    do {                                                                    \
        if (!(cond)) {                                                      \
            frhip::set_error(msg);                                          \
            return 2;                                                       \
        }                                                                   \
    } while (0)

// Launch-error check after a kernel launch.
#define FR_LAUNCHED()                                                       \
    do {                                                                    \
        hipError_t _e = hipGetLastError();                                  \
        if (_e != hipSuccess) {                                             \
            frhip::set_error(std::string("kernel launch: ") +               \
                             hipGetErrorString(_e));                        \
            return 3;                                                       \
        }                                                                   \
    } while (0)

// Catch-all so that no C++ exception crosses the ABI.
#define FR_API_BEGIN try {
#define FR_API_END                                                          \
    } catch (const std::exception &ex) {                                    \
        frhip::set_error(std::string("exception: ") + ex.what());           \
        return 9;                                                           \
    } catch (...) {                                                         \
        frhip::set_error("unknown exception");                              \
        return 9;                                                           \
    }

constexpr int SOA = 64;  // SoA width = gfx950 wavefront size

// Index of (element e, variable v) within one row of an AoSoA array.
__host__ __device__ inline long soa_ix(long e, int v, int nvar)
{
    return ((e >> 6) * nvar + v) * SOA + (e & 63);
}

inline hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

inline unsigned int cdiv(long a, long b) { return (unsigned int)((a + b - 1) / b); }

// Per-device address of the workspace-free reduction's scratch array
// (frhip_blasext.hip); frhip_init looks it up for its device.  Returns a
// hipError_t.
int red_scratch_init(int device);

}  // namespace frhip
