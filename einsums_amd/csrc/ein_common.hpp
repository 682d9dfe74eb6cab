// Shared helpers for the gfx950 einsum backend: status/error plumbing, per-stream
// workspace, fast integer division for index decoding, and launch-shape helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "einsums_amd.h"

namespace ein {

// ------------------------------------------------------------------ errors
void set_error(std::string const &msg);
int fail(int code, std::string const &msg);

#define EIN_HIP_CHECK(expr)                                                                                            \
    do {                                                                                                               \
        hipError_t _e = (expr);                                                                                        \
        if (_e != hipSuccess) {                                                                                        \
            return ::ein::fail(EIN_ERROR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));                      \
        }                                                                                                              \
    } while (0)

#define EIN_LAUNCH_CHECK()                                                                                             \
    do {                                                                                                               \
        hipError_t _e = hipGetLastError();                                                                             \
        if (_e != hipSuccess) {                                                                                        \
            return ::ein::fail(EIN_ERROR_HIP, std::string("kernel launch: ") + hipGetErrorString(_e));                 \
        }                                                                                                              \
    } while (0)

#define EIN_TRY(expr)                                                                                                  \
    do {                                                                                                               \
        int _s = (expr);                                                                                               \
        if (_s != EIN_SUCCESS) return _s;                                                                              \
    } while (0)

// --------------------------------------------------------------- workspace
// Returns a device buffer of at least `bytes` bytes private to `stream`. Grows with
// hipMalloc outside stream capture; inside capture it fails with EIN_ERROR_WORKSPACE
// unless ein_workspace_reserve() sized it beforehand.
int workspace(hipStream_t stream, size_t bytes, void **out);
// Same for a numbered buffer of the stream (slot 0 is the one above); a call that keeps
// data in slot 1 can run nested calls that use slot 0.
int workspace_slot(hipStream_t stream, int slot, size_t bytes, void **out);
// Whether `stream` is being captured into a graph.
bool stream_capturing(hipStream_t stream);
// Whether slot `slot` of `stream` already holds `bytes` (no allocation needed).
bool workspace_fits(hipStream_t stream, int slot, size_t bytes);

inline hipStream_t as_stream(ein_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Number of compute units on the current device (cached).
int num_cus();

// Workgroups of `kernel` (launched with `threads` threads, no dynamic LDS) resident per
// CU at once, from the occupancy calculator (cached per kernel; at least 1).
int resident_blocks(void const *kernel, int threads);

// ------------------------------------------------------ fast 32-bit divmod
// Division by an invariant divisor d >= 1 for numerators n < 2^31 using the
// round-up multiply-shift method (Granlund & Montgomery, PLDI 1994).
struct FastDiv {
    uint32_t d;
    uint32_t mul;
    uint32_t shift;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((uint64_t{1} << l) < d) ++l;
    f.shift = l;
    // mul = floor(2^32 * (2^l - d) / d) + 1
    f.mul = static_cast<uint32_t>(((uint64_t{1} << 32) * ((uint64_t{1} << l) - d)) / d + 1);
    return f;
}

__host__ __device__ __forceinline__ uint32_t fast_div(uint32_t n, FastDiv const &f) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t t = __umulhi(n, f.mul);
#else
    uint32_t t = static_cast<uint32_t>((uint64_t{n} * f.mul) >> 32);
#endif
    return (t + n) >> f.shift;
}

// ---------------------------------------------------- small device helpers
template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Maps a linear workgroup id to one that keeps consecutive logical tiles on the same
// XCD: the hardware dispatches workgroup ids round-robin over the 8 XCDs, so the
// XCD of id w is w % 8. The returned id enumerates all ids of XCD 0 first, then XCD 1...
__host__ __device__ __forceinline__ uint32_t xcd_remap(uint32_t wg, uint32_t nwg) {
    constexpr uint32_t NXCD = 8;
    uint32_t xcd  = wg % NXCD;
    uint32_t idx  = wg / NXCD;
    uint32_t base = nwg / NXCD;
    uint32_t rem  = nwg % NXCD;
    return xcd * base + (xcd < rem ? xcd : rem) + idx;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

} // namespace ein
