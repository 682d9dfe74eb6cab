// Bandwidth-bound leaves for gfx950: gemv, dot, ger, direct product, scale, axpby and
// whole-view tensor scaling. All reductions are deterministic (fixed-shape trees, no
// atomics) so repeated calls give bit-identical results.
//
// Reference leaves replaced (column-major BLAS semantics kept):
//   blas::gpu::gemv  -> hipblasDgemv_64   (hipBLASVendor/src/gemv.hip:65-113)
//   blas::gpu::dot   -> hipblasDdot_64    (hipBLASVendor/src/dot.hip:29-108) and the
//                       atomic dot_kernel  (LinearAlgebra/GPULinearAlgebra.hpp:36-100)
//   blas::gpu::ger   -> hipblasDger_64    (hipBLASVendor/src/ger.hip:46-119)
//   blas::gpu::dirprod -> ddirprod_kernel (hipBLASVendor/src/direct_prod.hip:20-271)
//   blas::gpu::scal / axpby               (hipBLASVendor/src/scal.hip:25-68)
#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>

#include "blas1.hpp"

namespace ein {

// ------------------------------------------------------------------- gemv
// y[i] (+)= alpha * sum_j A[i + j*lda] x[j]; block = 64 rows x 4 waves over j.
template <typename T>
__global__ __launch_bounds__(256) void gemv_n_kernel(int64_t m, int64_t n, T alpha, T const *__restrict__ A, int64_t lda,
                                                     T const *__restrict__ x, int64_t incx, T beta, T *__restrict__ y,
                                                     int64_t incy, int64_t jchunk, T *__restrict__ partial) {
    __shared__ T red[4][64];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int64_t const i = int64_t(blockIdx.x) * 64 + lane;
    int64_t const jb = int64_t(blockIdx.y) * jchunk;
    int64_t const je = std::min<int64_t>(n, jb + jchunk);
    T acc = T(0);
    if (i < m) {
        for (int64_t j = jb + wave; j < je; j += 4) acc += A[i + j * lda] * x[j * incx];
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && i < m) {
        T s = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
        if (partial != nullptr) {
            partial[int64_t(blockIdx.y) * m + i] = s;
        } else {
            T v = alpha * s;
            if (beta != T(0)) v += beta * y[i * incy];
            y[i * incy] = v;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void gemv_n_finish(int64_t m, int splits, T alpha, T const *__restrict__ partial,
                                                     T beta, T *__restrict__ y, int64_t incy) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    T s = T(0);
    for (int k = 0; k < splits; ++k) s += partial[int64_t(k) * m + i];
    T v = alpha * s;
    if (beta != T(0)) v += beta * y[i * incy];
    y[i * incy] = v;
}

// y[j] (+)= alpha * sum_i A[i + j*lda] x[i]. A block is four waves. Column j is cut into
// splits*wpc contiguous segments, one per wave: the wpc waves of a block that share a column
// take consecutive segments and add their sums in LDS in fixed order, and with splits > 1
// each block row writes a partial [split][j] that gemv_n_finish adds in fixed order. Cutting
// columns keeps many short waves in flight, so the grid has no long low-occupancy tail (one
// wave per column on a 9216 x 9216 matrix is 1.1 rounds of resident waves). With unit incx,
// lda a multiple of the vector width and 16-byte aligned A and x: 16-byte loads, four of
// each operand in flight per lane.
template <typename T>
__global__ __launch_bounds__(256) void gemv_t_kernel(int64_t m, int64_t n, T alpha, T const *__restrict__ A, int64_t lda,
                                                     T const *__restrict__ x, int64_t incx, T beta, T *__restrict__ y,
                                                     int64_t incy, int wpc, int splits, T *__restrict__ partial) {
    __shared__ T red[4];
    int const lane   = threadIdx.x & 63;
    int const wave   = threadIdx.x >> 6;
    int64_t const j  = int64_t(blockIdx.x) * (4 / wpc) + wave / wpc;
    int const seg    = int(blockIdx.y) * wpc + wave % wpc;
    int const nseg   = splits * wpc;
    T acc[4]         = {T(0), T(0), T(0), T(0)};
    constexpr int VW = 16 / int(sizeof(T));
    bool const vec   = incx == 1 && lda % VW == 0 &&
                     (reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(x)) % 16 == 0;
    if (j < n) {
        T const *col = A + j * lda;
        if (vec) {
            using V = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
            V const *cv = reinterpret_cast<V const *>(col);
            V const *xv = reinterpret_cast<V const *>(x);
            int64_t const mv    = m / VW;
            int64_t const chunk = (mv + nseg - 1) / nseg;
            int64_t const v1    = std::min<int64_t>(mv, (seg + 1) * chunk);
            auto dotv           = [](V const &a, V const &b) -> T {
                if constexpr (sizeof(T) == 8) return a.x * b.x + a.y * b.y;
                else return (a.x * b.x + a.y * b.y) + (a.z * b.z + a.w * b.w);
            };
            int64_t i = seg * chunk + lane;
            for (; i + 192 < v1; i += 256) {
                V a[4], b[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    a[u] = cv[i + 64 * u];
                    b[u] = xv[i + 64 * u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) acc[u] += dotv(a[u], b[u]);
            }
            for (; i < v1; i += 64) acc[0] += dotv(cv[i], xv[i]);
            int64_t const t = mv * VW + lane;
            if (seg == nseg - 1 && t < m) acc[1] += col[t] * x[t];
        } else {
            int64_t const chunk = (m + nseg - 1) / nseg;
            int64_t const e1    = std::min<int64_t>(m, (seg + 1) * chunk);
            int64_t i           = seg * chunk + lane;
            for (; i + 64 < e1; i += 128) {
                acc[0] += col[i] * x[i * incx];
                acc[1] += col[i + 64] * x[(i + 64) * incx];
            }
            if (i < e1) acc[0] += col[i] * x[i * incx];
        }
    }
    T s = wave_reduce_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
    if (wpc > 1) {
        if (lane == 0) red[wave] = s;
        __syncthreads();
        if (wave % wpc != 0) return;
        for (int u = 1; u < wpc; ++u) s += red[wave + u];
    }
    if (j >= n || lane != 0) return;
    if (splits > 1) {
        partial[int64_t(blockIdx.y) * n + j] = s;
        return;
    }
    T v = alpha * s;
    if (beta != T(0)) v += beta * y[j * incy];
    y[j * incy] = v;
}

template <typename T>
int gemv_impl(char trans, int64_t m, int64_t n, T alpha, T const *A, int64_t lda, T const *x, int64_t incx, T beta, T *y,
              int64_t incy, hipStream_t stream) {
    bool const notrans = trans == 'N' || trans == 'n';
    if (!notrans && !(trans == 'T' || trans == 't' || trans == 'C' || trans == 'c')) return fail(-1, "gemv: invalid trans");
    if (m < 0) return fail(-2, "gemv: m < 0");
    if (n < 0) return fail(-3, "gemv: n < 0");
    if (lda < std::max<int64_t>(1, m)) return fail(-6, "gemv: lda too small");
    if (incx == 0) return fail(-8, "gemv: incx == 0");
    if (incy == 0) return fail(-11, "gemv: incy == 0");
    int64_t const leny = notrans ? m : n;
    int64_t const lenx = notrans ? n : m;
    if (leny == 0) return EIN_SUCCESS;
    // negative increments address from the far end, as in BLAS
    if (incx < 0) x = x - (lenx - 1) * incx;
    if (incy < 0) y = y - (leny - 1) * incy;
    if (lenx == 0 || alpha == T(0)) return scale_strided<T>(leny, beta, y, incy, stream);
    if (notrans) {
        int64_t rb     = ceil_div(m, 64);
        int64_t target = int64_t(num_cus()) * 4;
        int splits     = int(std::max<int64_t>(1, std::min<int64_t>(ceil_div(target, rb), ceil_div(n, 64))));
        splits         = std::min(splits, 1024);
        int64_t jchunk = ceil_div(n, splits);
        splits         = int(ceil_div(n, jchunk));
        T *partial     = nullptr;
        if (splits > 1) {
            void *ws = nullptr;
            EIN_TRY(workspace(stream, size_t(splits) * size_t(m) * sizeof(T), &ws));
            partial = static_cast<T *>(ws);
        }
        dim3 grid{uint32_t(rb), uint32_t(splits)};
        gemv_n_kernel<T><<<grid, 256, 0, stream>>>(m, n, alpha, A, lda, x, incx, beta, y, incy, jchunk, partial);
        EIN_LAUNCH_CHECK();
        if (splits > 1) {
            gemv_n_finish<T><<<uint32_t(ceil_div(m, 256)), 256, 0, stream>>>(m, splits, alpha, partial, beta, y, incy);
            EIN_LAUNCH_CHECK();
        }
    } else {
        // waves per column (1, 2, 4) until about four rounds of resident waves, keeping
        // >= 256 elements per wave; then split across block rows when even that is short
        int64_t const resident = int64_t(num_cus()) * 32;
        int wpc                = 1;
        while (wpc < 4 && n * wpc < 4 * resident && m / (2 * wpc) >= 256) wpc *= 2;
        int splits = 1;
        if (wpc == 4 && n * 4 < resident)
            splits = int(std::clamp<int64_t>(std::min<int64_t>(ceil_div(4 * resident, n * 4), m / 4096), 1, 1024));
        T *partial = nullptr;
        if (splits > 1) {
            void *ws = nullptr;
            EIN_TRY(workspace(stream, size_t(splits) * size_t(n) * sizeof(T), &ws));
            partial = static_cast<T *>(ws);
        }
        dim3 grid{uint32_t(ceil_div(n, 4 / wpc)), uint32_t(splits)};
        gemv_t_kernel<T><<<grid, 256, 0, stream>>>(m, n, alpha, A, lda, x, incx, beta, y, incy, wpc, splits, partial);
        EIN_LAUNCH_CHECK();
        if (splits > 1) {
            gemv_n_finish<T><<<uint32_t(ceil_div(n, 256)), 256, 0, stream>>>(n, splits, alpha, partial, beta, y, incy);
            EIN_LAUNCH_CHECK();
        }
    }
    return EIN_SUCCESS;
}

// -------------------------------------------------------------------- dot
// Unit strides: workgroup b sums the contiguous elements [b*DOT_CHUNK, (b+1)*DOT_CHUNK), so
// consecutive workgroups stream consecutive memory (6.3 TB/s measured on 2.4 GB against
// 5.3 TB/s for a grid-stride loop), with 16-byte loads when both operands are aligned.
// Other strides: a grid-stride loop. A second launch sums the partials in a fixed order
// (deterministic). One launch whose last-arriving workgroup sums the partials was measured
// slower on MI355X (85 MB: 17.8 us for the two launches against 20.6 us with write-through
// partials and 32-35 us with agent-scope release/acquire counting), because making the
// partials visible across XCDs costs more than the second launch.
constexpr int DOT_CHUNK = 256 * 32; // elements per workgroup: 64 KB of each FP64 operand

template <typename T>
__global__ __launch_bounds__(256) void dot_partial_kernel(int64_t n, T const *__restrict__ x, int64_t incx,
                                                          T const *__restrict__ y, int64_t incy, T *__restrict__ partial) {
    __shared__ T red[4];
    T acc[4] = {T(0), T(0), T(0), T(0)};
    if (incx == 1 && incy == 1) {
        int64_t const b0 = int64_t(blockIdx.x) * DOT_CHUNK;
        int64_t const b1 = b0 + DOT_CHUNK < n ? b0 + DOT_CHUNK : n;
        bool const vec   = (reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) % 16 == 0;
        if (vec && b1 - b0 == DOT_CHUNK) {
            constexpr int VW = 16 / int(sizeof(T));
            using V          = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
            V const *xv = reinterpret_cast<V const *>(x + b0);
            V const *yv = reinterpret_cast<V const *>(y + b0);
            constexpr int NV = DOT_CHUNK / VW / 256; // vectors per thread
#pragma unroll
            for (int g = 0; g < NV; g += 8) {
                V a[8], c[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    a[u] = xv[(g + u) * 256 + threadIdx.x];
                    c[u] = yv[(g + u) * 256 + threadIdx.x];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    T d;
                    if constexpr (sizeof(T) == 8) d = a[u].x * c[u].x + a[u].y * c[u].y;
                    else d = (a[u].x * c[u].x + a[u].y * c[u].y) + (a[u].z * c[u].z + a[u].w * c[u].w);
                    acc[u & 3] += d;
                }
            }
        } else {
            for (int64_t e = b0 + threadIdx.x; e < b1; e += 256) acc[0] += x[e] * y[e];
        }
    } else {
        int64_t const stride = int64_t(gridDim.x) * blockDim.x;
        for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += stride) acc[0] += x[e * incx] * y[e * incy];
    }
    T s = wave_reduce_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Sums the partials in a fixed order: 1024 threads, eight independent loads in flight per
// thread (tens of thousands of partials for a 2.4 GB dot), then result = alpha*sum + beta*result.
template <typename T>
__global__ __launch_bounds__(1024) void dot_finish_kernel(int parts, T const *__restrict__ partial, T alpha, T beta,
                                                          T *__restrict__ result) {
    __shared__ T red[16];
    T acc[8] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    int k = threadIdx.x;
    for (; k + 7 * 1024 < parts; k += 8 * 1024) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = partial[k + u * 1024];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += v[u];
    }
    for (; k < parts; k += 1024) acc[0] += partial[k];
    T s = wave_reduce_sum(((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7])));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        T tot = T(0);
        for (int w = 0; w < 16; ++w) tot += red[w];
        T v = alpha * tot;
        if (beta != T(0)) v += beta * result[0];
        result[0] = v;
    }
}

// Workgroups of a dot over n elements (the number of partials it writes).
inline int64_t dot_parts(int64_t n, int64_t incx, int64_t incy) {
    int64_t const chunks = ceil_div(n, DOT_CHUNK);
    return incx == 1 && incy == 1 ? std::max<int64_t>(1, chunks)
                                  : std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256 * 8), int64_t(num_cus()) * 4));
}

// Both launches of a dot whose (already normalised, non-negative-stride) operands are given;
// `partial` holds at least dot_parts(n, incx, incy) elements.
template <typename T>
int dot_launch(int64_t n, T alpha, T const *x, int64_t incx, T const *y, int64_t incy, T beta, T *result, T *partial,
               int64_t parts, hipStream_t stream) {
    dot_partial_kernel<T><<<uint32_t(parts), 256, 0, stream>>>(n, x, incx, y, incy, partial);
    EIN_LAUNCH_CHECK();
    dot_finish_kernel<T><<<1, 1024, 0, stream>>>(int(parts), partial, alpha, beta, result);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

int dot_check(int64_t n, void const *result, int64_t incx, int64_t incy) {
    if (n < 0) return fail(-1, "dot: n < 0");
    if (result == nullptr) return fail(-8, "dot: null result");
    if (incx == 1 && incy == 1 && ceil_div(n, DOT_CHUNK) > (int64_t{1} << 30)) return fail(EIN_ERROR_UNSUPPORTED, "dot: vector too long");
    return EIN_SUCCESS;
}

template <typename T>
int dot_impl(int64_t n, T alpha, T const *x, int64_t incx, T const *y, int64_t incy, T beta, T *result,
             hipStream_t stream) {
    EIN_TRY(dot_check(n, result, incx, incy));
    if (n > 0 && incx < 0) x = x - (n - 1) * incx;
    if (n > 0 && incy < 0) y = y - (n - 1) * incy;
    int64_t const parts = dot_parts(n, incx, incy);
    void *ws            = nullptr;
    EIN_TRY(workspace(stream, size_t(parts) * sizeof(T), &ws));
    return dot_launch<T>(n, alpha, x, incx, y, incy, beta, result, static_cast<T *>(ws), parts, stream);
}

// ------------------------------------------------------- allocation-free dot
// ein_dot_f64_noalloc keeps its partial sums in this file-scope device array instead of
// the per-stream workspace, so it never allocates. Every GPU holds its own copy of the
// array at its own address; host code must use the address hipGetSymbolAddress returns
// for the current device (the array's name is not a device address on the host). Calls
// on one GPU share the array: calls that overlap in time on one GPU race. Nothing in it
// needs to start at zero (each call writes every partial it reads).
static __device__ __align__(256) double scratch[EIN_DOT_NOALLOC_MAX_PARTS];

namespace {
std::mutex g_scratch_mutex;
constexpr int MAX_DEVICES = 64;
double *g_scratch_ptr[MAX_DEVICES] = {};
} // namespace

// The current device's address of `scratch` (looked up once per device).
int noalloc_scratch(double **out) {
    int dev = 0;
    EIN_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(EIN_ERROR_UNSUPPORTED, "dot_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            return fail(EIN_ERROR_HIP, std::string("dot_noalloc: hipGetSymbolAddress of the scratch array failed: ") +
                                           hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return EIN_SUCCESS;
}

// ein_dot_f64 with the partials in `scratch`: same kernels, so the same bits.
int dot_noalloc_f64(int64_t n, double alpha, double const *x, int64_t incx, double const *y, int64_t incy, double beta,
                    double *result, hipStream_t stream) {
    EIN_TRY(dot_check(n, result, incx, incy));
    int64_t const parts = dot_parts(n, incx, incy);
    if (parts > EIN_DOT_NOALLOC_MAX_PARTS)
        return fail(EIN_ERROR_UNSUPPORTED, "dot_noalloc: n = " + std::to_string(n) + " needs " + std::to_string(parts) +
                                               " partial sums, more than the " +
                                               std::to_string(EIN_DOT_NOALLOC_MAX_PARTS) + " of its static scratch");
    double *partial = nullptr;
    EIN_TRY(noalloc_scratch(&partial));
    if (n > 0 && incx < 0) x = x - (n - 1) * incx;
    if (n > 0 && incy < 0) y = y - (n - 1) * incy;
    return dot_launch<double>(n, alpha, x, incx, y, incy, beta, result, partial, parts, stream);
}

// -------------------------------------------------------------------- ger
template <typename T>
__global__ __launch_bounds__(256) void ger_kernel(int64_t m, int64_t n, T alpha, T const *__restrict__ x, int64_t incx,
                                                  T const *__restrict__ y, int64_t incy, T *__restrict__ A, int64_t lda) {
    int64_t const i = int64_t(blockIdx.x) * 64 + (threadIdx.x & 63);
    if (i >= m) return;
    T const xi = alpha * x[i * incx];
    for (int64_t j = int64_t(blockIdx.y) * 4 + (threadIdx.x >> 6); j < n; j += int64_t(gridDim.y) * 4)
        A[i + j * lda] += xi * y[j * incy];
}

template <typename T>
int ger_impl(int64_t m, int64_t n, T alpha, T const *x, int64_t incx, T const *y, int64_t incy, T *A, int64_t lda,
             hipStream_t stream) {
    if (m < 0) return fail(-1, "ger: m < 0");
    if (n < 0) return fail(-2, "ger: n < 0");
    if (incx == 0) return fail(-5, "ger: incx == 0");
    if (incy == 0) return fail(-7, "ger: incy == 0");
    if (lda < std::max<int64_t>(1, m)) return fail(-9, "ger: lda too small");
    if (m == 0 || n == 0 || alpha == T(0)) return EIN_SUCCESS;
    if (incx < 0) x = x - (m - 1) * incx;
    if (incy < 0) y = y - (n - 1) * incy;
    int64_t rb = ceil_div(m, 64);
    int64_t cb = std::min<int64_t>(ceil_div(n, 4), std::max<int64_t>(1, int64_t(num_cus()) * 8 / rb));
    cb         = std::min<int64_t>(cb, 65535);
    ger_kernel<T><<<dim3(uint32_t(rb), uint32_t(cb)), 256, 0, stream>>>(m, n, alpha, x, incx, y, incy, A, lda);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

// ------------------------------------------------- dirprod / scale / axpby
template <typename T>
__global__ __launch_bounds__(256) void dirprod_kernel(int64_t n, T alpha, T const *__restrict__ x, int64_t incx,
                                                      T const *__restrict__ y, int64_t incy, T beta, T *__restrict__ z,
                                                      int64_t incz) {
    int64_t const stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += stride) {
        T v = alpha * (x[e * incx] * y[e * incy]);
        if (beta != T(0)) v += beta * z[e * incz];
        z[e * incz] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(int64_t n, T alpha, T *__restrict__ x, int64_t incx) {
    int64_t const stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += stride)
        x[e * incx] = alpha == T(0) ? T(0) : alpha * x[e * incx];
}

template <typename T>
__global__ __launch_bounds__(256) void axpby_kernel(int64_t n, T alpha, T const *__restrict__ x, int64_t incx, T beta,
                                                    T *__restrict__ y, int64_t incy) {
    int64_t const stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += stride) {
        T v = alpha * x[e * incx];
        if (beta != T(0)) v += beta * y[e * incy];
        y[e * incy] = v;
    }
}

inline uint32_t stream_blocks(int64_t n) {
    int64_t b = ceil_div(n, 256);
    return uint32_t(std::max<int64_t>(1, std::min<int64_t>(b, int64_t(num_cus()) * 16)));
}

template <typename T>
int scale_strided(int64_t n, T alpha, T *x, int64_t incx, hipStream_t stream) {
    if (n <= 0 || alpha == T(1)) return EIN_SUCCESS;
    if (incx < 0) x = x - (n - 1) * incx;
    scale_kernel<T><<<stream_blocks(n), 256, 0, stream>>>(n, alpha, x, incx);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

template <typename T>
int dirprod_impl(int64_t n, T alpha, T const *x, int64_t incx, T const *y, int64_t incy, T beta, T *z, int64_t incz,
                 hipStream_t stream) {
    if (n < 0) return fail(-1, "dirprod: n < 0");
    if (n == 0) return EIN_SUCCESS;
    if (incx < 0) x = x - (n - 1) * incx;
    if (incy < 0) y = y - (n - 1) * incy;
    if (incz < 0) z = z - (n - 1) * incz;
    dirprod_kernel<T><<<stream_blocks(n), 256, 0, stream>>>(n, alpha, x, incx, y, incy, beta, z, incz);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

template <typename T>
int axpby_impl(int64_t n, T alpha, T const *x, int64_t incx, T beta, T *y, int64_t incy, hipStream_t stream) {
    if (n < 0) return fail(-1, "axpby: n < 0");
    if (n == 0) return EIN_SUCCESS;
    if (incx < 0) x = x - (n - 1) * incx;
    if (incy < 0) y = y - (n - 1) * incy;
    axpby_kernel<T><<<stream_blocks(n), 256, 0, stream>>>(n, alpha, x, incx, beta, y, incy);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

// --------------------------------------------------- whole-view tensor scale
template <typename T>
__global__ __launch_bounds__(256) void tensor_scale_kernel(TensorShape S, T alpha, T *__restrict__ c) {
    int64_t const stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < S.total; e += stride) {
        int64_t off = 0, rem = e;
        for (int d = S.rank - 1; d >= 0; --d) {
            int64_t q = rem / S.dims[d];
            off += (rem - q * S.dims[d]) * S.strides[d];
            rem = q;
        }
        c[off] = alpha == T(0) ? T(0) : alpha * c[off];
    }
}

template <typename T>
int tensor_scale(ein_tensor_desc const &C, T *c, T alpha, hipStream_t stream) {
    if (alpha == T(1)) return EIN_SUCCESS;
    TensorShape S{};
    EIN_TRY(make_shape(C, S));
    if (S.total == 0) return EIN_SUCCESS;
    if (S.rank <= 1) return scale_strided<T>(S.total, alpha, c, S.rank == 0 ? 1 : S.strides[0], stream);
    tensor_scale_kernel<T><<<stream_blocks(S.total), 256, 0, stream>>>(S, alpha, c);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

// Collapses a descriptor to its distinct axes, merging axes that are contiguous with
// the next one (row-major sense) so the decode loop is as short as possible.
int make_shape(ein_tensor_desc const &C, TensorShape &S) {
    if (C.rank < 0 || C.rank > EIN_MAX_RANK) return fail(EIN_ERROR_RANK, "tensor rank out of range");
    S.rank  = 0;
    S.total = 1;
    for (int d = 0; d < C.rank; ++d) {
        if (C.dims[d] < 0) return fail(EIN_ERROR_VALUE, "negative tensor extent");
        S.total *= C.dims[d];
        if (C.dims[d] == 1) continue;
        if (S.rank > 0 && S.strides[S.rank - 1] == C.strides[d] * C.dims[d]) {
            S.dims[S.rank - 1] *= C.dims[d];
            S.strides[S.rank - 1] = C.strides[d];
        } else {
            S.dims[S.rank]    = C.dims[d];
            S.strides[S.rank] = C.strides[d];
            ++S.rank;
        }
    }
    if (S.rank == 0 && S.total == 1) {
        S.rank       = 1;
        S.dims[0]    = 1;
        S.strides[0] = 1;
    }
    return EIN_SUCCESS;
}

template int gemv_impl<double>(char, int64_t, int64_t, double, double const *, int64_t, double const *, int64_t, double,
                               double *, int64_t, hipStream_t);
template int gemv_impl<float>(char, int64_t, int64_t, float, float const *, int64_t, float const *, int64_t, float,
                              float *, int64_t, hipStream_t);
template int dot_impl<double>(int64_t, double, double const *, int64_t, double const *, int64_t, double, double *,
                              hipStream_t);
template int dot_impl<float>(int64_t, float, float const *, int64_t, float const *, int64_t, float, float *, hipStream_t);
template int ger_impl<double>(int64_t, int64_t, double, double const *, int64_t, double const *, int64_t, double *,
                              int64_t, hipStream_t);
template int ger_impl<float>(int64_t, int64_t, float, float const *, int64_t, float const *, int64_t, float *, int64_t,
                             hipStream_t);
template int dirprod_impl<double>(int64_t, double, double const *, int64_t, double const *, int64_t, double, double *,
                                  int64_t, hipStream_t);
template int dirprod_impl<float>(int64_t, float, float const *, int64_t, float const *, int64_t, float, float *, int64_t,
                                 hipStream_t);
template int axpby_impl<double>(int64_t, double, double const *, int64_t, double, double *, int64_t, hipStream_t);
template int axpby_impl<float>(int64_t, float, float const *, int64_t, float, float *, int64_t, hipStream_t);
template int scale_strided<double>(int64_t, double, double *, int64_t, hipStream_t);
template int scale_strided<float>(int64_t, float, float *, int64_t, hipStream_t);
template int tensor_scale<double>(ein_tensor_desc const &, double *, double, hipStream_t);
template int tensor_scale<float>(ein_tensor_desc const &, float *, float, hipStream_t);

} // namespace ein

extern "C" {

int ein_gemv_f64(char trans, int64_t m, int64_t n, double alpha, const double *A, int64_t lda, const double *x, int64_t incx,
                 double beta, double *y, int64_t incy, ein_stream_t stream) {
    return ein::gemv_impl<double>(trans, m, n, alpha, A, lda, x, incx, beta, y, incy, ein::as_stream(stream));
}
int ein_gemv_f32(char trans, int64_t m, int64_t n, float alpha, const float *A, int64_t lda, const float *x, int64_t incx,
                 float beta, float *y, int64_t incy, ein_stream_t stream) {
    return ein::gemv_impl<float>(trans, m, n, alpha, A, lda, x, incx, beta, y, incy, ein::as_stream(stream));
}
int ein_dot_f64(int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy, double beta,
                double *result, ein_stream_t stream) {
    return ein::dot_impl<double>(n, alpha, x, incx, y, incy, beta, result, ein::as_stream(stream));
}
int ein_dot_f32(int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy, float beta, float *result,
                ein_stream_t stream) {
    return ein::dot_impl<float>(n, alpha, x, incx, y, incy, beta, result, ein::as_stream(stream));
}
int ein_dot_f64_noalloc(int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy,
                        double beta, double *result, ein_stream_t stream) {
    return ein::dot_noalloc_f64(n, alpha, x, incx, y, incy, beta, result, ein::as_stream(stream));
}
int ein_ger_f64(int64_t m, int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy, double *A,
                int64_t lda, ein_stream_t stream) {
    return ein::ger_impl<double>(m, n, alpha, x, incx, y, incy, A, lda, ein::as_stream(stream));
}
int ein_ger_f32(int64_t m, int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy, float *A,
                int64_t lda, ein_stream_t stream) {
    return ein::ger_impl<float>(m, n, alpha, x, incx, y, incy, A, lda, ein::as_stream(stream));
}
int ein_dirprod_f64(int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy, double beta,
                    double *z, int64_t incz, ein_stream_t stream) {
    return ein::dirprod_impl<double>(n, alpha, x, incx, y, incy, beta, z, incz, ein::as_stream(stream));
}
int ein_dirprod_f32(int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy, float beta, float *z,
                    int64_t incz, ein_stream_t stream) {
    return ein::dirprod_impl<float>(n, alpha, x, incx, y, incy, beta, z, incz, ein::as_stream(stream));
}
int ein_scale_f64(int64_t n, double alpha, double *x, int64_t incx, ein_stream_t stream) {
    if (n < 0) return ein::fail(-1, "scale: n < 0");
    return ein::scale_strided<double>(n, alpha, x, incx, ein::as_stream(stream));
}
int ein_scale_f32(int64_t n, float alpha, float *x, int64_t incx, ein_stream_t stream) {
    if (n < 0) return ein::fail(-1, "scale: n < 0");
    return ein::scale_strided<float>(n, alpha, x, incx, ein::as_stream(stream));
}
int ein_axpby_f64(int64_t n, double alpha, const double *x, int64_t incx, double beta, double *y, int64_t incy,
                  ein_stream_t stream) {
    return ein::axpby_impl<double>(n, alpha, x, incx, beta, y, incy, ein::as_stream(stream));
}
int ein_axpby_f32(int64_t n, float alpha, const float *x, int64_t incx, float beta, float *y, int64_t incy,
                  ein_stream_t stream) {
    return ein::axpby_impl<float>(n, alpha, x, incx, beta, y, incy, ein::as_stream(stream));
}
int ein_tensor_scale_f64(const ein_tensor_desc *C, double *c, double alpha, ein_stream_t stream) {
    if (C == nullptr) return -1;
    return ein::tensor_scale<double>(*C, c, alpha, ein::as_stream(stream));
}
int ein_tensor_scale_f32(const ein_tensor_desc *C, float *c, float alpha, ein_stream_t stream) {
    if (C == nullptr) return -1;
    return ein::tensor_scale<float>(*C, c, alpha, ein::as_stream(stream));
}

} // extern "C"
