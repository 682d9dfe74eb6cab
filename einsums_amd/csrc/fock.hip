// Coulomb and exchange matrices of a four-index tensor in one pass over it (the Fock build
// of a Hartree-Fock iteration):
//   F = beta*F + alpha_j*J + alpha_k*K,
//   J(a,b) = sum_cd g(a,b,c,d) D(c,d),   K(a,c) = sum_bd g(a,b,c,d) D(b,d),
// with g row-major n^4 and D, F row-major n x n (leading dimensions ldd, ldf).
//
// The reference forms these with two einsum calls that each read all of g:
//   einsum(1, {i,j}, &F,  2, {k,l}, D, {i,j,k,l}, TEI)   J, the GEMV case
//   einsum(1, {i,j}, &F, -1, {k,l}, D, {i,k,j,l}, TEI)   K, the generic case
// (libs/Einsums/TensorAlgebra/tests/unit/QChemTests.cpp:299-304; the J and K builds of
// devtools/profiling/profile_einsums.cpp:23-30). Element g(a,b,c,d) feeds J(a,b) with D(c,d)
// and K(a,c) with D(b,d), so both can be accumulated from one read of g: half the HBM traffic
// of the two calls, which is what bounds them (0.25 flop per byte).
//
// Kernel 1: workgroup (a, one or four b's) reads the slabs g(a,b,:,:) of its b's. Wave w
// takes the rows c = w, w+W, ... (W = 4 or 16 waves); its lanes run over d. Per lane it keeps
// K partials per row; J(a,b) is complete inside the workgroup and is written directly, K is
// written as one partial row per chunk. Kernel 2 adds the chunks' K partials in chunk order
// and forms F (deterministic: fixed summation trees, no atomics). n <= 128. For even n <= 64
// with 16-byte aligned rows, kernel 1 reads two elements per lane (fock_jk16_kernel).
#include <algorithm>
#include <string>

#include "ein_common.hpp"

namespace ein {

namespace {

constexpr int FOCK_MAX_N = 128;

// NPL = elements of a row per lane (1 for n <= 64, 2 for n <= 128); WAVES = 4 or 16;
// MR >= ceil(n / WAVES) rows c per wave; BC = b's per workgroup. Each lane keeps the
// D(c, d) of its own (c, d) positions in registers (the same for every slab), so no LDS
// staging or barrier precedes the slab loads: all of a workgroup's slab reads are in
// flight at once, and the many resident workgroups keep HBM busy.
template <int NPL, int WAVES, int MR, int BC>
__global__ __launch_bounds__(64 * WAVES) void fock_jk_kernel(int n, double const *__restrict__ g,
                                                             double const *__restrict__ D, int64_t ldd,
                                                             double *__restrict__ J, double *__restrict__ Kp) {
    __shared__ double red[WAVES][BC];
    int const a    = int(blockIdx.x);
    int const ch   = int(blockIdx.y);
    int const wave = int(threadIdx.x) >> 6;
    int const lane = int(threadIdx.x) & 63;
    int const b0   = ch * BC;
    int const rows = (n - wave + WAVES - 1) / WAVES; // rows c = wave + WAVES*i, i < rows
    int64_t const nn = int64_t(n) * n;

    double x[BC][MR][NPL];
#pragma unroll
    for (int q = 0; q < BC; ++q) {
        bool const b_in = b0 + q < n;
        double const *__restrict__ slab = g + (int64_t(a) * n + (b_in ? b0 + q : b0)) * nn;
#pragma unroll
        for (int i = 0; i < MR; ++i)
#pragma unroll
            for (int u = 0; u < NPL; ++u) {
                int const c = wave + WAVES * i, d = lane + 64 * u;
                x[q][i][u]  = (b_in && i < rows && d < n) ? slab[int64_t(c) * n + d] : 0.0;
            }
    }
    double dc[MR][NPL], db[BC][NPL];
#pragma unroll
    for (int i = 0; i < MR; ++i)
#pragma unroll
        for (int u = 0; u < NPL; ++u) {
            int const c = wave + WAVES * i, d = lane + 64 * u;
            dc[i][u]    = (i < rows && d < n) ? D[int64_t(c) * ldd + d] : 0.0;
        }
#pragma unroll
    for (int q = 0; q < BC; ++q)
#pragma unroll
        for (int u = 0; u < NPL; ++u) {
            int const d = lane + 64 * u;
            db[q][u]    = (b0 + q < n && d < n) ? D[int64_t(b0 + q) * ldd + d] : 0.0;
        }

    double kacc[MR];
#pragma unroll
    for (int i = 0; i < MR; ++i) kacc[i] = 0.0;
#pragma unroll
    for (int q = 0; q < BC; ++q) {
        double jsum = 0.0;
#pragma unroll
        for (int i = 0; i < MR; ++i)
#pragma unroll
            for (int u = 0; u < NPL; ++u) {
                jsum += x[q][i][u] * dc[i][u];
                kacc[i] += x[q][i][u] * db[q][u];
            }
        // J(a, b): lanes here, the waves after the barrier
        jsum = wave_reduce_sum(jsum);
        if (lane == 0) red[wave][q] = jsum;
    }
    // K partial row of this chunk: Kp[(a * chunks + ch) * n + c]
    double *__restrict__ kp = Kp + (int64_t(a) * gridDim.y + ch) * n;
#pragma unroll
    for (int i = 0; i < MR; ++i) {
        double const s = wave_reduce_sum(kacc[i]);
        if (lane == 0 && i < rows) kp[wave + WAVES * i] = s;
    }
    __syncthreads();
    if (int(threadIdx.x) < BC && b0 + int(threadIdx.x) < n) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) s += red[w][threadIdx.x];
        J[int64_t(a) * n + b0 + int(threadIdx.x)] = s;
    }
}

// The same pass with 16-byte loads, for even n <= 64 with 16-byte aligned rows of g and D
// (even ldd): lane l reads d = 2(l%32), 2(l%32)+1 of row c = 2p + l/32 (lanes with
// 2(l%32) >= n idle), so one load covers two adjacent rows of a slab; wave w takes the row
// pairs p = w, w+4, ... (MP >= ceil(n/8) of them). K's row sums reduce within each 32-lane
// half. Measured at n = 48 (HF): 18.0 against 21.1 us cold for the two launches.
template <int MP, int BC>
__global__ __launch_bounds__(256) void fock_jk16_kernel(int n, double const *__restrict__ g, double const *__restrict__ D,
                                                        int64_t ldd, double *__restrict__ J, double *__restrict__ Kp) {
    constexpr int WAVES = 4;
    __shared__ double red[WAVES][BC];
    int const a      = int(blockIdx.x);
    int const ch     = int(blockIdx.y);
    int const wave   = int(threadIdx.x) >> 6;
    int const lane   = int(threadIdx.x) & 63;
    int const h      = lane >> 5;
    int const d0     = 2 * (lane & 31);
    bool const d_in  = d0 < n;
    int const b0     = ch * BC;
    int const np     = (n / 2 - wave + WAVES - 1) / WAVES; // row pairs p = wave + WAVES*i, i < np
    int64_t const nn = int64_t(n) * n;

    double2 x[BC][MP];
#pragma unroll
    for (int q = 0; q < BC; ++q) {
        bool const b_in = b0 + q < n;
        double const *__restrict__ slab = g + (int64_t(a) * n + (b_in ? b0 + q : b0)) * nn;
#pragma unroll
        for (int i = 0; i < MP; ++i) {
            int const c = 2 * (wave + WAVES * i) + h;
            x[q][i]     = (b_in && i < np && d_in) ? *reinterpret_cast<double2 const *>(slab + int64_t(c) * n + d0)
                                                   : double2{0.0, 0.0};
        }
    }
    double2 dc[MP], db[BC];
#pragma unroll
    for (int i = 0; i < MP; ++i) {
        int const c = 2 * (wave + WAVES * i) + h;
        dc[i] = (i < np && d_in) ? *reinterpret_cast<double2 const *>(D + int64_t(c) * ldd + d0) : double2{0.0, 0.0};
    }
#pragma unroll
    for (int q = 0; q < BC; ++q)
        db[q] = (b0 + q < n && d_in) ? *reinterpret_cast<double2 const *>(D + int64_t(b0 + q) * ldd + d0)
                                     : double2{0.0, 0.0};

    double kacc[MP];
#pragma unroll
    for (int i = 0; i < MP; ++i) kacc[i] = 0.0;
#pragma unroll
    for (int q = 0; q < BC; ++q) {
        double jsum = 0.0;
#pragma unroll
        for (int i = 0; i < MP; ++i) {
            jsum += x[q][i].x * dc[i].x + x[q][i].y * dc[i].y;
            kacc[i] += x[q][i].x * db[q].x + x[q][i].y * db[q].y;
        }
        jsum = wave_reduce_sum(jsum);
        if (lane == 0) red[wave][q] = jsum;
    }
    double *__restrict__ kp = Kp + (int64_t(a) * gridDim.y + ch) * n;
#pragma unroll
    for (int i = 0; i < MP; ++i) {
        double s = kacc[i];
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) s += __shfl_xor(s, o, 64); // within the 32-lane half
        if ((lane & 31) == 0 && i < np) kp[2 * (wave + WAVES * i) + h] = s;
    }
    __syncthreads();
    if (int(threadIdx.x) < BC && b0 + int(threadIdx.x) < n) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) s += red[w][threadIdx.x];
        J[int64_t(a) * n + b0 + int(threadIdx.x)] = s;
    }
}

// F(a,c) = beta*F(a,c) + alpha_j*J(a,c) + alpha_k*sum_chunk Kp(a,chunk,c), chunks in order.
__global__ __launch_bounds__(256) void fock_finish_kernel(int n, int chunks, double const *__restrict__ J,
                                                          double const *__restrict__ Kp, double alpha_j, double alpha_k,
                                                          double beta, double *__restrict__ F, int64_t ldf) {
    int64_t const e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= int64_t(n) * n) return;
    int const a = int(e / n), c = int(e % n);
    // four independent partial sums so the chunk loads are in flight together
    double const *__restrict__ kp = Kp + int64_t(a) * chunks * n + c;
    double k4[4] = {0.0, 0.0, 0.0, 0.0};
    int ch = 0;
    for (; ch + 3 < chunks; ch += 4)
#pragma unroll
        for (int u = 0; u < 4; ++u) k4[u] += kp[int64_t(ch + u) * n];
    for (; ch < chunks; ++ch) k4[0] += kp[int64_t(ch) * n];
    double const k = (k4[0] + k4[1]) + (k4[2] + k4[3]);
    double *fp = F + int64_t(a) * ldf + c;
    double v   = alpha_j * J[e] + alpha_k * k;
    if (beta != 0.0) v += beta * *fp;
    *fp = v;
}

} // namespace

int fock_jk_f64(int64_t n, double const *g, double const *D, int64_t ldd, double alpha_j, double alpha_k, double beta,
                double *F, int64_t ldf, hipStream_t stream) {
    if (n < 0) return fail(-1, "fock_jk: n < 0");
    if (ldd < std::max<int64_t>(1, n)) return fail(-4, "fock_jk: ldd < n");
    if (ldf < std::max<int64_t>(1, n)) return fail(-9, "fock_jk: ldf < n");
    if (n > FOCK_MAX_N)
        return fail(EIN_ERROR_UNSUPPORTED, "fock_jk: n = " + std::to_string(n) + " exceeds " + std::to_string(FOCK_MAX_N) +
                                               " (use the two einsum calls)");
    if (n == 0) return EIN_SUCCESS;
    if (g == nullptr || D == nullptr || F == nullptr) return fail(-2, "fock_jk: null pointer");
    int const bc     = n <= 64 ? 4 : 1; // the kernels' BC
    int const chunks = int(ceil_div(n, bc));
    void *ws = nullptr;
    size_t const j_bytes = size_t(n) * size_t(n) * sizeof(double);
    size_t const k_bytes = size_t(n) * size_t(chunks) * size_t(n) * sizeof(double);
    EIN_TRY(workspace(stream, j_bytes + k_bytes, &ws));
    double *J  = static_cast<double *>(ws);
    double *Kp = J + size_t(n) * size_t(n);
    dim3 grid{uint32_t(n), uint32_t(chunks)};
    int const ni = int(n);
    // two elements per lane for even n <= 64 (a one-row-per-load form for 64 < n <= 128
    // measured slower than the kernel below it: 215 against 201 us cold at n = 96)
    bool const wide = n % 2 == 0 && n <= 64 && ldd % 2 == 0 && reinterpret_cast<uintptr_t>(g) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(D) % 16 == 0;
    if (wide && n <= 32)
        fock_jk16_kernel<4, 4><<<grid, 256, 0, stream>>>(ni, g, D, ldd, J, Kp);
    else if (wide && n <= 48)
        fock_jk16_kernel<6, 4><<<grid, 256, 0, stream>>>(ni, g, D, ldd, J, Kp);
    else if (wide)
        fock_jk16_kernel<8, 4><<<grid, 256, 0, stream>>>(ni, g, D, ldd, J, Kp);
    else if (n <= 32)
        fock_jk_kernel<1, 4, 8, 4><<<grid, 256, 0, stream>>>(ni, g, D, ldd, J, Kp);
    else if (n <= 48)
        fock_jk_kernel<1, 4, 12, 4><<<grid, 256, 0, stream>>>(ni, g, D, ldd, J, Kp);
    else if (n <= 64)
        fock_jk_kernel<1, 4, 16, 4><<<grid, 256, 0, stream>>>(ni, g, D, ldd, J, Kp);
    else if (n <= 96)
        fock_jk_kernel<2, 16, 6, 1><<<grid, 1024, 0, stream>>>(ni, g, D, ldd, J, Kp);
    else
        fock_jk_kernel<2, 16, 8, 1><<<grid, 1024, 0, stream>>>(ni, g, D, ldd, J, Kp);
    EIN_LAUNCH_CHECK();
    fock_finish_kernel<<<uint32_t(ceil_div(n * n, 256)), 256, 0, stream>>>(int(n), chunks, J, Kp, alpha_j, alpha_k, beta,
                                                                         F, ldf);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

} // namespace ein

extern "C" {

int ein_fock_jk_f64(int64_t n, const double *g, const double *D, int64_t ldd, double alpha_j, double alpha_k,
                    double beta, double *F, int64_t ldf, ein_stream_t stream) {
    return ein::fock_jk_f64(n, g, D, ldd, alpha_j, alpha_k, beta, F, ldf, ein::as_stream(stream));
}

} // extern "C"
