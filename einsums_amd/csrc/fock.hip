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
// Kernel 1: workgroup (a, chunk of b) streams the slabs g(a,b,:,:) of its b's. Wave w takes
// the rows c = w, w+4, ...; its lanes run over d. Per lane it keeps J partials per b and K
// partials per row; J(a,b) is complete inside the workgroup and is written directly, K is
// written as one partial row per chunk. Kernel 2 adds the chunks' K partials in chunk order
// and forms F (deterministic: fixed summation trees, no atomics).
#include <algorithm>
#include <string>

#include "ein_common.hpp"

namespace ein {

namespace {

constexpr int FOCK_MAX_N  = 128; // D staged in LDS (128 KiB at n = 128)
constexpr int FOCK_MAX_BC = 4;   // b's per workgroup

// NPL = elements of a row per lane (1 for n <= 64, 2 for n <= 128); MR = rows c per wave.
template <int NPL, int MR = 16 * NPL>
__global__ __launch_bounds__(256) void fock_jk_kernel(int n, int bc, double const *__restrict__ g,
                                                      double const *__restrict__ D, int64_t ldd, double *__restrict__ J,
                                                      double *__restrict__ Kp) {
    extern __shared__ double Ds[]; // n x n, row stride n
    __shared__ double red[4][FOCK_MAX_BC];
    int const a    = int(blockIdx.x);
    int const ch   = int(blockIdx.y);
    int const wave = int(threadIdx.x) >> 6;
    int const lane = int(threadIdx.x) & 63;
    for (int e = int(threadIdx.x); e < n * n; e += 256) Ds[e] = D[int64_t(e / n) * ldd + e % n];
    __syncthreads();

    int const b0   = ch * bc;
    int const b1   = b0 + bc < n ? b0 + bc : n;
    int const rows = (n - wave + 3) / 4; // rows c = wave + 4*i, i < rows
    double kacc[MR];
#pragma unroll
    for (int i = 0; i < MR; ++i) kacc[i] = 0.0;
    int64_t const nn = int64_t(n) * n;
    // one b at a time (not unrolled, so the loads of one slab stay in flight together
    // without the compiler hoisting every slab's loads at once)
#pragma unroll 1
    for (int q = 0; q < FOCK_MAX_BC; ++q) {
        double jsum = 0.0;
        int const b      = b0 + q;
        bool const b_in  = b < b1;
        double const *__restrict__ slab = g + (int64_t(a) * n + (b_in ? b : b0)) * nn;
        // D(b, d) of this lane's d's
        double db[NPL];
#pragma unroll
        for (int u = 0; u < NPL; ++u) {
            int const d = lane + 64 * u;
            db[u]       = (b_in && d < n) ? Ds[b * n + d] : 0.0;
        }
        // rows in groups of RG: every load of a group is issued before any is used
        constexpr int RG = 8;
#pragma unroll
        for (int i0 = 0; i0 < MR; i0 += RG) {
            double x[RG][NPL];
#pragma unroll
            for (int i = 0; i < RG; ++i)
#pragma unroll
                for (int u = 0; u < NPL; ++u) {
                    int const c = wave + 4 * (i0 + i), d = lane + 64 * u;
                    x[i][u]     = (b_in && i0 + i < rows && d < n) ? slab[int64_t(c) * n + d] : 0.0;
                }
#pragma unroll
            for (int i = 0; i < RG; ++i) {
                int const c = wave + 4 * (i0 + i);
#pragma unroll
                for (int u = 0; u < NPL; ++u) {
                    int const d     = lane + 64 * u;
                    double const dc = (i0 + i < rows && d < n) ? Ds[c * n + d] : 0.0;
                    jsum += x[i][u] * dc;
                    kacc[i0 + i] += x[i][u] * db[u];
                }
            }
        }
        // J(a, b): lanes here, the four waves after the loop
        jsum = wave_reduce_sum(jsum);
        if (lane == 0) red[wave][q] = jsum;
    }
    // K partial row of this chunk: Kp[(a * chunks + ch) * n + c]
    double *__restrict__ kp = Kp + (int64_t(a) * gridDim.y + ch) * n;
#pragma unroll
    for (int i = 0; i < MR; ++i) {
        double const s = wave_reduce_sum(kacc[i]);
        if (lane == 0 && i < rows) kp[wave + 4 * i] = s;
    }
    __syncthreads();
    if (threadIdx.x < unsigned(b1 - b0))
        J[int64_t(a) * n + b0 + int(threadIdx.x)] =
            (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// F(a,c) = beta*F(a,c) + alpha_j*J(a,c) + alpha_k*sum_chunk Kp(a,chunk,c), chunks in order.
__global__ __launch_bounds__(256) void fock_finish_kernel(int n, int chunks, double const *__restrict__ J,
                                                          double const *__restrict__ Kp, double alpha_j, double alpha_k,
                                                          double beta, double *__restrict__ F, int64_t ldf) {
    int64_t const e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= int64_t(n) * n) return;
    int const a = int(e / n), c = int(e % n);
    double k = 0.0;
    for (int ch = 0; ch < chunks; ++ch) k += Kp[(int64_t(a) * chunks + ch) * n + c];
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
    // b's per workgroup: enough workgroups for four per CU where n allows it
    int const cus = num_cus();
    int bc        = 1;
    while (bc < FOCK_MAX_BC && n * ceil_div(n, bc * 2) >= int64_t(cus) * 4) bc *= 2;
    int const chunks = int(ceil_div(n, bc));
    void *ws = nullptr;
    size_t const j_bytes = size_t(n) * size_t(n) * sizeof(double);
    size_t const k_bytes = size_t(n) * size_t(chunks) * size_t(n) * sizeof(double);
    EIN_TRY(workspace(stream, j_bytes + k_bytes, &ws));
    double *J  = static_cast<double *>(ws);
    double *Kp = J + size_t(n) * size_t(n);
    size_t const lds = size_t(n) * size_t(n) * sizeof(double);
    dim3 grid{uint32_t(n), uint32_t(chunks)};
    if (n <= 64) {
        fock_jk_kernel<1><<<grid, 256, lds, stream>>>(int(n), bc, g, D, ldd, J, Kp);
    } else {
        // more than 64 KiB of dynamic LDS has to be allowed explicitly
        static hipError_t const attr = hipFuncSetAttribute(reinterpret_cast<void const *>(&fock_jk_kernel<2>),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                                           int(FOCK_MAX_N * FOCK_MAX_N * sizeof(double)));
        if (attr != hipSuccess) return fail(EIN_ERROR_HIP, "fock_jk: cannot enable 128 KiB of LDS");
        fock_jk_kernel<2><<<grid, 256, lds, stream>>>(int(n), bc, g, D, ldd, J, Kp);
    }
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
