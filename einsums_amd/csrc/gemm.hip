// Dense strided GEMM on MFMA for gfx950 (FP64 and FP32), with strided batching and
// deterministic split-K.
//
// C(i,j) = alpha * sum_p A(i,p) B(p,j) + beta * C(i,j) with
//   A(i,p) = A[i*a_sm + p*a_sk + b*a_sb], B(p,j) = B[p*b_sk + j*b_sn + b*b_sb],
//   C(i,j) = C[i*c_sm + j*c_sn + b*c_sb].
// Column-major BLAS calls (ein_gemm_*) and the folded 2-D views of gemm-mapped einsums
// both land here. Reference leaf replaced: blas::gpu::gemm -> hipblasDgemm_64
// (libs/Einsums/hipBLAS/include/Einsums/hipBLAS.hpp:64, hipBLASVendor/src/gemm.hip:82-155).
#include <algorithm>
#include <string>

#include "gemm.hpp"

namespace ein {

template <typename T, int BM, int BN, int BK, bool A_I_FAST, bool B_J_FAST>
__global__ __launch_bounds__(GEMM_THREADS, 2) void gemm_dense_kernel(GemmProblem<T> P) {
    using Cfg = TileCfg<T, BM, BN, BK>;
    __shared__ T As[2][BK * Cfg::LDA];
    __shared__ T Bs[2][BK * Cfg::LDB];

    int const tid  = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const wm   = wave & 1;
    int const wn   = wave >> 1;

    uint32_t const ntiles = P.tiles_m * P.tiles_n;
    uint32_t const logical = xcd_remap(blockIdx.x, ntiles);
    uint32_t tm, tn;
    tile_coords(logical, P.tiles_m, P.tiles_n, tm, tn);
    int64_t const i0 = int64_t(tm) * BM;
    int64_t const j0 = int64_t(tn) * BN;

    int const split      = blockIdx.y;
    int64_t const kbeg   = int64_t(split) * P.k_chunk;
    int64_t const kend   = kbeg + P.k_chunk < P.k ? kbeg + P.k_chunk : P.k;
    int64_t const batch  = blockIdx.z;
    T const *__restrict__ Ab = P.A + batch * P.a_sb;
    T const *__restrict__ Bb = P.B + batch * P.b_sb;

    // ---- per-thread load coordinates
    // A slab: BM x BK. i-fast: consecutive threads walk i; else walk p.
    constexpr int A_STEP = A_I_FAST ? GEMM_THREADS / BM : GEMM_THREADS / BK;
    int const a_ii = A_I_FAST ? tid % BM : tid / BK;
    int const a_pp = A_I_FAST ? tid / BM : tid % BK;
    constexpr int B_STEP = B_J_FAST ? GEMM_THREADS / BN : GEMM_THREADS / BK;
    int const b_jj = B_J_FAST ? tid % BN : tid / BK;
    int const b_pp = B_J_FAST ? tid / BN : tid % BK;

    T ra[Cfg::LA], rb[Cfg::LB];

    auto load_slab = [&](int64_t k0) {
#pragma unroll
        for (int r = 0; r < Cfg::LA; ++r) {
            int ii    = A_I_FAST ? a_ii : a_ii + r * A_STEP;
            int pp    = A_I_FAST ? a_pp + r * A_STEP : a_pp;
            int64_t i = i0 + ii, p = k0 + pp;
            ra[r]     = (i < P.m && p < kend) ? Ab[i * P.a_sm + p * P.a_sk] : T(0);
        }
#pragma unroll
        for (int r = 0; r < Cfg::LB; ++r) {
            int jj    = B_J_FAST ? b_jj : b_jj + r * B_STEP;
            int pp    = B_J_FAST ? b_pp + r * B_STEP : b_pp;
            int64_t j = j0 + jj, p = k0 + pp;
            rb[r]     = (j < P.n && p < kend) ? Bb[p * P.b_sk + j * P.b_sn] : T(0);
        }
    };
    auto store_slab = [&](int buf) {
#pragma unroll
        for (int r = 0; r < Cfg::LA; ++r) {
            int ii = A_I_FAST ? a_ii : a_ii + r * A_STEP;
            int pp = A_I_FAST ? a_pp + r * A_STEP : a_pp;
            As[buf][pp * Cfg::LDA + ii] = ra[r];
        }
#pragma unroll
        for (int r = 0; r < Cfg::LB; ++r) {
            int jj = B_J_FAST ? b_jj : b_jj + r * B_STEP;
            int pp = B_J_FAST ? b_pp + r * B_STEP : b_pp;
            Bs[buf][pp * Cfg::LDB + jj] = rb[r];
        }
    };

    WaveAcc<T, Cfg::FM, Cfg::FN> acc;
    acc.zero();

    int64_t const nslabs = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
    if (nslabs > 0) {
        load_slab(kbeg);
        store_slab(0);
        __syncthreads();
        for (int64_t s = 0; s < nslabs; ++s) {
            int const buf = int(s & 1);
            if (s + 1 < nslabs) load_slab(kbeg + (s + 1) * BK);
            mfma_slab<T, BM, BN, BK>(As[buf], Bs[buf], acc, wm, wn, lane);
            if (s + 1 < nslabs) store_slab(buf ^ 1);
            __syncthreads();
        }
    }

    // ---- epilogue
    int const r = lane & 15;
    int const q = lane >> 4;
    if (P.splits > 1) {
        T *__restrict__ part = P.partial + (int64_t(split) * P.batch + batch) * P.m * P.n;
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
            for (int b = 0; b < Cfg::FN; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    int64_t i = i0 + wm * Cfg::WM + a * 16 + r;
                    int64_t j = j0 + wn * Cfg::WN + b * 16 + Mfma<T>::row(q, v);
                    if (i < P.m && j < P.n) part[i + j * P.m] = acc.c[a][b][v];
                }
        return;
    }
    T *__restrict__ Cb = P.C + batch * P.c_sb;
#pragma unroll
    for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
        for (int b = 0; b < Cfg::FN; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int64_t i = i0 + wm * Cfg::WM + a * 16 + r;
                int64_t j = j0 + wn * Cfg::WN + b * 16 + Mfma<T>::row(q, v);
                if (i < P.m && j < P.n) {
                    T *cp = Cb + i * P.c_sm + j * P.c_sn;
                    T val = P.alpha * acc.c[a][b][v];
                    if (P.beta != T(0)) val += P.beta * *cp;
                    *cp = val;
                }
            }
}

// C = alpha * sum_s partial[s] + beta * C, one thread per element, i fastest.
template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(GemmProblem<T> P) {
    int64_t const mn    = P.m * P.n;
    int64_t const total = mn * P.batch;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        int64_t b   = e / mn;
        int64_t rem = e - b * mn;
        int64_t j   = rem / P.m;
        int64_t i   = rem - j * P.m;
        T sum       = T(0);
        for (int s = 0; s < P.splits; ++s) sum += P.partial[(int64_t(s) * P.batch + b) * mn + rem];
        T *cp = P.C + b * P.c_sb + i * P.c_sm + j * P.c_sn;
        T val = P.alpha * sum;
        if (P.beta != T(0)) val += P.beta * *cp;
        *cp = val;
    }
}

// C = beta*C over the m x n (x batch) view (k == 0 or alpha == 0 case); beta == 0 writes zeros.
template <typename T>
__global__ __launch_bounds__(256) void scale_2d_kernel(GemmProblem<T> P) {
    int64_t const mn    = P.m * P.n;
    int64_t const total = mn * P.batch;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        int64_t b   = e / mn;
        int64_t rem = e - b * mn;
        int64_t j   = rem / P.m;
        int64_t i   = rem - j * P.m;
        T *cp       = P.C + b * P.c_sb + i * P.c_sm + j * P.c_sn;
        *cp         = P.beta == T(0) ? T(0) : P.beta * *cp;
    }
}

namespace {

template <typename T, int BM, int BN, int BK>
void launch_dense(GemmProblem<T> const &P, dim3 grid, hipStream_t stream) {
    bool const a_i_fast = (P.a_sm == 1) || (P.a_sk != 1 && P.a_sm <= P.a_sk);
    bool const b_j_fast = (P.b_sn == 1) || (P.b_sk != 1 && P.b_sn <= P.b_sk);
    if (a_i_fast && b_j_fast)
        gemm_dense_kernel<T, BM, BN, BK, true, true><<<grid, GEMM_THREADS, 0, stream>>>(P);
    else if (a_i_fast && !b_j_fast)
        gemm_dense_kernel<T, BM, BN, BK, true, false><<<grid, GEMM_THREADS, 0, stream>>>(P);
    else if (!a_i_fast && b_j_fast)
        gemm_dense_kernel<T, BM, BN, BK, false, true><<<grid, GEMM_THREADS, 0, stream>>>(P);
    else
        gemm_dense_kernel<T, BM, BN, BK, false, false><<<grid, GEMM_THREADS, 0, stream>>>(P);
}

} // namespace

GemmConfig choose_gemm_config(int64_t m, int64_t n, int64_t k, int64_t batch, int elem_bytes) {
    (void)elem_bytes;
    GemmConfig cfg;
    int const cus  = num_cus();
    int64_t t128   = ceil_div(m, 128) * ceil_div(n, 128) * batch;
    int64_t t64    = ceil_div(m, 64) * ceil_div(n, 64) * batch;
    // Two workgroups are resident per CU; aim for at least ~1.5 waves of workgroups.
    int64_t target = int64_t(cus) * 3;
    if (t128 >= target || (m >= 128 && n >= 128 && t128 >= int64_t(cus) * 2)) {
        cfg.bm = 128;
        cfg.bn = 128;
    } else {
        cfg.bm = 64;
        cfg.bn = 64;
    }
    int64_t tiles = cfg.bm == 128 ? t128 : t64;
    cfg.splits    = 1;
    if (tiles < target && k >= 512) {
        int64_t s = ceil_div(target, tiles);
        s         = std::min<int64_t>(s, k / 256);
        s         = std::max<int64_t>(s, 1);
        cfg.splits = int(std::min<int64_t>(s, 64));
    }
    return cfg;
}

template <typename T>
int gemm_strided(GemmProblem<T> P, GemmConfig const *force, hipStream_t stream) {
    if (P.m <= 0 || P.n <= 0 || P.batch <= 0) return EIN_SUCCESS;
    // Normalise so that C's unit-stride direction is i (lane-contiguous stores):
    // C^T = B^T A^T swaps the roles of the operands.
    bool const c_row_fast = (P.c_sn == 1 && P.c_sm != 1) || (P.c_sm != 1 && P.c_sn != 1 && P.c_sn < P.c_sm);
    if (c_row_fast) {
        GemmProblem<T> Q = P;
        Q.m = P.n;
        Q.n = P.m;
        Q.A = P.B;
        Q.a_sm = P.b_sn;
        Q.a_sk = P.b_sk;
        Q.a_sb = P.b_sb;
        Q.B = P.A;
        Q.b_sk = P.a_sk;
        Q.b_sn = P.a_sm;
        Q.b_sb = P.a_sb;
        Q.c_sm = P.c_sn;
        Q.c_sn = P.c_sm;
        P      = Q;
    }
    int64_t const total = P.m * P.n * P.batch;
    if (P.k <= 0 || P.alpha == T(0)) {
        int blocks = int(std::min<int64_t>(ceil_div(total, 256), 65536));
        scale_2d_kernel<T><<<blocks, 256, 0, stream>>>(P);
        EIN_LAUNCH_CHECK();
        return EIN_SUCCESS;
    }
    if (P.batch > 65535) return fail(EIN_ERROR_UNSUPPORTED, "gemm batch count above 65535");
    GemmConfig cfg = force ? *force : choose_gemm_config(P.m, P.n, P.k, P.batch, int(sizeof(T)));
    constexpr int BK = 16;
    P.tiles_m        = uint32_t(ceil_div(P.m, cfg.bm));
    P.tiles_n        = uint32_t(ceil_div(P.n, cfg.bn));
    P.splits         = std::max(cfg.splits, 1);
    P.k_chunk        = ceil_div(ceil_div(P.k, P.splits), BK) * BK;
    P.splits         = int(ceil_div(P.k, P.k_chunk));
    P.partial        = nullptr;
    if (P.splits > 1) {
        void *ws     = nullptr;
        size_t bytes = size_t(P.splits) * size_t(P.batch) * size_t(P.m) * size_t(P.n) * sizeof(T);
        EIN_TRY(workspace(stream, bytes, &ws));
        P.partial = static_cast<T *>(ws);
    }
    int64_t ntiles = int64_t(P.tiles_m) * P.tiles_n;
    if (ntiles > 0x7fffffff) return fail(EIN_ERROR_UNSUPPORTED, "gemm tile grid too large");
    dim3 grid{uint32_t(ntiles), uint32_t(P.splits), uint32_t(P.batch)};
    if (cfg.bm == 128 && cfg.bn == 128)
        launch_dense<T, 128, 128, BK>(P, grid, stream);
    else if (cfg.bm == 128 && cfg.bn == 64)
        launch_dense<T, 128, 64, BK>(P, grid, stream);
    else if (cfg.bm == 64 && cfg.bn == 128)
        launch_dense<T, 64, 128, BK>(P, grid, stream);
    else
        launch_dense<T, 64, 64, BK>(P, grid, stream);
    EIN_LAUNCH_CHECK();
    if (P.splits > 1) {
        int blocks = int(std::min<int64_t>(ceil_div(total, 256), 65536));
        splitk_reduce_kernel<T><<<blocks, 256, 0, stream>>>(P);
        EIN_LAUNCH_CHECK();
    }
    return EIN_SUCCESS;
}

template int gemm_strided<double>(GemmProblem<double>, GemmConfig const *, hipStream_t);
template int gemm_strided<float>(GemmProblem<float>, GemmConfig const *, hipStream_t);

namespace {

bool valid_trans(char t) { return t == 'N' || t == 'n' || t == 'T' || t == 't' || t == 'C' || t == 'c'; }
bool is_notrans(char t) { return t == 'N' || t == 'n'; }

// Column-major BLAS argument checks (same order and codes as the reference's
// hipBLASVendor/src/gemm.hip:23-80 and reference BLAS xGEMM).
int check_gemm_args(char ta, char tb, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc) {
    if (!valid_trans(ta)) return fail(-1, "gemm: invalid transa");
    if (!valid_trans(tb)) return fail(-2, "gemm: invalid transb");
    if (m < 0) return fail(-3, "gemm: m < 0");
    if (n < 0) return fail(-4, "gemm: n < 0");
    if (k < 0) return fail(-5, "gemm: k < 0");
    int64_t nrowa = is_notrans(ta) ? m : k;
    int64_t nrowb = is_notrans(tb) ? k : n;
    if (lda < std::max<int64_t>(1, nrowa)) return fail(-8, "gemm: lda too small");
    if (ldb < std::max<int64_t>(1, nrowb)) return fail(-10, "gemm: ldb too small");
    if (ldc < std::max<int64_t>(1, m)) return fail(-13, "gemm: ldc too small");
    return EIN_SUCCESS;
}

template <typename T>
int gemm_colmajor(char ta, char tb, int64_t m, int64_t n, int64_t k, T alpha, T const *A, int64_t lda, int64_t sa,
                  T const *B, int64_t ldb, int64_t sb, T beta, T *C, int64_t ldc, int64_t sc, int64_t batch,
                  hipStream_t stream) {
    EIN_TRY(check_gemm_args(ta, tb, m, n, k, lda, ldb, ldc));
    if (batch < 0) return fail(-17, "gemm: batch < 0");
    GemmProblem<T> P{};
    P.m     = m;
    P.n     = n;
    P.k     = k;
    P.batch = batch;
    P.A     = A;
    P.B     = B;
    P.C     = C;
    P.alpha = alpha;
    P.beta  = beta;
    if (is_notrans(ta)) {
        P.a_sm = 1;
        P.a_sk = lda;
    } else {
        P.a_sm = lda;
        P.a_sk = 1;
    }
    if (is_notrans(tb)) {
        P.b_sk = 1;
        P.b_sn = ldb;
    } else {
        P.b_sk = ldb;
        P.b_sn = 1;
    }
    P.c_sm = 1;
    P.c_sn = ldc;
    P.a_sb = sa;
    P.b_sb = sb;
    P.c_sb = sc;
    return gemm_strided<T>(P, nullptr, stream);
}

} // namespace
} // namespace ein

extern "C" {

int ein_gemm_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                 const double *B, int64_t ldb, double beta, double *C, int64_t ldc, ein_stream_t stream) {
    return ein::gemm_colmajor<double>(transa, transb, m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1,
                                      ein::as_stream(stream));
}

int ein_gemm_f32(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda,
                 const float *B, int64_t ldb, float beta, float *C, int64_t ldc, ein_stream_t stream) {
    return ein::gemm_colmajor<float>(transa, transb, m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1,
                                     ein::as_stream(stream));
}

int ein_gemm_strided_batched_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                                 int64_t lda, int64_t strideA, const double *B, int64_t ldb, int64_t strideB, double beta,
                                 double *C, int64_t ldc, int64_t strideC, int64_t batch, ein_stream_t stream) {
    return ein::gemm_colmajor<double>(transa, transb, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                      strideC, batch, ein::as_stream(stream));
}

int ein_gemm_strided_batched_f32(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                                 int64_t lda, int64_t strideA, const float *B, int64_t ldb, int64_t strideB, float beta,
                                 float *C, int64_t ldc, int64_t strideC, int64_t batch, ein_stream_t stream) {
    return ein::gemm_colmajor<float>(transa, transb, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                     strideC, batch, ein::as_stream(stream));
}

} // extern "C"
