// Dense symmetric eigensolver and LU solve for the small matrices of the SCF loop around the
// einsum path (Fock matrices of nbf x nbf, DIIS systems of a few dozen unknowns), so that a
// whole SCF iteration stays on the device and can be captured in a graph. Both run on one
// 1024-thread workgroup and report LAPACK's `info` in device memory instead of returning it,
// which keeps the calls asynchronous.
//
// Reference interfaces (paths relative to the Einsums source tree):
//   syev: blas::syev(job, uplo, n, a, lda, w, work, lwork)  libs/Einsums/BLAS/include/Einsums/BLAS.hpp:306
//         linear_algebra::syev(&A, &W)                      libs/Einsums/LinearAlgebra/include/Einsums/LinearAlgebra.hpp:362
//   gesv: blas::gesv(n, nrhs, a, lda, ipiv, b, ldb)         libs/Einsums/BLAS/include/Einsums/BLAS.hpp:529
//         linear_algebra::gesv(&A, &B)                      libs/Einsums/LinearAlgebra/include/Einsums/LinearAlgebra.hpp:495
// used by the reference HF/DIIS test, libs/Einsums/TensorAlgebra/tests/unit/QChemTests.cpp:157,266,312.
//
// syev: cyclic two-sided Jacobi (Golub & Van Loan, Matrix Computations, 4th ed., 8.5: the
// symmetric Schur decomposition of each 2x2 block, Algorithm 8.5.1) with the round-robin
// ordering of Brent and Luk, so every step rotates n/2 disjoint planes at once. A sweep is
// n-1 steps; the loop stops once the off-diagonal Frobenius norm is below n*eps times that of
// A (converged), or after 60 sweeps (info = 1). Eigenvalues come out ascending, eigenvectors
// as the columns of A, as from LAPACK dsyev.
// gesv: right-looking LU with partial pivoting (the first largest magnitude in the column, as
// LAPACK's idamax picks), then unit-lower and upper triangular solves of B; ipiv is 1-based.
#include <cmath>

#include "ein_common.hpp"

namespace ein {

namespace {

constexpr int LA_THREADS = 1024;
constexpr int LA_WAVES   = LA_THREADS / 64;
constexpr int SYEV_MAX_N = 1024; // n/2 rotation pairs held in LDS
constexpr int GESV_MAX_N = 4096;
constexpr int SYEV_MAX_SWEEPS = 60;

// Sum over the workgroup in a fixed order (deterministic).
template <typename T>
__device__ T block_sum(T v, T *red) {
    v              = wave_reduce_sum(v);
    int const lane = threadIdx.x & 63;
    int const w    = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    T s = T(0);
#pragma unroll
    for (int i = 0; i < LA_WAVES; ++i) s += red[i];
    __syncthreads();
    return s;
}

// Work layout: M (n x n, column-major, the full symmetric matrix being diagonalised) and V
// (n x n, column-major, the accumulated rotations) in the workspace.
template <typename T>
__global__ __launch_bounds__(LA_THREADS) void syev_jacobi_kernel(int n, int upper, int vectors, T *__restrict__ A,
                                                                 int64_t lda, T *__restrict__ W, T *__restrict__ M,
                                                                 T *__restrict__ V, int32_t *__restrict__ info) {
    __shared__ T cs[SYEV_MAX_N / 2], sn[SYEV_MAX_N / 2];
    __shared__ int pp[SYEV_MAX_N / 2], qq[SYEV_MAX_N / 2];
    __shared__ int rank_of[SYEV_MAX_N];
    __shared__ T red[LA_WAVES];
    int const tid    = threadIdx.x;
    int64_t const nn = int64_t(n) * n;
    // the referenced triangle, mirrored
    T fro = T(0);
    for (int64_t e = tid; e < nn; e += LA_THREADS) {
        int const i = int(e % n), j = int(e / n);
        bool const in_tri = upper ? i <= j : i >= j;
        T const v         = in_tri ? A[i + int64_t(j) * lda] : A[j + int64_t(i) * lda];
        M[e]              = v;
        fro += v * v;
        if (vectors) V[e] = i == j ? T(1) : T(0);
    }
    fro = block_sum(fro, red); // also orders the writes above before the reads below
    T const eps   = sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920929e-7);
    T const limit = T(n) * T(n) * eps * eps * fro;
    int const m      = n + (n & 1); // an odd n gets a dummy index n that never rotates
    int const npairs = m / 2;
    bool conv        = false;
    for (int sweep = 0; sweep <= SYEV_MAX_SWEEPS; ++sweep) {
        T off = T(0);
        for (int64_t e = tid; e < nn; e += LA_THREADS)
            if (e % n != e / n) off += M[e] * M[e];
        off = block_sum(off, red);
        if (off <= limit) {
            conv = true;
            break;
        }
        if (sweep == SYEV_MAX_SWEEPS) break;
        for (int r = 0; r < m - 1; ++r) {
            // round-robin pairing: index m-1 fixed, the others rotate by one per step
            for (int i = tid; i < npairs; i += LA_THREADS) {
                int const a = i == 0 ? m - 1 : (r + i) % (m - 1);
                int const b = (r - i + (m - 1)) % (m - 1);
                int const p = a < b ? a : b, q = a < b ? b : a;
                T c = T(1), s = T(0);
                if (q < n) {
                    T const apq = M[p + int64_t(q) * n];
                    if (apq != T(0)) {
                        T const tau = (M[q + int64_t(q) * n] - M[p + int64_t(p) * n]) / (T(2) * apq);
                        T const t   = (tau >= T(0) ? T(1) : T(-1)) / (std::fabs(tau) + std::sqrt(T(1) + tau * tau));
                        c           = T(1) / std::sqrt(T(1) + t * t);
                        s           = t * c;
                    }
                }
                pp[i] = p;
                qq[i] = q;
                cs[i] = c;
                sn[i] = s;
            }
            __syncthreads();
            // rows p and q of M <- J^T M
            for (int64_t e = tid; e < int64_t(npairs) * n; e += LA_THREADS) {
                int const i = int(e / n), k = int(e % n);
                T const s   = sn[i];
                if (s == T(0)) continue;
                T const c    = cs[i];
                int64_t const xp = pp[i] + int64_t(k) * n, xq = qq[i] + int64_t(k) * n;
                T const x = M[xp], y = M[xq];
                M[xp]     = c * x - s * y;
                M[xq]     = s * x + c * y;
            }
            __syncthreads();
            // columns p and q of M <- M J, and of V <- V J
            for (int64_t e = tid; e < int64_t(npairs) * n; e += LA_THREADS) {
                int const i = int(e / n), k = int(e % n);
                T const s   = sn[i];
                if (s == T(0)) continue;
                T const c    = cs[i];
                int64_t const xp = k + int64_t(pp[i]) * n, xq = k + int64_t(qq[i]) * n;
                T const x = M[xp], y = M[xq];
                M[xp]     = c * x - s * y;
                M[xq]     = s * x + c * y;
                if (vectors) {
                    T const u = V[xp], v = V[xq];
                    V[xp]     = c * u - s * v;
                    V[xq]     = s * u + c * v;
                }
            }
            __syncthreads();
        }
    }
    // ascending order, ties by index (stable)
    for (int i = tid; i < n; i += LA_THREADS) {
        T const d = M[i + int64_t(i) * n];
        int pos   = 0;
        for (int j = 0; j < n; ++j) {
            T const e = M[j + int64_t(j) * n];
            pos += (e < d || (e == d && j < i)) ? 1 : 0;
        }
        rank_of[i] = pos;
        W[pos]     = d;
    }
    __syncthreads();
    if (vectors) {
        for (int64_t e = tid; e < nn; e += LA_THREADS) {
            int const k = int(e % n), j = int(e / n);
            A[k + int64_t(rank_of[j]) * lda] = V[e];
        }
    }
    if (tid == 0) *info = conv ? 0 : 1;
}

template <typename T>
__global__ __launch_bounds__(LA_THREADS) void gesv_kernel(int n, int nrhs, T *__restrict__ A, int64_t lda,
                                                          int64_t *__restrict__ ipiv, T *__restrict__ B, int64_t ldb,
                                                          int32_t *__restrict__ info) {
    __shared__ T bv[LA_WAVES];
    __shared__ int bi[LA_WAVES];
    int const tid  = threadIdx.x;
    int const lane = tid & 63;
    int const w    = tid >> 6;
    int first_zero = 0; // 1-based column of the first exactly zero pivot
    for (int j = 0; j < n; ++j) {
        // pivot: the first row of largest magnitude in column j at or below the diagonal
        T best  = T(-1);
        int row = n;
        for (int i = j + tid; i < n; i += LA_THREADS) {
            T const v = std::fabs(A[i + int64_t(j) * lda]);
            if (v > best) {
                best = v;
                row  = i;
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            T const ob  = __shfl_xor(best, off, 64);
            int const o = __shfl_xor(row, off, 64);
            if (ob > best || (ob == best && o < row)) {
                best = ob;
                row  = o;
            }
        }
        if (lane == 0) {
            bv[w] = best;
            bi[w] = row;
        }
        __syncthreads();
        best = bv[0];
        row  = bi[0];
        for (int k = 1; k < LA_WAVES; ++k)
            if (bv[k] > best || (bv[k] == best && bi[k] < row)) {
                best = bv[k];
                row  = bi[k];
            }
        __syncthreads();
        int const p = row < n ? row : j; // a column of NaNs compares false everywhere: keep row j
        if (tid == 0) ipiv[j] = p + 1;
        if (p != j) {
            for (int k = tid; k < n; k += LA_THREADS) {
                T const t               = A[j + int64_t(k) * lda];
                A[j + int64_t(k) * lda] = A[p + int64_t(k) * lda];
                A[p + int64_t(k) * lda] = t;
            }
        }
        __syncthreads();
        T const ujj = A[j + int64_t(j) * lda];
        if (ujj == T(0)) {
            if (first_zero == 0) first_zero = j + 1;
            continue; // LAPACK completes the factorization past an exactly singular pivot
        }
        for (int i = j + 1 + tid; i < n; i += LA_THREADS) A[i + int64_t(j) * lda] /= ujj;
        __syncthreads();
        int const t = n - j - 1;
        for (int64_t e = tid; e < int64_t(t) * t; e += LA_THREADS) {
            int const i = j + 1 + int(e % t), k = j + 1 + int(e / t);
            A[i + int64_t(k) * lda] -= A[i + int64_t(j) * lda] * A[j + int64_t(k) * lda];
        }
        __syncthreads();
    }
    if (first_zero == 0) {
        // P b (the interchanges in order), L y = P b (unit lower), then U x = y; B is left
        // untouched when the factor is singular, as by LAPACK
        for (int j = 0; j < n; ++j) {
            int const p = int(ipiv[j]) - 1;
            if (p != j) {
                for (int r = tid; r < nrhs; r += LA_THREADS) {
                    T const t               = B[j + int64_t(r) * ldb];
                    B[j + int64_t(r) * ldb] = B[p + int64_t(r) * ldb];
                    B[p + int64_t(r) * ldb] = t;
                }
            }
            __syncthreads();
        }
        for (int j = 0; j < n; ++j) {
            int const t = n - j - 1;
            for (int64_t e = tid; e < int64_t(t) * nrhs; e += LA_THREADS) {
                int const i = j + 1 + int(e % t), r = int(e / t);
                B[i + int64_t(r) * ldb] -= A[i + int64_t(j) * lda] * B[j + int64_t(r) * ldb];
            }
            __syncthreads();
        }
        for (int j = n - 1; j >= 0; --j) {
            for (int r = tid; r < nrhs; r += LA_THREADS) B[j + int64_t(r) * ldb] /= A[j + int64_t(j) * lda];
            __syncthreads();
            for (int64_t e = tid; e < int64_t(j) * nrhs; e += LA_THREADS) {
                int const i = int(e % j), r = int(e / j);
                B[i + int64_t(r) * ldb] -= A[i + int64_t(j) * lda] * B[j + int64_t(r) * ldb];
            }
            __syncthreads();
        }
    }
    if (tid == 0) *info = first_zero;
}

// n == 0: only the status (a kernel rather than a memset node, so that it can be captured)
__global__ void clear_info_kernel(int32_t *info) { *info = 0; }

} // namespace

template <typename T>
int syev_impl(char job, char uplo, int64_t n, T *A, int64_t lda, T *W, int32_t *info, hipStream_t stream) {
    bool const vectors = job == 'V' || job == 'v';
    if (!vectors && job != 'N' && job != 'n') return fail(-1, "syev: job must be 'V' or 'N'");
    bool const upper = uplo == 'U' || uplo == 'u';
    if (!upper && uplo != 'L' && uplo != 'l') return fail(-2, "syev: uplo must be 'U' or 'L'");
    if (n < 0) return fail(-3, "syev: n < 0");
    if (n > SYEV_MAX_N) return fail(EIN_ERROR_UNSUPPORTED, "syev: n above 1024 (a one-workgroup solver for small matrices)");
    if (lda < std::max<int64_t>(1, n)) return fail(-5, "syev: lda < max(1, n)");
    if (info == nullptr) return fail(-7, "syev: info must point to device memory");
    if (n == 0) {
        clear_info_kernel<<<1, 1, 0, stream>>>(info);
        EIN_LAUNCH_CHECK();
        return EIN_SUCCESS;
    }
    void *ws = nullptr;
    EIN_TRY(workspace(stream, size_t(2) * size_t(n) * size_t(n) * sizeof(T), &ws));
    T *M = static_cast<T *>(ws);
    syev_jacobi_kernel<T><<<1, LA_THREADS, 0, stream>>>(int(n), upper ? 1 : 0, vectors ? 1 : 0, A, lda, W, M,
                                                        M + n * n, info);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

template <typename T>
int gesv_impl(int64_t n, int64_t nrhs, T *A, int64_t lda, int64_t *ipiv, T *B, int64_t ldb, int32_t *info,
              hipStream_t stream) {
    if (n < 0) return fail(-1, "gesv: n < 0");
    if (nrhs < 0) return fail(-2, "gesv: nrhs < 0");
    if (lda < std::max<int64_t>(1, n)) return fail(-4, "gesv: lda < max(1, n)");
    if (ldb < std::max<int64_t>(1, n)) return fail(-7, "gesv: ldb < max(1, n)");
    if (info == nullptr) return fail(-8, "gesv: info must point to device memory");
    if (n > GESV_MAX_N || nrhs > (int64_t{1} << 30))
        return fail(EIN_ERROR_UNSUPPORTED, "gesv: n above 4096 (a one-workgroup solver for small systems)");
    if (n == 0) {
        clear_info_kernel<<<1, 1, 0, stream>>>(info);
        EIN_LAUNCH_CHECK();
        return EIN_SUCCESS;
    }
    gesv_kernel<T><<<1, LA_THREADS, 0, stream>>>(int(n), int(nrhs), A, lda, ipiv, B, ldb, info);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

} // namespace ein

extern "C" {

int ein_syev_f64(char job, char uplo, int64_t n, double *A, int64_t lda, double *W, int32_t *info,
                 ein_stream_t stream) {
    return ein::syev_impl<double>(job, uplo, n, A, lda, W, info, ein::as_stream(stream));
}
int ein_syev_f32(char job, char uplo, int64_t n, float *A, int64_t lda, float *W, int32_t *info, ein_stream_t stream) {
    return ein::syev_impl<float>(job, uplo, n, A, lda, W, info, ein::as_stream(stream));
}
int ein_gesv_f64(int64_t n, int64_t nrhs, double *A, int64_t lda, int64_t *ipiv, double *B, int64_t ldb, int32_t *info,
                 ein_stream_t stream) {
    return ein::gesv_impl<double>(n, nrhs, A, lda, ipiv, B, ldb, info, ein::as_stream(stream));
}
int ein_gesv_f32(int64_t n, int64_t nrhs, float *A, int64_t lda, int64_t *ipiv, float *B, int64_t ldb, int32_t *info,
                 ein_stream_t stream) {
    return ein::gesv_impl<float>(n, nrhs, A, lda, ipiv, B, ldb, info, ein::as_stream(stream));
}

} // extern "C"
