// Host-side interface of the bandwidth-bound leaves.
#pragma once

#include "ein_common.hpp"

namespace ein {

// Distinct (merged) axes of a strided view, innermost last.
struct TensorShape {
    int32_t rank;
    int64_t total;
    int64_t dims[EIN_MAX_RANK];
    int64_t strides[EIN_MAX_RANK];
};

int make_shape(ein_tensor_desc const &C, TensorShape &S);

template <typename T>
int gemv_impl(char trans, int64_t m, int64_t n, T alpha, T const *A, int64_t lda, T const *x, int64_t incx, T beta, T *y,
              int64_t incy, hipStream_t stream);
template <typename T>
int dot_impl(int64_t n, T alpha, T const *x, int64_t incx, T const *y, int64_t incy, T beta, T *result,
             hipStream_t stream);
template <typename T>
int ger_impl(int64_t m, int64_t n, T alpha, T const *x, int64_t incx, T const *y, int64_t incy, T *A, int64_t lda,
             hipStream_t stream);
template <typename T>
int dirprod_impl(int64_t n, T alpha, T const *x, int64_t incx, T const *y, int64_t incy, T beta, T *z, int64_t incz,
                 hipStream_t stream);
template <typename T>
int axpby_impl(int64_t n, T alpha, T const *x, int64_t incx, T beta, T *y, int64_t incy, hipStream_t stream);
template <typename T>
int scale_strided(int64_t n, T alpha, T *x, int64_t incx, hipStream_t stream);
template <typename T>
int tensor_scale(ein_tensor_desc const &C, T *c, T alpha, hipStream_t stream);

// C = beta*C + alpha*perm(A): C axis i is A axis perm[i] (permute.hip).
template <typename T>
int permute_tensor(int32_t rank, int64_t const *dims_A, int32_t const *perm, T alpha, T const *A,
                   int64_t const *strides_A, T beta, T *C, int64_t const *strides_C, hipStream_t stream);

} // namespace ein
