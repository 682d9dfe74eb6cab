// Device syev and gesv on DeviceTensors, with the reference's linear_algebra signatures:
//   syev<ComputeEigenvectors>(&A, &W)  libs/Einsums/LinearAlgebra/include/Einsums/LinearAlgebra.hpp:362
//   gesv(&A, &B) -> info               libs/Einsums/LinearAlgebra/include/Einsums/LinearAlgebra.hpp:495
// Both queue their kernels on einsums::gpu::stream() and then wait for it, because the
// reference returns info (gesv) or throws on a failed iteration (syev). The kernels themselves
// report info in device memory (ein_syev_* / ein_gesv_* of einsums_amd.h), for callers that
// stay asynchronous. DeviceTensors are row-major: the calls transpose through the device
// permute where the column-major solvers need it, so A(i, j) keeps its meaning.
#pragma once

#include <algorithm>
#include <array>
#include <memory>
#include <stdexcept>
#include <type_traits>

#include "DeviceTensor.hpp"

namespace einsums::linear_algebra {

namespace detail {

// A device int32 for the solver status, read back after the stream has drained.
struct DeviceInfo {
    DeviceInfo() { check(ein_malloc(&ptr, 64), "info allocation"); }
    ~DeviceInfo() { (void)ein_free(ptr); }
    DeviceInfo(DeviceInfo const &)            = delete;
    DeviceInfo &operator=(DeviceInfo const &) = delete;
    int32_t read() const {
        gpu::stream_wait();
        int32_t v = 0;
        check(hipMemcpy(&v, ptr, sizeof(v), hipMemcpyDeviceToHost) == hipSuccess ? 0 : EIN_ERROR_HIP, "info read");
        return v;
    }
    void *ptr = nullptr;
};

// dst = src with axes permuted (dst axis i = src axis perm[i]), through the device permute.
template <typename T>
void permute2(T const *src, std::array<int64_t, 2> sdims, std::array<int64_t, 2> ss, T *dst, std::array<int64_t, 2> ds,
              std::array<int32_t, 2> perm) {
    if constexpr (std::is_same_v<T, double>)
        check(ein_permute_f64(2, sdims.data(), perm.data(), 1.0, src, ss.data(), 0.0, dst, ds.data(), gpu::stream()),
              "permute");
    else
        check(ein_permute_f32(2, sdims.data(), perm.data(), 1.0f, src, ss.data(), 0.0f, dst, ds.data(), gpu::stream()),
              "permute");
}

// dst (rows x cols, strides ds) = src^T (cols x rows, strides ss).
template <typename T>
void transpose(T const *src, std::array<int64_t, 2> sdims, std::array<int64_t, 2> ss, T *dst,
               std::array<int64_t, 2> ds) {
    permute2<T>(src, sdims, ss, dst, ds, {1, 0});
}

} // namespace detail

// On exit A holds the orthonormal eigenvectors as columns (A(:, j) belongs to W(j)) and W the
// eigenvalues in ascending order; the upper triangle of A is read. Throws like the reference:
// dimension_error for a non-square A, tensor_compat_error when W's length differs,
// std::runtime_error when the iteration does not converge.
template <bool ComputeEigenvectors = true, typename AType, typename WType>
    requires(AType::Rank == 2 && WType::Rank == 1 && std::is_same_v<typename AType::ValueType, typename WType::ValueType>)
void syev(AType *A, WType *W) {
    using T         = typename AType::ValueType;
    int64_t const n = A->dim(0);
    if (A->dim(1) != n) throw dimension_error("syev: the matrix must be square");
    if (W->dim(0) != n) throw tensor_compat_error("syev: W must have as many entries as A has rows");
    if (n != 0 && W->stride(0) != 1) throw tensor_compat_error("syev: W must be contiguous");
    // a row-major n x n buffer: the solver reads it as column-major A^T = A, so the upper
    // triangle of A is its lower one, and its eigenvector columns land as rows
    DeviceTensor<T, 2> work("syev work", n, n);
    if (n > 0) detail::permute2<T>(A->data(), A->dims(), A->strides(), work.data(), work.strides(), {0, 1});
    detail::DeviceInfo info;
    char const job = ComputeEigenvectors ? 'V' : 'N';
    if constexpr (std::is_same_v<T, double>)
        check(ein_syev_f64(job, 'L', n, work.data(), std::max<int64_t>(1, n), W->data(),
                           static_cast<int32_t *>(info.ptr), gpu::stream()),
              "syev");
    else
        check(ein_syev_f32(job, 'L', n, work.data(), std::max<int64_t>(1, n), W->data(),
                           static_cast<int32_t *>(info.ptr), gpu::stream()),
              "syev");
    if constexpr (ComputeEigenvectors) {
        // work row j = eigenvector j; A(i, j) = work(j, i)
        if (n > 0) detail::transpose<T>(work.data(), {n, n}, {n, 1}, A->data(), A->strides());
    }
    if (info.read() != 0) throw std::runtime_error("syev: the Jacobi iteration did not converge");
}

// Solves A X = B: A is overwritten with its LU factors (unit L below the diagonal, U on and
// above, rows in pivoted order), B (n, or n x nrhs) with X. Returns LAPACK's info: 0, or
// i > 0 when U(i, i) is exactly zero (then B is left as it was).
template <typename AType, typename BType>
    requires(AType::Rank == 2 && (BType::Rank == 1 || BType::Rank == 2) &&
             std::is_same_v<typename AType::ValueType, typename BType::ValueType>)
int gesv(AType *A, BType *B) {
    using T         = typename AType::ValueType;
    int64_t const n = A->dim(0);
    if (A->dim(1) != n) throw dimension_error("gesv: the matrix must be square");
    if (B->dim(0) != n) throw dimension_error("gesv: B must have as many rows as A");
    int64_t const nrhs = BType::Rank == 2 ? B->dim(1) : 1;
    // column-major copies: At(j, i) = A(i, j) row-major is A column-major, same for B
    DeviceTensor<T, 2> At("gesv A", n, n);
    DeviceTensor<T, 2> Bt("gesv B", nrhs, n);
    if (n > 0) {
        detail::transpose<T>(A->data(), A->dims(), A->strides(), At.data(), At.strides());
        if constexpr (BType::Rank == 2)
            detail::transpose<T>(B->data(), B->dims(), B->strides(), Bt.data(), Bt.strides());
        else
            detail::transpose<T>(B->data(), {n, 1}, {B->stride(0), 1}, Bt.data(), Bt.strides());
    }
    void *ipiv = nullptr;
    check(ein_malloc(&ipiv, size_t(std::max<int64_t>(n, 1)) * sizeof(int64_t)), "gesv ipiv");
    std::shared_ptr<void> keep(ipiv, [](void *p) { (void)ein_free(p); });
    detail::DeviceInfo info;
    int64_t const ld = std::max<int64_t>(1, n);
    if constexpr (std::is_same_v<T, double>)
        check(ein_gesv_f64(n, nrhs, At.data(), ld, static_cast<int64_t *>(ipiv), Bt.data(), ld,
                           static_cast<int32_t *>(info.ptr), gpu::stream()),
              "gesv");
    else
        check(ein_gesv_f32(n, nrhs, At.data(), ld, static_cast<int64_t *>(ipiv), Bt.data(), ld,
                           static_cast<int32_t *>(info.ptr), gpu::stream()),
              "gesv");
    if (n > 0) {
        detail::transpose<T>(At.data(), {n, n}, {n, 1}, A->data(), A->strides());
        if constexpr (BType::Rank == 2)
            detail::transpose<T>(Bt.data(), {nrhs, n}, {n, 1}, B->data(), B->strides());
        else
            detail::transpose<T>(Bt.data(), {1, n}, {n, 1}, B->data(), std::array<int64_t, 2>{B->stride(0), 1});
    }
    return info.read();
}

} // namespace einsums::linear_algebra
