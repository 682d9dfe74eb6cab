// Tensors on disk: DiskTensor<T, Rank>, the reference's HDF5-backed container
// (libs/Einsums/Tensor/include/Einsums/Tensor/DiskTensor.hpp:200-306): one chunked dataset per
// tensor, its dims in row-major order, created when absent and opened (extents and element
// type checked) when present. get() reads the whole tensor into a row-major host Tensor and
// put() writes one back. The reference's constructor takes an open HDF5 file handle; this one
// takes the file's path (the backend opens it per call, see einsums_amd.h).
//
// linear_algebra::gemm<TransA, TransB>(alpha, A, B, beta, &C) with C a DiskTensor streams C
// through the GPU in panels of whole leading slices: the out-of-core update of a v^4 CCD
// intermediate kept on disk. A and B are row-major device matrices.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <type_traits>

#include "Tensor.hpp"

namespace einsums {

template <typename T, size_t Rank>
class DiskTensor {
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>, "DiskTensor holds double or float");
    static_assert(Rank >= 1, "DiskTensor has rank 1 or more");

  public:
    using ValueType              = T;
    static constexpr size_t rank = Rank;

    template <typename... Dims>
        requires(sizeof...(Dims) == Rank && (std::is_integral_v<Dims> && ...))
    DiskTensor(std::string path, std::string name, Dims... dims)
        : DiskTensor(std::move(path), std::move(name), std::array<int64_t, Rank>{int64_t(dims)...}) {}

    DiskTensor(std::string path, std::string name, std::array<int64_t, Rank> dims)
        : path_(std::move(path)), name_(std::move(name)), dims_(dims) {
        if constexpr (std::is_same_v<T, double>)
            check(ein_disk_create_f64(path_.c_str(), name_.c_str(), int32_t(Rank), dims_.data()), "DiskTensor");
        else
            check(ein_disk_create_f32(path_.c_str(), name_.c_str(), int32_t(Rank), dims_.data()), "DiskTensor");
    }

    std::string const &path() const { return path_; }
    std::string const &name() const { return name_; }
    std::array<int64_t, Rank> const &dims() const { return dims_; }
    int64_t dim(size_t i) const { return dims_[i]; }
    int64_t size() const {
        int64_t s = 1;
        for (auto d : dims_) s *= d;
        return s;
    }

    // The whole tensor, read into a new row-major host tensor.
    Tensor<T, Rank> get() const {
        Tensor<T, Rank> t(true, name_, dims_);
        if constexpr (std::is_same_v<T, double>)
            check(ein_disk_read_f64(path_.c_str(), name_.c_str(), int32_t(Rank), dims_.data(), t.data()), "DiskTensor::get");
        else
            check(ein_disk_read_f32(path_.c_str(), name_.c_str(), int32_t(Rank), dims_.data(), t.data()), "DiskTensor::get");
        return t;
    }

    // Writes `t` (same dims, row-major and contiguous) over the whole tensor.
    void put(Tensor<T, Rank> const &t) {
        if (t.dims() != dims_) throw dimension_error("DiskTensor::put: dims differ");
        int64_t s = 1;
        for (int d = int(Rank) - 1; d >= 0; --d) {
            if (t.stride(size_t(d)) != s) throw tensor_compat_error("DiskTensor::put: tensor is not row-major contiguous");
            s *= dims_[size_t(d)];
        }
        if constexpr (std::is_same_v<T, double>)
            check(ein_disk_write_f64(path_.c_str(), name_.c_str(), int32_t(Rank), dims_.data(), t.data()), "DiskTensor::put");
        else
            check(ein_disk_write_f32(path_.c_str(), name_.c_str(), int32_t(Rank), dims_.data(), t.data()), "DiskTensor::put");
    }

  private:
    std::string path_, name_;
    std::array<int64_t, Rank> dims_;
};

namespace linear_algebra {

// C = alpha * op(A) op(B) + beta * C, row-major: op(A) is m x k, op(B) k x n, and C's leading
// dims multiply to m, the rest to n. Runs on gpu::stream() and returns when C is on disk.
template <bool TransA, bool TransB, typename T, size_t Rank>
void gemm(T alpha, DeviceTensor<T, 2> const &A, DeviceTensor<T, 2> const &B, T beta, DiskTensor<T, Rank> *C,
          size_t panel_bytes = 0) {
    if (A.stride(1) != 1 || B.stride(1) != 1) throw tensor_compat_error("gemm: A and B must be row-major");
    int64_t const m = TransA ? A.dim(1) : A.dim(0);
    int64_t const k = TransA ? A.dim(0) : A.dim(1);
    int64_t const n = TransB ? B.dim(0) : B.dim(1);
    if ((TransB ? B.dim(1) : B.dim(0)) != k || C->size() != m * n) throw dimension_error("gemm: shapes do not match");
    // row-major C = op(A) op(B) is column-major C^T = op(B)^T op(A)^T: the operands swap
    char const ta = TransA ? 'T' : 'N', tb = TransB ? 'T' : 'N';
    if constexpr (std::is_same_v<T, double>)
        check(ein_gemm_disk_c_f64(tb, ta, n, m, k, alpha, B.data(), B.stride(0), A.data(), A.stride(0), beta,
                                  C->path().c_str(), C->name().c_str(), panel_bytes, 0, 0, gpu::stream()),
              "gemm");
    else
        check(ein_gemm_disk_c_f32(tb, ta, n, m, k, alpha, B.data(), B.stride(0), A.data(), A.stride(0), beta,
                                  C->path().c_str(), C->name().c_str(), panel_bytes, gpu::stream()),
              "gemm");
}

} // namespace linear_algebra
} // namespace einsums
