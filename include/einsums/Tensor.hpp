// Host tensors: Tensor<T, Rank> and TensorView<T, Rank>, the reference's host containers
// (libs/Einsums/Tensor/include/Einsums/Tensor/Tensor.hpp:82 GeneralTensor / Tensor,
// :1092 Tensor<T, 0>, :1253 TensorView). Storage is a host std::vector; the layout is
// column-major unless asked for row-major, as the reference's default configuration
// (GlobalConfigMap "row-major" = false). An einsum or permute over host tensors runs on the
// GPU: the operands' memory spans are copied to device memory, the device path runs, and C
// is copied back (include/einsums/TensorAlgebra.hpp).
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "DeviceTensor.hpp"

namespace einsums {

namespace config {
// Layout of host tensors constructed without an explicit layout (the reference reads the
// "row-major" entry of its global configuration; default false = column-major).
inline bool &row_major_default() {
    static bool v = false;
    return v;
}
} // namespace config

template <typename T, size_t Rank>
class TensorView;

template <typename T, size_t R>
class HostTensorBase {
  public:
    using ValueType                   = T;
    static constexpr size_t Rank      = R;
    static constexpr bool is_device   = false;

    int64_t dim(size_t i) const { return dims_[i]; }
    int64_t stride(size_t i) const { return strides_[i]; }
    std::array<int64_t, R> const &dims() const { return dims_; }
    std::array<int64_t, R> const &strides() const { return strides_; }
    int64_t size() const {
        int64_t s = 1;
        for (auto d : dims_) s *= d;
        return s;
    }
    T *data() const { return ptr_; }
    std::string const &name() const { return name_; }
    void set_name(std::string n) { name_ = std::move(n); }
    bool is_row_major() const {
        if constexpr (R < 2) return true;
        else return strides_[R - 1] == 1;
    }

    template <typename... Idx>
        requires(sizeof...(Idx) == R && (std::is_integral_v<Idx> && ...))
    T &operator()(Idx... idx) const {
        std::array<int64_t, R> const i{int64_t(idx)...};
        int64_t off = 0;
        for (size_t d = 0; d < R; ++d) {
            int64_t v = i[d] < 0 ? i[d] + dims_[d] : i[d]; // negative indices count from the end
            if (v < 0 || v >= dims_[d]) throw std::out_of_range("Tensor index out of range");
            off += v * strides_[d];
        }
        return ptr_[off];
    }

    // Range / All per axis: a rank-preserving view.
    template <typename... Args>
        requires(sizeof...(Args) == R && ((std::is_same_v<Args, Range> || std::is_same_v<Args, AllT>) && ...))
    TensorView<T, R> operator()(Args... args) const;

    void set_all(T value) {
        for_each_offset([&](int64_t off) { ptr_[off] = value; });
    }
    void zero() { set_all(T(0)); }
    HostTensorBase &operator*=(T alpha) {
        for_each_offset([&](int64_t off) { ptr_[off] *= alpha; });
        return *this;
    }

    // Calls f(offset) for every element, last index fastest.
    template <typename F>
    void for_each_offset(F &&f) const {
        int64_t const n = size();
        std::array<int64_t, R> i{};
        int64_t off = 0;
        for (int64_t e = 0; e < n; ++e) {
            f(off);
            for (int d = int(R) - 1; d >= 0; --d) {
                off += strides_[size_t(d)];
                if (++i[size_t(d)] < dims_[size_t(d)]) break;
                off -= dims_[size_t(d)] * strides_[size_t(d)];
                i[size_t(d)] = 0;
            }
        }
    }

  protected:
    std::string name_;
    T *ptr_ = nullptr;
    std::array<int64_t, R> dims_{};
    std::array<int64_t, R> strides_{};
};

template <typename T, size_t Rank>
class TensorView : public HostTensorBase<T, Rank> {
  public:
    TensorView() = default;
    TensorView(std::string name, T *ptr, std::array<int64_t, Rank> dims, std::array<int64_t, Rank> strides) {
        this->name_    = std::move(name);
        this->ptr_     = ptr;
        this->dims_    = dims;
        this->strides_ = strides;
    }
};

template <typename T, size_t R>
template <typename... Args>
    requires(sizeof...(Args) == R && ((std::is_same_v<Args, Range> || std::is_same_v<Args, AllT>) && ...))
TensorView<T, R> HostTensorBase<T, R>::operator()(Args... args) const {
    std::array<int64_t, R> dims = dims_;
    T *p                        = ptr_;
    size_t ax                   = 0;
    auto one                    = [&](auto a) {
        if constexpr (std::is_same_v<decltype(a), Range>) {
            if (a.first < 0 || a.last > dims_[ax] || a.first > a.last) throw std::out_of_range("Range out of bounds");
            p += a.first * strides_[ax];
            dims[ax] = a.last - a.first;
        }
        ++ax;
    };
    (one(args), ...);
    return TensorView<T, R>(name_ + " view", p, dims, strides_);
}

template <typename T, size_t Rank>
class Tensor : public HostTensorBase<T, Rank> {
  public:
    Tensor() = default;

    template <typename... Dims>
        requires(sizeof...(Dims) == Rank && (std::is_integral_v<Dims> && ...))
    explicit Tensor(std::string name, Dims... dims) {
        init(config::row_major_default(), std::move(name), std::array<int64_t, Rank>{int64_t(dims)...});
    }
    template <typename... Dims>
        requires(sizeof...(Dims) == Rank && (std::is_integral_v<Dims> && ...))
    Tensor(bool row_major, std::string name, Dims... dims) {
        init(row_major, std::move(name), std::array<int64_t, Rank>{int64_t(dims)...});
    }
    Tensor(bool row_major, std::string name, std::array<int64_t, Rank> dims) { init(row_major, std::move(name), dims); }

    Tensor(Tensor const &o) { *this = o; }
    Tensor &operator=(Tensor const &o) {
        if (this == &o) return *this;
        this->name_    = o.name_;
        this->dims_    = o.dims_;
        this->strides_ = o.strides_;
        data_          = o.data_;
        this->ptr_     = data_.data();
        return *this;
    }
    Tensor(Tensor &&o) noexcept { *this = std::move(o); }
    Tensor &operator=(Tensor &&o) noexcept {
        this->name_    = std::move(o.name_);
        this->dims_    = o.dims_;
        this->strides_ = o.strides_;
        data_          = std::move(o.data_);
        this->ptr_     = data_.data();
        o.ptr_         = nullptr;
        return *this;
    }

    // Rank-0 tensors hold one value (reference Tensor<T, 0>, Tensor.hpp:1092).
    operator T() const
        requires(Rank == 0)
    {
        return data_[0];
    }
    Tensor &operator=(T value)
        requires(Rank == 0)
    {
        data_[0] = value;
        return *this;
    }

    std::vector<T> &vector_data() { return data_; }
    std::vector<T> const &vector_data() const { return data_; }

    TensorView<T, Rank> view() const { return TensorView<T, Rank>(this->name_, this->ptr_, this->dims_, this->strides_); }

  private:
    void init(bool row_major, std::string name, std::array<int64_t, Rank> dims) {
        this->name_ = std::move(name);
        this->dims_ = dims;
        int64_t s   = 1;
        if (row_major) {
            for (int d = int(Rank) - 1; d >= 0; --d) {
                this->strides_[size_t(d)] = s;
                s *= dims[size_t(d)];
            }
        } else {
            for (size_t d = 0; d < Rank; ++d) {
                this->strides_[d] = s;
                s *= dims[d];
            }
        }
        for (auto d : dims)
            if (d < 0) throw std::domain_error("Tensor: negative extent");
        data_.assign(size_t(std::max<int64_t>(s, 1)), T(0));
        this->ptr_ = data_.data();
    }
    std::vector<T> data_;
};

namespace detail {

// Device copy of a host view's memory span [data + lo, data + hi]: a device view with the
// same strides. `upload` copies the span; `download` copies it back (the span may include
// elements outside the view: they travel both ways unchanged).
template <typename T, size_t R>
struct StagedView {
    DeviceTensor<T, 1> buf;
    DeviceTensorView<T, R> view;
    T *host_lo = nullptr;
    int64_t span = 0;

    explicit StagedView(HostTensorBase<T, R> const &h, bool upload) {
        int64_t lo = 0, hi = 0;
        bool empty = false;
        for (size_t i = 0; i < R; ++i) {
            if (h.dim(i) == 0) empty = true;
            int64_t const e = (h.dim(i) - 1) * h.stride(i);
            (e < 0 ? lo : hi) += e;
        }
        span    = empty ? 0 : hi - lo + 1;
        buf     = DeviceTensor<T, 1>("staged", std::max<int64_t>(span, 1));
        host_lo = h.data() + lo;
        view    = DeviceTensorView<T, R>(h.name(), buf.data() - lo, h.dims(), h.strides());
        if (upload && span > 0) {
            gpu::stream_wait();
            check(hipMemcpy(buf.data(), host_lo, size_t(span) * sizeof(T), hipMemcpyHostToDevice) == hipSuccess
                      ? EIN_SUCCESS
                      : EIN_ERROR_HIP,
                  "host tensor upload");
        }
    }
    void download() {
        if (span == 0) return;
        gpu::stream_wait();
        check(hipMemcpy(host_lo, buf.data(), size_t(span) * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess ? EIN_SUCCESS
                                                                                                             : EIN_ERROR_HIP,
              "host tensor download");
    }
};

} // namespace detail

} // namespace einsums
