// Device tensors for the MI355X einsum backend.
//
// Reference types mirrored: DeviceTensor / DeviceTensorView (libs/Einsums/Tensor/include/
// Einsums/Tensor/DeviceTensor.hpp:327,543, Backends/DeviceTensor.hpp:62-113) -- always
// row-major on the device, sliced with Range/All. Storage comes from the backend's pooled
// allocator (ein_malloc); all work is queued on einsums::gpu::stream().
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "einsums_amd.h"

namespace einsums {

// ------------------------------------------------------------------ errors
struct dimension_error : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct rank_error : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct tensor_compat_error : std::logic_error {
    using std::logic_error::logic_error;
};
struct gpu_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// A call whose arguments cannot work together (reference: einsums::bad_logic, thrown by the
// batched einsum for lists of different lengths, Dispatch.hpp:1270).
struct bad_logic : std::logic_error {
    using std::logic_error::logic_error;
};

inline void check(int status, char const *what) {
    if (status == EIN_SUCCESS) return;
    std::string msg = std::string(what) + ": " + ein_last_error_string();
    if (status < 0) throw std::invalid_argument(msg);
    if (status == EIN_ERROR_DIMENSION) throw dimension_error(msg);
    if (status == EIN_ERROR_RANK) throw rank_error(msg);
    if (status == EIN_ERROR_VALUE) throw std::domain_error(msg);
    throw gpu_error(msg);
}

namespace gpu {
// The stream all device work of the calling thread is queued on (null stream by default).
inline ein_stream_t &stream() {
    thread_local ein_stream_t s = nullptr;
    return s;
}
inline void set_stream(ein_stream_t s) { stream() = s; }
inline void stream_wait() { check(ein_stream_synchronize(stream()), "stream_wait"); }
} // namespace gpu

namespace detail {
// Complex element types run as real operations on the real and imaginary parts: views of
// the same storage with every stride doubled (in real elements), the imaginary part one
// real element in.
template <typename T>
struct is_complex : std::false_type {};
template <typename R>
struct is_complex<std::complex<R>> : std::true_type {};
template <typename T>
inline constexpr bool is_complex_v = is_complex<T>::value;
template <typename T>
struct real_of {
    using type = T;
};
template <typename R>
struct real_of<std::complex<R>> {
    using type = R;
};
template <typename T>
using real_t = typename real_of<T>::type;

inline ein_tensor_desc part_desc(ein_tensor_desc d) {
    for (int32_t i = 0; i < d.rank; ++i) d.strides[i] *= 2;
    return d;
}

// C = beta*C + alpha*A over one real part, element for element (an identity permute).
template <typename R>
int part_axpby(ein_tensor_desc const &d, R alpha, R const *a, ein_tensor_desc const &da, R beta, R *c,
               ein_stream_t s) {
    int32_t perm[EIN_MAX_RANK];
    for (int32_t i = 0; i < d.rank; ++i) perm[i] = i;
    if constexpr (std::is_same_v<R, double>)
        return ein_permute_f64(d.rank, d.dims, perm, alpha, a, da.strides, beta, c, d.strides, s);
    else
        return ein_permute_f32(d.rank, d.dims, perm, alpha, a, da.strides, beta, c, d.strides, s);
}
} // namespace detail

struct Range {
    int64_t first, last; // [first, last)
    constexpr Range(int64_t f, int64_t l) : first(f), last(l) {}
};
struct AllT {};
inline constexpr AllT All{};

template <typename T, size_t Rank>
class DeviceTensorView;

// Common strided device view: pointer, extents, element strides.
template <typename T, size_t R>
class DeviceTensorBase {
  public:
    using ValueType                 = T;
    static constexpr size_t Rank    = R;
    static constexpr bool is_device = true;

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
    T *gpu_data() const { return ptr_; }
    std::string const &name() const { return name_; }
    void set_name(std::string n) { name_ = std::move(n); }

    bool is_contiguous() const {
        int64_t expect = 1;
        for (int d = int(R) - 1; d >= 0; --d) {
            if (dims_[d] != 1 && strides_[d] != expect) return false;
            expect *= dims_[d];
        }
        return true;
    }

    ein_tensor_desc desc(int32_t const *labels) const {
        ein_tensor_desc d{};
        d.rank = int32_t(R);
        for (size_t i = 0; i < R; ++i) {
            d.labels[i]  = labels ? labels[i] : int32_t(i);
            d.dims[i]    = dims_[i];
            d.strides[i] = strides_[i];
        }
        return d;
    }

    // C = value (value == 0) or C *= alpha over the view.
    void zero() { scale(T(0)); }
    void scale(T alpha) {
        auto d = desc(nullptr);
        if constexpr (detail::is_complex_v<T>) {
            scale_complex(alpha);
        } else if constexpr (std::is_same_v<T, double>) {
            check(ein_tensor_scale_f64(&d, ptr_, alpha, gpu::stream()), "scale");
        } else {
            check(ein_tensor_scale_f32(&d, ptr_, alpha, gpu::stream()), "scale");
        }
    }
    DeviceTensorBase &operator*=(T alpha) {
        scale(alpha);
        return *this;
    }

    // Synchronous host transfers (row-major order of the view).
    std::vector<T> to_host() const {
        std::vector<T> out(static_cast<size_t>(size()));
        if (size() == 0) return out;
        gpu::stream_wait();
        if (is_contiguous()) {
            check(hipMemcpy(out.data(), ptr_, out.size() * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess ? 0 : EIN_ERROR_HIP,
                  "to_host");
            return out;
        }
        // strided view: copy the enclosing span and gather on the host
        int64_t lo = 0, hi = 0;
        for (size_t i = 0; i < R; ++i) {
            int64_t e = (dims_[i] - 1) * strides_[i];
            (e < 0 ? lo : hi) += e;
        }
        std::vector<T> span(size_t(hi - lo + 1));
        check(hipMemcpy(span.data(), ptr_ + lo, span.size() * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess ? 0 : EIN_ERROR_HIP,
              "to_host");
        for (int64_t e = 0; e < size(); ++e) {
            int64_t off = 0, rem = e;
            for (int d = int(R) - 1; d >= 0; --d) {
                off += (rem % dims_[d]) * strides_[d];
                rem /= dims_[d];
            }
            out[size_t(e)] = span[size_t(off - lo)];
        }
        return out;
    }

    void from_host(std::vector<T> const &v) {
        if (int64_t(v.size()) != size()) throw dimension_error("from_host: size mismatch");
        if (!is_contiguous()) throw tensor_compat_error("from_host: destination view is not contiguous");
        gpu::stream_wait();
        check(hipMemcpy(ptr_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess ? 0 : EIN_ERROR_HIP,
              "from_host");
    }

    template <typename... Args>
    DeviceTensorView<T, R> operator()(Args... args) const;

  protected:
    // C = alpha*C for complex elements: a real alpha scales both parts; a complex one mixes
    // them, through a contiguous copy of the two parts.
    void scale_complex(T alpha) {
        using Re   = detail::real_t<T>;
        auto pd    = detail::part_desc(desc(nullptr));
        Re *re     = reinterpret_cast<Re *>(ptr_);
        Re *im     = re + 1;
        if (alpha.imag() == Re(0)) {
            if constexpr (std::is_same_v<Re, double>) {
                check(ein_tensor_scale_f64(&pd, re, alpha.real(), gpu::stream()), "scale");
                check(ein_tensor_scale_f64(&pd, im, alpha.real(), gpu::stream()), "scale");
            } else {
                check(ein_tensor_scale_f32(&pd, re, alpha.real(), gpu::stream()), "scale");
                check(ein_tensor_scale_f32(&pd, im, alpha.real(), gpu::stream()), "scale");
            }
            return;
        }
        int64_t const n = size();
        if (n == 0) return;
        void *tmp = nullptr;
        check(ein_malloc(&tmp, size_t(n) * sizeof(T)), "scale workspace");
        std::shared_ptr<void> hold(tmp, [](void *q) { (void)ein_free_async(q, gpu::stream()); });
        ein_tensor_desc td = desc(nullptr);
        int64_t st = 1;
        for (int d = int(R) - 1; d >= 0; --d) {
            td.strides[d] = st;
            st *= dims_[size_t(d)];
        }
        td       = detail::part_desc(td);
        Re *tre  = static_cast<Re *>(tmp);
        Re *tim  = tre + 1;
        auto run = [&](int status) { check(status, "scale"); };
        run(detail::part_axpby<Re>(td, Re(1), re, pd, Re(0), tre, gpu::stream()));
        run(detail::part_axpby<Re>(td, Re(1), im, pd, Re(0), tim, gpu::stream()));
        run(detail::part_axpby<Re>(pd, alpha.real(), tre, td, Re(0), re, gpu::stream()));
        run(detail::part_axpby<Re>(pd, -alpha.imag(), tim, td, Re(1), re, gpu::stream()));
        run(detail::part_axpby<Re>(pd, alpha.real(), tim, td, Re(0), im, gpu::stream()));
        run(detail::part_axpby<Re>(pd, alpha.imag(), tre, td, Re(1), im, gpu::stream()));
    }

    std::string name_;
    T *ptr_ = nullptr;
    std::array<int64_t, R> dims_{};
    std::array<int64_t, R> strides_{};
};

template <typename T, size_t Rank>
class DeviceTensorView : public DeviceTensorBase<T, Rank> {
  public:
    DeviceTensorView() = default;
    DeviceTensorView(std::string name, T *ptr, std::array<int64_t, Rank> dims, std::array<int64_t, Rank> strides) {
        this->name_    = std::move(name);
        this->ptr_     = ptr;
        this->dims_    = dims;
        this->strides_ = strides;
    }
};

// Slicing: each argument is an integer extent-preserving All, a Range, or (for views)
// nothing else; the rank is preserved (the reference's Range/All slicing).
template <typename T, size_t R>
template <typename... Args>
DeviceTensorView<T, R> DeviceTensorBase<T, R>::operator()(Args... args) const {
    static_assert(sizeof...(Args) == R, "one Range/All per axis");
    std::array<int64_t, R> dims = dims_, strides = strides_;
    T *p   = ptr_;
    size_t ax = 0;
    auto one = [&](auto a) {
        if constexpr (std::is_same_v<decltype(a), Range>) {
            if (a.first < 0 || a.last > dims_[ax] || a.first > a.last) throw std::out_of_range("Range out of bounds");
            p += a.first * strides_[ax];
            dims[ax] = a.last - a.first;
        }
        ++ax;
    };
    (one(args), ...);
    return DeviceTensorView<T, R>(name_ + " view", p, dims, strides);
}

template <typename T, size_t Rank>
class DeviceTensor : public DeviceTensorBase<T, Rank> {
  public:
    DeviceTensor() = default;

    template <typename... Dims>
        requires(sizeof...(Dims) == Rank && (std::is_integral_v<Dims> && ...))
    explicit DeviceTensor(std::string name, Dims... dims) {
        init(std::move(name), std::array<int64_t, Rank>{int64_t(dims)...});
    }
    DeviceTensor(std::string name, std::array<int64_t, Rank> dims) { init(std::move(name), dims); }
    DeviceTensor(std::string name, std::array<int64_t, Rank> dims, std::vector<T> const &host) {
        init(std::move(name), dims);
        this->from_host(host);
    }

    // Scalar (rank-0) conveniences.
    operator T() const
        requires(Rank == 0)
    {
        return this->to_host()[0];
    }
    DeviceTensor &operator=(T value)
        requires(Rank == 0)
    {
        this->from_host(std::vector<T>{value});
        return *this;
    }

    DeviceTensorView<T, Rank> view() const {
        return DeviceTensorView<T, Rank>(this->name_, this->ptr_, this->dims_, this->strides_);
    }

  private:
    void init(std::string name, std::array<int64_t, Rank> dims) {
        this->name_ = std::move(name);
        this->dims_ = dims;
        int64_t s   = 1;
        for (int d = int(Rank) - 1; d >= 0; --d) {
            this->strides_[size_t(d)] = s;
            s *= dims[size_t(d)];
        }
        size_t bytes = size_t(std::max<int64_t>(s, 1)) * sizeof(T);
        void *p      = nullptr;
        check(ein_malloc(&p, bytes), "DeviceTensor allocation");
        storage_   = std::shared_ptr<void>(p, [](void *q) { (void)ein_free(q); });
        this->ptr_ = static_cast<T *>(p);
    }
    std::shared_ptr<void> storage_;
};

} // namespace einsums
