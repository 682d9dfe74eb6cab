// einsum / permute on device tensors with compile-time pattern analysis.
//
// Reference API kept: einsum(C_prefactor, Indices{...}, &C, AB_prefactor, Indices{...}, A,
// Indices{...}, B, &algorithm_choice), the default-prefactor form, the smart-pointer form and
// the batched form over collections
// (libs/Einsums/TensorAlgebra/include/Einsums/TensorAlgebra/TensorAlgebra.hpp:48-146), with
// device tensors, host tensors (staged through device memory) or a scalar C,
// permute(C_prefactor, Indices{...}, &C, A_prefactor, Indices{...}, A) (Permute.hpp:330-407),
// and detail::AlgorithmChoice (Detail/Utilities.hpp:26-34).
//
// The label sets are analysed at compile time: every label tag carries a constexpr id and
// the predicates of Dispatch.hpp:194-262,407-439,565-609 (precedence :894-955) are
// evaluated as constexpr functions over fixed-capacity id lists. The chosen algorithm is a
// constant of each instantiation; at run time the call only builds the strided views and
// hands them to the backend (ein_einsum_execute_*), which folds the index groups and falls
// back to the generic kernel when a view does not fold.
#pragma once

#include <tuple>
#include <type_traits>

#include "DeviceTensor.hpp"
#include "Index.hpp"
#include "Tensor.hpp"

namespace einsums::tensor_algebra {

namespace detail {

enum AlgorithmChoice { GENERIC, DOT, DIRECT, GER, GEMV, GEMM, INDETERMINATE };

inline char const *algorithm_name(AlgorithmChoice a) {
    constexpr char const *names[] = {"GENERIC", "DOT", "DIRECT", "GER", "GEMV", "GEMM", "INDETERMINATE"};
    return names[int(a)];
}

// Fixed-capacity list of label ids (or (label, position) pairs) usable in constexpr code.
struct LList {
    int32_t v[48] = {};
    int32_t p[48] = {};
    int n         = 0;
    constexpr void push(int32_t x, int32_t pos = -1) {
        v[n] = x;
        p[n] = pos;
        ++n;
    }
    constexpr bool has(int32_t x) const {
        for (int i = 0; i < n; ++i)
            if (v[i] == x) return true;
        return false;
    }
    constexpr bool same_labels(LList const &o) const {
        if (n != o.n) return false;
        for (int i = 0; i < n; ++i)
            if (v[i] != o.v[i]) return false;
        return true;
    }
};

template <typename... Ls>
constexpr LList make_list() {
    LList l;
    (l.push(std::remove_cvref_t<Ls>::id), ...);
    return l;
}

// IntersectT: elements of a that are also in b (order of a).
constexpr LList intersect(LList const &a, LList const &b) {
    LList o;
    for (int i = 0; i < a.n; ++i)
        if (b.has(a.v[i])) o.push(a.v[i]);
    return o;
}
// DifferenceT: elements of a not in b.
constexpr LList difference(LList const &a, LList const &b) {
    LList o;
    for (int i = 0; i < a.n; ++i)
        if (!b.has(a.v[i])) o.push(a.v[i]);
    return o;
}
// UniqueT: first occurrences.
constexpr LList unique(LList const &a) {
    LList o;
    for (int i = 0; i < a.n; ++i)
        if (!o.has(a.v[i])) o.push(a.v[i]);
    return o;
}
// find_type_with_position: for each label of a, every position where it occurs in b.
constexpr LList find_with_pos(LList const &a, LList const &b) {
    LList o;
    for (int i = 0; i < a.n; ++i)
        for (int j = 0; j < b.n; ++j)
            if (a.v[i] == b.v[j]) o.push(a.v[i], j);
    return o;
}
// contiguous_positions
constexpr bool contiguous(LList const &x) {
    for (int i = 0; i + 1 < x.n; ++i)
        if (x.p[i] != x.p[i + 1] - 1) return false;
    return true;
}
// is_same_ordering (false for empty lists, as in the reference)
constexpr bool same_ordering(LList const &x, LList const &y) {
    if (x.n == 0 || y.n == 0 || x.n != y.n) return false;
    for (int i = 0; i < x.n; ++i)
        if (x.v[i] != y.v[i]) return false;
    return true;
}

constexpr bool is_matrix_vector(LList const &C, LList const &X, LList const &Y) {
    LList Cu = unique(C);
    LList links = unique(difference(intersect(X, Y), C));
    LList lX = find_with_pos(links, X), lY = find_with_pos(links, Y);
    LList tX = find_with_pos(Cu, X), tY = find_with_pos(Cu, Y);
    LList X_in_C = find_with_pos(X, C), Y_in_C = find_with_pos(Y, C);
    return contiguous(lX) && contiguous(lY) && contiguous(tX) && same_ordering(lX, lY) && same_ordering(tX, X_in_C) &&
           !same_ordering(tY, Y_in_C) && Y_in_C.n == 0;
}

// detail::einsum precedence for real basic tensors (Dispatch.hpp:894-955).
constexpr AlgorithmChoice classify(LList const &C, LList const &A, LList const &B) {
    if (unique(C).n != C.n || unique(A).n != A.n || unique(B).n != B.n) return GENERIC; // Hadamard
    if (C.n == 0 && A.same_labels(B)) return DOT;
    if (C.same_labels(A) && C.same_labels(B)) return DIRECT;
    LList linksAB = intersect(A, B);
    LList Cu = unique(C);
    LList tA = find_with_pos(Cu, A), tB = find_with_pos(Cu, B);
    if (linksAB.n == 0 && contiguous(tA) && contiguous(tB)) return GER;
    if (is_matrix_vector(C, A, B) || is_matrix_vector(C, B, A)) return GEMV;
    if (C.n >= 2 && A.n >= 2 && B.n >= 2) {
        LList links = unique(difference(linksAB, C));
        LList lA = find_with_pos(links, A), lB = find_with_pos(links, B);
        LList A_in_C = find_with_pos(A, C), B_in_C = find_with_pos(B, C);
        bool cond = difference(C, A).n > 0 && difference(C, B).n > 0 && contiguous(lA) && contiguous(lB) &&
                    contiguous(tA) && contiguous(tB) && contiguous(A_in_C) && contiguous(B_in_C) &&
                    same_ordering(lA, lB) && same_ordering(tA, A_in_C) && same_ordering(tB, B_in_C);
        if (cond) return GEMM;
    }
    return GENERIC;
}

template <typename... Cs, typename... As, typename... Bs>
constexpr AlgorithmChoice classify(std::tuple<Cs...> const &, std::tuple<As...> const &, std::tuple<Bs...> const &) {
    return classify(make_list<Cs...>(), make_list<As...>(), make_list<Bs...>());
}

template <typename... Ls>
constexpr std::array<int32_t, sizeof...(Ls) == 0 ? 1 : sizeof...(Ls)> label_ids(std::tuple<Ls...> const &) {
    if constexpr (sizeof...(Ls) == 0)
        return {0};
    else
        return {std::remove_cvref_t<Ls>::id...};
}

template <typename T>
concept StridedTensor = requires(T const &t) {
    typename T::ValueType;
    { T::Rank } -> std::convertible_to<size_t>;
    { T::is_device } -> std::convertible_to<bool>;
    t.data();
    t.dims();
    t.strides();
};
// Tensors whose memory is on the GPU (DeviceTensor, DeviceTensorView, block views).
template <typename T>
concept DeviceTensorConcept = StridedTensor<T> && T::is_device;
// Host tensors (Tensor, TensorView): their einsums stage through device memory.
template <typename T>
concept HostTensorConcept = StridedTensor<T> && !T::is_device;

// Smart pointers to tensors (reference concept SmartPointer, Concepts/SmartPointer.hpp).
template <typename T>
concept SmartPointer = requires(T const &p) {
    typename T::element_type;
    { p.get() } -> std::convertible_to<typename T::element_type *>;
};
// Collections of tensors for the batched einsum (reference concept Container).
template <typename T>
concept Container = !StridedTensor<T> && !SmartPointer<T> && requires(T const &c) {
    typename T::value_type;
    { c.size() } -> std::convertible_to<size_t>;
    c.at(size_t{0});
};

template <typename T>
struct value_of {
    using type = T; // scalars
};
template <StridedTensor T>
struct value_of<T> {
    using type = typename T::ValueType;
};
template <SmartPointer T>
struct value_of<T> {
    using type = typename value_of<typename T::element_type>::type;
};
template <Container T>
struct value_of<T> {
    using type = typename value_of<typename T::value_type>::type;
};
template <typename T>
using value_of_t = typename value_of<std::remove_cvref_t<T>>::type;

template <typename T>
decltype(auto) deref(T const &t) {
    if constexpr (SmartPointer<T>) return *t;
    else return (t);
}
template <typename T>
auto deref_c(T *p) {
    if constexpr (SmartPointer<T>) return p->get();
    else return p;
}

// C = c*C + a * sum op(A) op(B) for complex elements (op = conj where ConjA / ConjB), as
// real einsums on the parts: with A = Ar + i sA Ai and B = Br + i sB Bi (s = -1 for a
// conjugated operand), Re P = Ar Br - sA sB Ai Bi and Im P = sB Ar Bi + sA Ai Br; the terms
// accumulate into the parts of C with beta = 1 after the first. A complex c first scales C.
template <typename R>
int real_einsum(int32_t alg, R cp, ein_tensor_desc const &cd, R *c, R abp, ein_tensor_desc const &ad, R const *a,
                ein_tensor_desc const &bd, R const *b, int32_t *used) {
    if constexpr (std::is_same_v<R, double>)
        return ein_einsum_execute_f64(alg, cp, &cd, c, abp, &ad, a, &bd, b, used, gpu::stream());
    else
        return ein_einsum_execute_f32(alg, cp, &cd, c, abp, &ad, a, &bd, b, used, gpu::stream());
}

template <bool ConjA, bool ConjB, typename CType, typename T>
void complex_einsum(int32_t alg, T c, ein_tensor_desc const &cd, CType *C, T a, ein_tensor_desc const &ad, T const *A,
                    ein_tensor_desc const &bd, T const *B, int32_t *used) {
    using R            = ::einsums::detail::real_t<T>;
    R const sA         = ConjA ? R(-1) : R(1);
    R const sB         = ConjB ? R(-1) : R(1);
    ein_tensor_desc pc = ::einsums::detail::part_desc(cd), pa = ::einsums::detail::part_desc(ad), pb = ::einsums::detail::part_desc(bd);
    R *Cr              = reinterpret_cast<R *>(C->data());
    R const *Ar        = reinterpret_cast<R const *>(A);
    R const *Br        = reinterpret_cast<R const *>(B);
    R beta             = c.real();
    if (c.imag() != R(0)) {
        C->scale(c);
        beta = R(1);
    }
    struct Term {
        R pre;
        R const *x;
        R const *y;
    };
    Term const re[4] = {{a.real(), Ar, Br}, {-a.real() * sA * sB, Ar + 1, Br + 1}, {-a.imag() * sB, Ar, Br + 1},
                        {-a.imag() * sA, Ar + 1, Br}};
    Term const im[4] = {{a.real() * sB, Ar, Br + 1}, {a.real() * sA, Ar + 1, Br}, {a.imag(), Ar, Br},
                        {-a.imag() * sA * sB, Ar + 1, Br + 1}};
    for (int part = 0; part < 2; ++part) {
        Term const *terms = part == 0 ? re : im;
        R *target         = Cr + part;
        for (int t = 0; t < 4; ++t) {
            if (t > 0 && terms[t].pre == R(0)) continue;
            check(real_einsum<R>(alg, t == 0 ? beta : R(1), pc, target, terms[t].pre, pa, terms[t].x, pb, terms[t].y,
                                 used),
                  "einsum");
        }
    }
}

} // namespace detail

template <bool ConjA = false, bool ConjB = false, detail::DeviceTensorConcept CType, detail::DeviceTensorConcept AType,
          detail::DeviceTensorConcept BType, typename U, typename... CIndices, typename... AIndices, typename... BIndices>
void einsum(U const C_prefactor, std::tuple<CIndices...> const &C_indices, CType *C, U const AB_prefactor,
            std::tuple<AIndices...> const &A_indices, AType const &A, std::tuple<BIndices...> const &B_indices,
            BType const &B, detail::AlgorithmChoice *algorithm_choice = nullptr) {
    using T = typename CType::ValueType;
    static_assert(std::is_same_v<T, typename AType::ValueType> && std::is_same_v<T, typename BType::ValueType>,
                  "mixed element types are not supported on the device path");
    static_assert(std::is_same_v<::einsums::detail::real_t<T>, double> || std::is_same_v<::einsums::detail::real_t<T>, float>,
                  "double, float and their complex types only");
    static_assert(::einsums::detail::is_complex_v<T> || (!ConjA && !ConjB), "conjugation applies to complex tensors");
    static_assert(sizeof...(CIndices) == CType::Rank, "Rank of C does not match Indices given for C.");
    static_assert(sizeof...(AIndices) == AType::Rank, "Rank of A does not match Indices given for A.");
    static_assert(sizeof...(BIndices) == BType::Rank, "Rank of B does not match Indices given for B.");

    constexpr detail::AlgorithmChoice alg = detail::classify(std::tuple<CIndices...>{}, std::tuple<AIndices...>{},
                                                             std::tuple<BIndices...>{});
    constexpr auto c_ids = detail::label_ids(std::tuple<CIndices...>{});
    constexpr auto a_ids = detail::label_ids(std::tuple<AIndices...>{});
    constexpr auto b_ids = detail::label_ids(std::tuple<BIndices...>{});
    (void)C_indices;
    (void)A_indices;
    (void)B_indices;
    ein_tensor_desc cd = C->desc(c_ids.data()), ad = A.desc(a_ids.data()), bd = B.desc(b_ids.data());
    int32_t used = -1;
    if constexpr (::einsums::detail::is_complex_v<T>) {
        detail::complex_einsum<ConjA, ConjB>(int32_t(alg), T(C_prefactor), cd, C, T(AB_prefactor), ad, A.data(), bd,
                                             B.data(), &used);
    } else {
        int st;
        if constexpr (std::is_same_v<T, double>)
            st = ein_einsum_execute_f64(int32_t(alg), double(C_prefactor), &cd, C->data(), double(AB_prefactor), &ad,
                                        A.data(), &bd, B.data(), &used, gpu::stream());
        else
            st = ein_einsum_execute_f32(int32_t(alg), float(C_prefactor), &cd, C->data(), float(AB_prefactor), &ad,
                                        A.data(), &bd, B.data(), &used, gpu::stream());
        check(st, "einsum");
    }
    if (algorithm_choice != nullptr) *algorithm_choice = detail::AlgorithmChoice(used);
}

// Host tensors (Tensor / TensorView): the operands' memory spans go to device memory, the
// device path runs, and C comes back; the call returns with C updated on the host.
template <bool ConjA = false, bool ConjB = false, detail::HostTensorConcept CType, detail::HostTensorConcept AType,
          detail::HostTensorConcept BType, typename U, typename... CIndices, typename... AIndices, typename... BIndices>
void einsum(U const C_prefactor, std::tuple<CIndices...> const &C_indices, CType *C, U const AB_prefactor,
            std::tuple<AIndices...> const &A_indices, AType const &A, std::tuple<BIndices...> const &B_indices,
            BType const &B, detail::AlgorithmChoice *algorithm_choice = nullptr) {
    using T = typename CType::ValueType;
    ::einsums::detail::StagedView<T, CType::Rank> c(*C, C_prefactor != U(0));
    ::einsums::detail::StagedView<typename AType::ValueType, AType::Rank> a(A, true);
    ::einsums::detail::StagedView<typename BType::ValueType, BType::Rank> b(B, true);
    einsum<ConjA, ConjB>(C_prefactor, C_indices, &c.view, AB_prefactor, A_indices, a.view, B_indices, b.view,
                         algorithm_choice);
    c.download();
}

// Scalar C (a plain T*, as the reference accepts for a rank-0 target, Dispatch.hpp:900-920):
// C = C_prefactor * C + AB_prefactor * (full contraction of A and B). Reads the result back
// (a synchronisation of the calling thread's stream).
template <bool ConjA = false, bool ConjB = false, typename CType, typename AType, typename BType, typename U,
          typename... AIndices, typename... BIndices>
    requires(std::is_arithmetic_v<CType> &&
             ((detail::DeviceTensorConcept<AType> && detail::DeviceTensorConcept<BType>) ||
              (detail::HostTensorConcept<AType> && detail::HostTensorConcept<BType>)))
void einsum(U const C_prefactor, std::tuple<> const &C_indices, CType *C, U const AB_prefactor,
            std::tuple<AIndices...> const &A_indices, AType const &A, std::tuple<BIndices...> const &B_indices,
            BType const &B, detail::AlgorithmChoice *algorithm_choice = nullptr) {
    if constexpr (detail::DeviceTensorConcept<AType>) {
        DeviceTensor<CType, 0> c("scalar C");
        if (C_prefactor != U(0)) c = *C;
        einsum<ConjA, ConjB>(C_prefactor, C_indices, &c, AB_prefactor, A_indices, A, B_indices, B, algorithm_choice);
        *C = CType(c);
    } else {
        Tensor<CType, 0> c("scalar C");
        c = *C;
        einsum<ConjA, ConjB>(C_prefactor, C_indices, &c, AB_prefactor, A_indices, A, B_indices, B, algorithm_choice);
        *C = CType(c);
    }
}

// Smart pointers to any of the operands (TensorAlgebra.hpp:116-132): dereferenced, then the
// matching overload above runs. C is passed as a pointer to its smart pointer.
template <bool ConjA = false, bool ConjB = false, typename AType, typename BType, typename CType, typename U,
          typename... CIndices, typename... AIndices, typename... BIndices>
    requires(detail::SmartPointer<AType> || detail::SmartPointer<BType> || detail::SmartPointer<CType>)
void einsum(U const C_prefactor, std::tuple<CIndices...> const &C_indices, CType *C, U const AB_prefactor,
            std::tuple<AIndices...> const &A_indices, AType const &A, std::tuple<BIndices...> const &B_indices,
            BType const &B, detail::AlgorithmChoice *algorithm_choice = nullptr) {
    einsum<ConjA, ConjB>(C_prefactor, C_indices, detail::deref_c(C), AB_prefactor, A_indices, detail::deref(A), B_indices,
                         detail::deref(B), algorithm_choice);
}

// Batched einsum over equally long collections (TensorAlgebra.hpp:57, Dispatch.hpp:1264-1286):
// entry i computes C[i] from A[i] and B[i]; the calls are queued back to back on the
// stream. *algorithm_choice (when given) is the first entry's.
template <bool ConjA = false, bool ConjB = false, detail::Container CType, detail::Container AType,
          detail::Container BType, typename CPrefactorType, typename ABPrefactorType, typename... CIndices,
          typename... AIndices, typename... BIndices>
void einsum(CPrefactorType const C_prefactor, std::tuple<CIndices...> const &C_indices, CType *C_list,
            ABPrefactorType const AB_prefactor, std::tuple<AIndices...> const &A_indices, AType const &A_list,
            std::tuple<BIndices...> const &B_indices, BType const &B_list,
            detail::AlgorithmChoice *algorithm_choice = nullptr) {
    if (C_list->size() != A_list.size() || C_list->size() != B_list.size())
        throw bad_logic("Lists passed to batched einsum call do not have the same size!");
    using T = detail::value_of_t<CType>;
    for (size_t i = 0; i < C_list->size(); ++i) {
        detail::AlgorithmChoice used = detail::INDETERMINATE;
        einsum<ConjA, ConjB>(T(C_prefactor), C_indices, &(C_list->at(i)), T(AB_prefactor), A_indices, A_list.at(i),
                             B_indices, B_list.at(i), &used);
        if (i == 0 && algorithm_choice != nullptr) *algorithm_choice = used;
    }
}

// Default prefactors: C = A*B (C_prefactor 0, AB_prefactor 1), for every operand kind above.
template <bool ConjA = false, bool ConjB = false, typename CType, typename AType, typename BType, typename... CIndices,
          typename... AIndices, typename... BIndices>
void einsum(std::tuple<CIndices...> const &C_indices, CType *C, std::tuple<AIndices...> const &A_indices, AType const &A,
            std::tuple<BIndices...> const &B_indices, BType const &B,
            detail::AlgorithmChoice *algorithm_choice = nullptr) {
    using T = detail::value_of_t<CType>;
    einsum<ConjA, ConjB>(T(0), C_indices, C, T(1), A_indices, A, B_indices, B, algorithm_choice);
}

// C = C_prefactor*C + A_prefactor*perm(A); the permutation is resolved at compile time.
template <detail::DeviceTensorConcept CType, detail::DeviceTensorConcept AType, typename U, typename... CIndices,
          typename... AIndices>
void permute(U const C_prefactor, std::tuple<CIndices...> const &, CType *C, U const A_prefactor,
             std::tuple<AIndices...> const &, AType const &A) {
    using T = typename CType::ValueType;
    static_assert(std::is_same_v<T, typename AType::ValueType>, "permute: element types differ");
    constexpr size_t R = sizeof...(AIndices);
    static_assert(sizeof...(CIndices) == R && CType::Rank == R && AType::Rank == R, "permute: rank mismatch");
    constexpr auto perm = [] {
        constexpr std::array<int32_t, R == 0 ? 1 : R> a{std::remove_cvref_t<AIndices>::id...};
        constexpr std::array<int32_t, R == 0 ? 1 : R> c{std::remove_cvref_t<CIndices>::id...};
        std::array<int32_t, R == 0 ? 1 : R> p{};
        for (size_t i = 0; i < R; ++i) {
            p[i] = -1;
            for (size_t j = 0; j < R; ++j)
                if (a[j] == c[i]) p[i] = int32_t(j);
        }
        return p;
    }();
    for (size_t i = 0; i < R; ++i)
        if (perm[i] < 0) throw std::invalid_argument("permute: index sets differ");
    std::array<int64_t, R == 0 ? 1 : R> dA{}, sA{}, sC{};
    for (size_t i = 0; i < R; ++i) {
        dA[i] = A.dim(i);
        sA[i] = A.stride(i);
        sC[i] = C->stride(i);
        if (C->dim(i) != A.dim(size_t(perm[i]))) throw dimension_error("permute: extents do not match");
    }
    if constexpr (::einsums::detail::is_complex_v<T>) {
        // the parts move independently; complex prefactors mix them
        using Re  = ::einsums::detail::real_t<T>;
        T const c = T(C_prefactor), a = T(A_prefactor);
        Re beta   = c.real();
        if (c.imag() != Re(0)) {
            C->scale(c);
            beta = Re(1);
        }
        std::array<int64_t, R == 0 ? 1 : R> pA{}, pC{};
        for (size_t i = 0; i < R; ++i) {
            pA[i] = 2 * sA[i];
            pC[i] = 2 * sC[i];
        }
        Re const *Ar = reinterpret_cast<Re const *>(A.data());
        Re *Cr       = reinterpret_cast<Re *>(C->data());
        auto run     = [&](Re alpha, Re const *x, Re b, Re *y) {
            int rc;
            if constexpr (std::is_same_v<Re, double>)
                rc = ein_permute_f64(int32_t(R), dA.data(), perm.data(), alpha, x, pA.data(), b, y, pC.data(), gpu::stream());
            else
                rc = ein_permute_f32(int32_t(R), dA.data(), perm.data(), alpha, x, pA.data(), b, y, pC.data(), gpu::stream());
            check(rc, "permute");
        };
        run(a.real(), Ar, beta, Cr); // Re C = beta Re C + a_r Re A - a_i Im A
        if (a.imag() != Re(0)) run(-a.imag(), Ar + 1, Re(1), Cr);
        run(a.real(), Ar + 1, beta, Cr + 1); // Im C = beta Im C + a_r Im A + a_i Re A
        if (a.imag() != Re(0)) run(a.imag(), Ar, Re(1), Cr + 1);
    } else {
        int st;
        if constexpr (std::is_same_v<T, double>)
            st = ein_permute_f64(int32_t(R), dA.data(), perm.data(), double(A_prefactor), A.data(), sA.data(),
                                 double(C_prefactor), C->data(), sC.data(), gpu::stream());
        else
            st = ein_permute_f32(int32_t(R), dA.data(), perm.data(), float(A_prefactor), A.data(), sA.data(),
                                 float(C_prefactor), C->data(), sC.data(), gpu::stream());
        check(st, "permute");
    }
}

// Host tensors: staged through device memory like the host einsum.
template <detail::HostTensorConcept CType, detail::HostTensorConcept AType, typename U, typename... CIndices,
          typename... AIndices>
void permute(U const C_prefactor, std::tuple<CIndices...> const &Ci, CType *C, U const A_prefactor,
             std::tuple<AIndices...> const &Ai, AType const &A) {
    ::einsums::detail::StagedView<typename CType::ValueType, CType::Rank> c(*C, C_prefactor != U(0));
    ::einsums::detail::StagedView<typename AType::ValueType, AType::Rank> a(A, true);
    permute(C_prefactor, Ci, &c.view, A_prefactor, Ai, a.view);
    c.download();
}

template <detail::StridedTensor CType, detail::StridedTensor AType, typename... CIndices, typename... AIndices>
void permute(std::tuple<CIndices...> const &Ci, CType *C, std::tuple<AIndices...> const &Ai, AType const &A) {
    permute(typename CType::ValueType(0), Ci, C, typename CType::ValueType(1), Ai, A);
}

} // namespace einsums::tensor_algebra
