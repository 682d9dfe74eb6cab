// Index labels and the Indices pack (reference: libs/Einsums/TensorAlgebra/include/
// Einsums/TensorAlgebra/Detail/Index.hpp:32-206).
//
// Every label is an empty tag type carrying its spelling and a compile-time id (FNV-1a
// of the spelling) that the C ABI uses as the label value.
#pragma once

#include <cstdint>
#include <tuple>
#include <type_traits>

namespace einsums {
namespace index {

struct LabelBase {};

constexpr int32_t label_id(char const *s) {
    uint32_t h = 2166136261u;
    while (*s) {
        h ^= static_cast<unsigned char>(*s++);
        h *= 16777619u;
    }
    return static_cast<int32_t>(h & 0x7fffffffu);
}

} // namespace index
} // namespace einsums

#define EINSUMS_MAKE_INDEX(x)                                                                                          \
    namespace einsums::index {                                                                                         \
    struct x : public LabelBase {                                                                                      \
        static constexpr char const *letter = #x;                                                                      \
        static constexpr int32_t id         = label_id(#x);                                                            \
        constexpr x()                       = default;                                                                 \
    };                                                                                                                 \
    inline constexpr struct x x;                                                                                       \
    }

// The reference's spelling of the macro (Detail/Index.hpp:32).
#ifndef MAKE_INDEX
#define MAKE_INDEX(x) EINSUMS_MAKE_INDEX(x)
#endif

EINSUMS_MAKE_INDEX(A)
EINSUMS_MAKE_INDEX(a)
EINSUMS_MAKE_INDEX(B)
EINSUMS_MAKE_INDEX(b)
EINSUMS_MAKE_INDEX(C)
EINSUMS_MAKE_INDEX(c)
EINSUMS_MAKE_INDEX(D)
EINSUMS_MAKE_INDEX(d)
EINSUMS_MAKE_INDEX(E)
EINSUMS_MAKE_INDEX(e)
EINSUMS_MAKE_INDEX(F)
EINSUMS_MAKE_INDEX(f)
EINSUMS_MAKE_INDEX(G)
EINSUMS_MAKE_INDEX(g)
EINSUMS_MAKE_INDEX(H)
EINSUMS_MAKE_INDEX(h)
EINSUMS_MAKE_INDEX(I)
EINSUMS_MAKE_INDEX(i)
EINSUMS_MAKE_INDEX(J)
EINSUMS_MAKE_INDEX(j)
EINSUMS_MAKE_INDEX(K)
EINSUMS_MAKE_INDEX(k)
EINSUMS_MAKE_INDEX(L)
EINSUMS_MAKE_INDEX(l)
EINSUMS_MAKE_INDEX(M)
EINSUMS_MAKE_INDEX(m)
EINSUMS_MAKE_INDEX(N)
EINSUMS_MAKE_INDEX(n)
EINSUMS_MAKE_INDEX(O)
EINSUMS_MAKE_INDEX(o)
EINSUMS_MAKE_INDEX(P)
EINSUMS_MAKE_INDEX(p)
EINSUMS_MAKE_INDEX(Q)
EINSUMS_MAKE_INDEX(q)
EINSUMS_MAKE_INDEX(R)
EINSUMS_MAKE_INDEX(r)
EINSUMS_MAKE_INDEX(S)
EINSUMS_MAKE_INDEX(s)
EINSUMS_MAKE_INDEX(T)
EINSUMS_MAKE_INDEX(t)
EINSUMS_MAKE_INDEX(U)
EINSUMS_MAKE_INDEX(u)
EINSUMS_MAKE_INDEX(V)
EINSUMS_MAKE_INDEX(v)
EINSUMS_MAKE_INDEX(W)
EINSUMS_MAKE_INDEX(w)
EINSUMS_MAKE_INDEX(X)
EINSUMS_MAKE_INDEX(x)
EINSUMS_MAKE_INDEX(Y)
EINSUMS_MAKE_INDEX(y)
EINSUMS_MAKE_INDEX(Z)
EINSUMS_MAKE_INDEX(z)
EINSUMS_MAKE_INDEX(alpha)
EINSUMS_MAKE_INDEX(beta)
EINSUMS_MAKE_INDEX(gamma)
EINSUMS_MAKE_INDEX(delta)
EINSUMS_MAKE_INDEX(epsilon)
EINSUMS_MAKE_INDEX(zeta)
EINSUMS_MAKE_INDEX(eta)
EINSUMS_MAKE_INDEX(theta)
EINSUMS_MAKE_INDEX(iota)
EINSUMS_MAKE_INDEX(kappa)
EINSUMS_MAKE_INDEX(lambda)
EINSUMS_MAKE_INDEX(mu)
EINSUMS_MAKE_INDEX(nu)
EINSUMS_MAKE_INDEX(xi)
EINSUMS_MAKE_INDEX(omicron)
EINSUMS_MAKE_INDEX(pi)
EINSUMS_MAKE_INDEX(rho)
EINSUMS_MAKE_INDEX(sigma)
EINSUMS_MAKE_INDEX(tau)
EINSUMS_MAKE_INDEX(upsilon)
EINSUMS_MAKE_INDEX(phi)
EINSUMS_MAKE_INDEX(chi)
EINSUMS_MAKE_INDEX(psi)
EINSUMS_MAKE_INDEX(omega)

namespace einsums {

// Indices{i, j, ...}: a tuple of label tags (Index.hpp:199-206 of the reference).
template <typename... Args>
    requires(std::is_base_of_v<index::LabelBase, Args> && ... && true)
struct Indices : std::tuple<Args...> {
    constexpr Indices(Args... args) : std::tuple<Args...>(args...) {}
};

template <typename... Args>
Indices(Args...) -> Indices<Args...>;

} // namespace einsums
