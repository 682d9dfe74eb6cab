// Compile-time checks of the C++ pattern analysis against the algorithm choices the
// reference tests assert (libs/Einsums/TensorAlgebra/tests/unit/Einsum*.cpp, Dot.cpp).
// Compiles only if every static_assert holds; runs without a GPU.
#include <cstdio>
#include <memory>
#include <vector>

#include "einsums/Einsums.hpp"

using namespace einsums;
using namespace einsums::index;
using einsums::tensor_algebra::detail::classify;
namespace ta = einsums::tensor_algebra::detail;

#define CHECK(Cs, As, Bs, ALG) static_assert(classify(std::tuple Cs, std::tuple As, std::tuple Bs) == ta::ALG)

CHECK((i, j), (i, k), (k, j), GEMM);                  // Einsum1.cpp:31
CHECK((i, l), (i, j, k), (j, k, l), GEMM);            // Einsum1.cpp:72
CHECK((i, j), (i, k), (j, k), GEMM);                  // Einsum3.cpp:42
CHECK((i), (i, j), (j), GEMV);                        // Einsum3.cpp:62
CHECK((i), (i, j, k), (j, i, k), GENERIC);            // Einsum3.cpp:79
CHECK((i, k), (i, j, k), (j, i, k), GENERIC);         // Einsum3.cpp:107
CHECK((i, j, k, l), (i, j, k, p), (p, l), GEMM);      // Einsum4.cpp:46
CHECK((i, j, k, l), (i, j, p, l), (p, k), GENERIC);   // Einsum4.cpp:71
CHECK((p, q, r, l), (p, q, r, s), (s, l), GEMM);      // Einsum5.cpp:28
CHECK((p, q, k, s), (p, q, r, s), (r, k), GENERIC);   // Einsum5.cpp:53
CHECK((i, a, W), (i, W), (a, W), GENERIC);            // Einsum7.cpp:26
CHECK((i, a, P), (i, a, W), (P, W), GEMM);            // Einsum7.cpp:37
CHECK((p, q), (p, q), (p, q), DIRECT);                // Einsum7.cpp:47
CHECK((Q, X), (Q, i, a), (i, a, X), GEMM);            // Einsum7.cpp:66
CHECK((Q, a, X), (Q, Y, X), (a, Y), GENERIC);         // Einsum7.cpp:96
CHECK((a, b), (a), (b), GER);                         // Einsum7.cpp:125
CHECK((a, X), (Q, a, X), (Q, X), GENERIC);            // Einsum7.cpp:169
CHECK((X), (i, a, X), (i, a), GEMV);                  // Einsum7.cpp:192
CHECK((), (i), (i), DOT);                             // Dot.cpp:30
CHECK((), (i, j, a, b), (i, j, a, b), DOT);           // Dot.cpp:84
CHECK((m, n, i, j), (i, j, e, f), (m, n, e, f), GEMM); // CCD Wmnij
CHECK((a, b, e, f), (m, n, e, f), (m, n, a, b), GEMM); // CCD Wabef
CHECK((m, b, e, j), (j, n, f, b), (m, n, e, f), GENERIC); // CCD Wmbej
CHECK((p, q), (p, q, r, s), (r, s), GEMV);            // HF J
CHECK((p, q), (p, r, q, s), (r, s), GENERIC);         // HF K
CHECK((i), (i, i), (i), GENERIC);                     // Hadamard

// Operand kinds of the overload set (TensorAlgebra.hpp:48-146 of the reference): device and
// host tensors are told apart, smart pointers and collections have their own overloads.
static_assert(ta::DeviceTensorConcept<DeviceTensor<double, 2>> && ta::DeviceTensorConcept<DeviceTensorView<float, 3>>);
static_assert(ta::HostTensorConcept<Tensor<double, 2>> && ta::HostTensorConcept<TensorView<double, 4>>);
static_assert(!ta::DeviceTensorConcept<Tensor<double, 2>> && !ta::HostTensorConcept<DeviceTensor<double, 2>>);
static_assert(ta::SmartPointer<std::shared_ptr<Tensor<double, 2>>> && !ta::SmartPointer<Tensor<double, 2>>);
static_assert(ta::Container<std::vector<DeviceTensor<double, 2>>> && !ta::Container<DeviceTensor<double, 2>>);
static_assert(std::is_same_v<ta::value_of_t<std::vector<Tensor<float, 2>>>, float>);
static_assert(std::is_same_v<ta::value_of_t<std::shared_ptr<DeviceTensor<double, 1>>>, double>);
static_assert(std::is_same_v<ta::value_of_t<double>, double>);

int main() {
    std::printf("compile-time dispatch checks passed\n");
    return 0;
}
