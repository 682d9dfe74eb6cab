// C++ API on the GPU with complex elements: einsum (GEMM, conjugated operands, generic,
// dot), permute and scaling with real and complex prefactors on DeviceTensor<std::complex<T>>,
// and a host Tensor einsum, each checked against a std::complex<long double> loop nest on
// the host. Exit status 0 = all checks passed. Needs a HIP device.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "einsums/Einsums.hpp"

using namespace einsums;
using namespace einsums::index;
using einsums::tensor_algebra::einsum;
using einsums::tensor_algebra::permute;
namespace ta = einsums::tensor_algebra::detail;
using cd = std::complex<double>;
using cf = std::complex<float>;
using cl = std::complex<long double>;

static int failures = 0;
#define EXPECT(cond, ...)                                                                                              \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            ++failures;                                                                                                \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                                           \
            std::printf(__VA_ARGS__);                                                                                  \
            std::printf("\n");                                                                                         \
        }                                                                                                              \
    } while (0)

template <typename T>
static std::vector<T> randc(size_t n, unsigned seed) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<T> v(n);
    for (auto &x : v) x = T(typename T::value_type(u(g)), typename T::value_type(u(g)));
    return v;
}

template <typename T>
static double rel(std::vector<T> const &got, std::vector<cl> const &ref) {
    long double num = 0, den = 0;
    for (size_t i = 0; i < got.size(); ++i) {
        cl const d = cl(got[i].real(), got[i].imag()) - ref[i];
        num += std::norm(d);
        den += std::norm(ref[i]);
    }
    return double(std::sqrt(num / den));
}

static cl lc(cd x) { return cl(x.real(), x.imag()); }

static void gemm_and_conj() {
    int const I = 37, J = 29, K = 41;
    auto a = randc<cd>(size_t(I) * K, 1), b = randc<cd>(size_t(K) * J, 2), c0 = randc<cd>(size_t(I) * J, 3);
    DeviceTensor<cd, 2> A("A", {I, K}, a), B("B", {K, J}, b);
    cd const alpha(1.0, -0.5);
    for (int conj = 0; conj < 4; ++conj) {
        for (cd beta : {cd(0.5, 0.0), cd(0.5, 0.25)}) {
            DeviceTensor<cd, 2> C("C", {I, J}, c0);
            ta::AlgorithmChoice alg = ta::INDETERMINATE;
            if (conj == 0) einsum<false, false>(beta, Indices{i, j}, &C, alpha, Indices{i, k}, A, Indices{k, j}, B, &alg);
            if (conj == 1) einsum<true, false>(beta, Indices{i, j}, &C, alpha, Indices{i, k}, A, Indices{k, j}, B, &alg);
            if (conj == 2) einsum<false, true>(beta, Indices{i, j}, &C, alpha, Indices{i, k}, A, Indices{k, j}, B, &alg);
            if (conj == 3) einsum<true, true>(beta, Indices{i, j}, &C, alpha, Indices{i, k}, A, Indices{k, j}, B, &alg);
            EXPECT(alg == ta::GEMM, "complex GEMM algorithm %s", ta::algorithm_name(alg));
            std::vector<cl> ref(size_t(I) * J);
            for (int x = 0; x < I; ++x)
                for (int y = 0; y < J; ++y) {
                    cl s = 0;
                    for (int z = 0; z < K; ++z) {
                        cl av = lc(a[size_t(x) * K + z]), bv = lc(b[size_t(z) * J + y]);
                        if (conj & 1) av = std::conj(av);
                        if (conj & 2) bv = std::conj(bv);
                        s += av * bv;
                    }
                    ref[size_t(x) * J + y] = lc(beta) * lc(c0[size_t(x) * J + y]) + lc(alpha) * s;
                }
            double const e = rel(C.to_host(), ref);
            EXPECT(e < 1e-14, "complex GEMM conj=%d beta=(%g,%g): rel err %g", conj, beta.real(), beta.imag(), e);
        }
    }
}

static void generic_float_and_dot() {
    // W(m,b,e,j) = sum_{n,f} t(j,n,f,b) g(m,n,e,f), complex<float>
    int const o = 4, v = 6;
    auto t = randc<cf>(size_t(o) * o * v * v, 4), g = randc<cf>(size_t(o) * o * v * v, 5);
    DeviceTensor<cf, 4> T("t", {o, o, v, v}, t), G("g", {o, o, v, v}, g), W("W", o, v, v, o);
    einsum(Indices{m, b, e, j}, &W, Indices{j, n, f, b}, T, Indices{m, n, e, f}, G);
    std::vector<cl> ref(size_t(o) * v * v * o);
    for (int mm = 0; mm < o; ++mm)
        for (int bb = 0; bb < v; ++bb)
            for (int ee = 0; ee < v; ++ee)
                for (int jj = 0; jj < o; ++jj) {
                    cl s = 0;
                    for (int nn = 0; nn < o; ++nn)
                        for (int ff = 0; ff < v; ++ff) {
                            cf tv = t[((size_t(jj) * o + nn) * v + ff) * v + bb], gv = g[((size_t(mm) * o + nn) * v + ee) * v + ff];
                            s += cl(tv.real(), tv.imag()) * cl(gv.real(), gv.imag());
                        }
                    ref[((size_t(mm) * v + bb) * v + ee) * o + jj] = s;
                }
    double const e = rel(W.to_host(), ref);
    EXPECT(e < 1e-5, "complex<float> generic rel err %g", e);
    // dot: E = sum A B (rank-0 C)
    auto x = randc<cd>(600, 6), y = randc<cd>(600, 7);
    DeviceTensor<cd, 2> X("X", {20, 30}, x), Y("Y", {20, 30}, y);
    DeviceTensor<cd, 0> E("E");
    ta::AlgorithmChoice alg = ta::INDETERMINATE;
    einsum(cd(0.0), Indices{}, &E, cd(0.0, 2.0), Indices{i, j}, X, Indices{i, j}, Y, &alg);
    cl s = 0;
    for (size_t q = 0; q < 600; ++q) s += lc(x[q]) * lc(y[q]);
    std::vector<cl> rdot{cl(0.0L, 2.0L) * s};
    EXPECT(alg == ta::DOT, "complex dot algorithm %s", ta::algorithm_name(alg));
    double const ed = rel(E.to_host(), rdot);
    EXPECT(ed < 1e-14, "complex dot rel err %g", ed);
}

static void permute_scale_host() {
    int const I = 6, J = 7, K = 8;
    auto a = randc<cd>(size_t(I) * J * K, 8), c0 = randc<cd>(size_t(K) * I * J, 9);
    DeviceTensor<cd, 3> A("A", {I, J, K}, a), C("C", {K, I, J}, c0);
    cd const beta(0.5, -1.0), alpha(2.0, 0.5);
    permute(beta, Indices{k, i, j}, &C, alpha, Indices{i, j, k}, A);
    std::vector<cl> ref(size_t(K) * I * J);
    for (int x = 0; x < I; ++x)
        for (int y = 0; y < J; ++y)
            for (int z = 0; z < K; ++z) {
                size_t const ci = (size_t(z) * I + x) * J + y;
                ref[ci]         = lc(beta) * lc(c0[ci]) + lc(alpha) * lc(a[(size_t(x) * J + y) * K + z]);
            }
    double const e = rel(C.to_host(), ref);
    EXPECT(e < 1e-15, "complex permute rel err %g", e);
    // C *= complex alpha
    DeviceTensor<cd, 3> S("S", {K, I, J}, c0);
    S *= cd(0.25, 0.75);
    std::vector<cl> rs(c0.size());
    for (size_t q = 0; q < c0.size(); ++q) rs[q] = cl(0.25L, 0.75L) * lc(c0[q]);
    double const es = rel(S.to_host(), rs);
    EXPECT(es < 1e-15, "complex scale rel err %g", es);
    // host tensors, staged through the device
    Tensor<cd, 2> HA(true, "HA", 5, 4), HB(true, "HB", 4, 3), HC(true, "HC", 5, 3);
    auto ha = randc<cd>(20, 10), hb = randc<cd>(12, 11);
    for (int q = 0; q < 20; ++q) HA.data()[q] = ha[size_t(q)];
    for (int q = 0; q < 12; ++q) HB.data()[q] = hb[size_t(q)];
    einsum(Indices{i, j}, &HC, Indices{i, k}, HA, Indices{k, j}, HB);
    std::vector<cd> got(15);
    std::vector<cl> rh(15);
    for (int x = 0; x < 5; ++x)
        for (int y = 0; y < 3; ++y) {
            cl s2 = 0;
            for (int z = 0; z < 4; ++z) s2 += lc(ha[size_t(x) * 4 + z]) * lc(hb[size_t(z) * 3 + y]);
            rh[size_t(x) * 3 + y]  = s2;
            got[size_t(x) * 3 + y] = HC(x, y);
        }
    double const eh = rel(got, rh);
    EXPECT(eh < 1e-14, "complex host einsum rel err %g", eh);
}

int main() {
    try {
        gemm_and_conj();
        generic_float_and_dot();
        permute_scale_host();
    } catch (std::exception const &e) {
        std::printf("FAIL exception: %s\n", e.what());
        ++failures;
    }
    if (failures == 0) std::printf("PASS test_complex\n");
    return failures ? 1 : 0;
}
