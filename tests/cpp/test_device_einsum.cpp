// C++ API on the GPU: DeviceTensor einsum/permute through the compile-time dispatch,
// checked against the integer known answers of Einsum1.cpp and against host loop nests.
// Exit status 0 = all checks passed. Needs a HIP device.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "einsums/Einsums.hpp"

using namespace einsums;
using namespace einsums::index;
using einsums::tensor_algebra::einsum;
using einsums::tensor_algebra::permute;
namespace ta = einsums::tensor_algebra::detail;

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

static std::vector<double> randv(size_t n, unsigned seed) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> v(n);
    for (auto &x : v) x = u(g);
    return v;
}

static double rel(std::vector<double> const &a, std::vector<double> const &b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

static void einsum1() {
    // Einsum1.cpp:17-49 (ik=ij,jk) and :51-88 (il=ijk,jkl), integer data 1..n
    std::vector<double> a9(9), a27(27);
    for (int i = 0; i < 9; ++i) a9[i] = i + 1;
    for (int i = 0; i < 27; ++i) a27[i] = i + 1;
    DeviceTensor<double, 2> A("A", {3, 3}, a9), B("B", {3, 3}, a9), C("C", 3, 3);
    ta::AlgorithmChoice alg = ta::INDETERMINATE;
    einsum(Indices{i, j}, &C, Indices{i, k}, A, Indices{k, j}, B, &alg);
    EXPECT(alg == ta::GEMM, "ik=ij,jk algorithm %s", ta::algorithm_name(alg));
    std::vector<double> e1{30, 36, 42, 66, 81, 96, 102, 126, 150};
    EXPECT(C.to_host() == e1, "ik=ij,jk values");
    DeviceTensor<double, 3> A3("A", {3, 3, 3}, a27), B3("B", {3, 3, 3}, a27);
    einsum(Indices{i, l}, &C, Indices{i, j, k}, A3, Indices{j, k, l}, B3, &alg);
    EXPECT(alg == ta::GEMM, "il=ijk,jkl algorithm %s", ta::algorithm_name(alg));
    std::vector<double> e2{765, 810, 855, 1818, 1944, 2070, 2871, 3078, 3285};
    EXPECT(C.to_host() == e2, "il=ijk,jkl values");
}

static void ccd_like() {
    // Wmbej += -0.5 t_jnfb g_mnef (generic) and Wmnij += 0.25 t_ijef g_mnef (GEMM)
    int const o = 7, v = 13;
    auto th = randv(size_t(o * o * v * v), 1), gh = randv(size_t(o * o * v * v), 2), wh = randv(size_t(o * v * v * o), 3);
    DeviceTensor<double, 4> t("t", {o, o, v, v}, th), g("g", {o, o, v, v}, gh), W("W", {o, v, v, o}, wh);
    ta::AlgorithmChoice alg;
    einsum(1.0, Indices{m, b, e, j}, &W, -0.5, Indices{j, n, f, b}, t, Indices{m, n, e, f}, g, &alg);
    EXPECT(alg == ta::GENERIC, "Wmbej algorithm %s", ta::algorithm_name(alg));
    std::vector<double> ref = wh;
    auto T = [&](int J, int N, int F, int Bb) { return th[((size_t(J) * o + N) * v + F) * v + Bb]; };
    auto G = [&](int M, int N, int E, int F) { return gh[((size_t(M) * o + N) * v + E) * v + F]; };
    for (int M = 0; M < o; ++M)
        for (int Bb = 0; Bb < v; ++Bb)
            for (int E = 0; E < v; ++E)
                for (int J = 0; J < o; ++J) {
                    double s = 0;
                    for (int N = 0; N < o; ++N)
                        for (int F = 0; F < v; ++F) s += T(J, N, F, Bb) * G(M, N, E, F);
                    ref[((size_t(M) * v + Bb) * v + E) * o + J] += -0.5 * s;
                }
    EXPECT(rel(W.to_host(), ref) < 1e-13, "Wmbej rel err %.3e", rel(W.to_host(), ref));

    auto w2 = randv(size_t(o * o * o * o), 4);
    DeviceTensor<double, 4> W2("W2", {o, o, o, o}, w2);
    einsum(1.0, Indices{m, n, i, j}, &W2, 0.25, Indices{i, j, e, f}, t, Indices{m, n, e, f}, g, &alg);
    EXPECT(alg == ta::GEMM, "Wmnij algorithm %s", ta::algorithm_name(alg));
    std::vector<double> r2 = w2;
    for (int M = 0; M < o; ++M)
        for (int N = 0; N < o; ++N)
            for (int I = 0; I < o; ++I)
                for (int J = 0; J < o; ++J) {
                    double s = 0;
                    for (int E = 0; E < v; ++E)
                        for (int F = 0; F < v; ++F) s += T(I, J, E, F) * G(M, N, E, F);
                    r2[((size_t(M) * o + N) * o + I) * o + J] += 0.25 * s;
                }
    EXPECT(rel(W2.to_host(), r2) < 1e-13, "Wmnij rel err %.3e", rel(W2.to_host(), r2));

    DeviceTensor<double, 0> E0("E");
    E0 = 0.0;
    einsum(0.0, Indices{}, &E0, 0.25, Indices{i, j, a, b}, t, Indices{i, j, a, b}, g, &alg);
    EXPECT(alg == ta::DOT, "energy algorithm %s", ta::algorithm_name(alg));
    double es = 0;
    for (size_t q = 0; q < th.size(); ++q) es += th[q] * gh[q];
    EXPECT(std::fabs(double(E0) - 0.25 * es) < 1e-13 * std::fabs(es), "energy %.17g vs %.17g", double(E0), 0.25 * es);

    DeviceTensor<double, 4> P("P", o, v, o, v);
    permute(Indices{i, a, j, b}, &P, Indices{i, j, a, b}, t);
    auto ph = P.to_host();
    bool same = true;
    for (int I = 0; I < o; ++I)
        for (int A_ = 0; A_ < v; ++A_)
            for (int J = 0; J < o; ++J)
                for (int Bb = 0; Bb < v; ++Bb)
                    same = same && ph[((size_t(I) * v + A_) * o + J) * v + Bb] == T(I, J, A_, Bb);
    EXPECT(same, "permute ijab->iajb not bit-exact");
}

static void views_and_gemv() {
    // GEMV (HF J shape) and a sub-view that cannot fold (Einsum7.cpp:197-209 -> GENERIC)
    int const nb = 9;
    auto gh = randv(size_t(nb * nb * nb * nb), 5), dh = randv(size_t(nb * nb), 6);
    DeviceTensor<double, 4> g("g", {nb, nb, nb, nb}, gh);
    DeviceTensor<double, 2> D("D", {nb, nb}, dh), J("J", nb, nb);
    ta::AlgorithmChoice alg;
    einsum(Indices{p, q}, &J, Indices{p, q, r, s}, g, Indices{r, s}, D, &alg);
    EXPECT(alg == ta::GEMV, "J algorithm %s", ta::algorithm_name(alg));
    std::vector<double> ref(size_t(nb * nb), 0.0);
    for (int P_ = 0; P_ < nb; ++P_)
        for (int Q_ = 0; Q_ < nb; ++Q_)
            for (int R_ = 0; R_ < nb; ++R_)
                for (int S_ = 0; S_ < nb; ++S_)
                    ref[size_t(P_ * nb + Q_)] += gh[((size_t(P_) * nb + Q_) * nb + R_) * nb + S_] * dh[size_t(R_ * nb + S_)];
    EXPECT(rel(J.to_host(), ref) < 1e-13, "J rel err");

    int const naux = 6, nocc = 3, nvirt = 4, ur = 5;
    auto bh = randv(size_t(naux * (nocc + nvirt) * (nocc + nvirt)), 7), xh = randv(size_t(nocc * nvirt * ur), 8);
    DeviceTensor<double, 3> Bq("B", {naux, nocc + nvirt, nocc + nvirt}, bh), Xt("X", {nocc, nvirt, ur}, xh);
    auto Qov = Bq(All, Range{0, nocc}, Range{nocc, nocc + nvirt});
    DeviceTensor<double, 2> Nt("N", naux, ur);
    einsum(Indices{Q, X}, &Nt, Indices{Q, i, a}, Qov, Indices{i, a, X}, Xt, &alg);
    EXPECT(alg == ta::GENERIC, "sub-view algorithm %s", ta::algorithm_name(alg));
    std::vector<double> rn(size_t(naux * ur), 0.0);
    for (int Q_ = 0; Q_ < naux; ++Q_)
        for (int I = 0; I < nocc; ++I)
            for (int A_ = 0; A_ < nvirt; ++A_)
                for (int X_ = 0; X_ < ur; ++X_)
                    rn[size_t(Q_ * ur + X_)] += bh[(size_t(Q_) * (nocc + nvirt) + I) * (nocc + nvirt) + nocc + A_] *
                                                xh[(size_t(I) * nvirt + A_) * ur + X_];
    EXPECT(rel(Nt.to_host(), rn) < 1e-13, "sub-view rel err");

    bool threw = false;
    try {
        DeviceTensor<double, 2> bad("bad", nb + 1, nb);
        einsum(Indices{p, q}, &bad, Indices{p, q, r, s}, g, Indices{r, s}, D);
    } catch (dimension_error const &) {
        threw = true;
    }
    EXPECT(threw, "dimension mismatch did not throw dimension_error");
}

int main() {
    check(ein_init(), "init");
    einsum1();
    ccd_like();
    views_and_gemv();
    if (failures == 0) std::printf("PASS test_device_einsum\n");
    return failures == 0 ? 0 : 1;
}
