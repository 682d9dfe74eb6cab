// C++ API on the GPU: linear_algebra::syev / gesv on DeviceTensors (row-major), checked on
// the host against the defining identities (A V = V diag(W), V^T V = I, A X = B) and against
// LAPACK's singular-pivot info. Exit status 0 = all checks passed. Needs a HIP device.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "einsums/Einsums.hpp"
#include "einsums/LinearAlgebra.hpp"

using namespace einsums;

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

int main() {
    std::mt19937_64 g(42);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int n : {1, 5, 48}) {
        std::vector<double> a(size_t(n) * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) a[size_t(i) * n + j] = a[size_t(j) * n + i] = u(g);
        DeviceTensor<double, 2> A("A", std::array<int64_t, 2>{n, n}, a);
        DeviceTensor<double, 1> W("W", n);
        linear_algebra::syev(&A, &W);
        auto V = A.to_host();
        auto w = W.to_host();
        double res = 0, orth = 0, norm = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double av = 0, vv = 0;
                for (int k = 0; k < n; ++k) {
                    av += a[size_t(i) * n + k] * V[size_t(k) * n + j];  // (A V)(i, j), V(k, j) = A(k, j)
                    vv += V[size_t(k) * n + i] * V[size_t(k) * n + j];  // (V^T V)(i, j)
                }
                res = std::max(res, std::fabs(av - V[size_t(i) * n + j] * w[size_t(j)]));
                orth = std::max(orth, std::fabs(vv - (i == j ? 1.0 : 0.0)));
                norm += a[size_t(i) * n + j] * a[size_t(i) * n + j];
            }
        for (int j = 1; j < n; ++j) EXPECT(w[size_t(j - 1)] <= w[size_t(j)], "n=%d eigenvalues not ascending", n);
        EXPECT(res <= 1e-12 * std::max(1.0, std::sqrt(norm)), "n=%d residual %g", n, res);
        EXPECT(orth <= 1e-12 * std::max(1, n), "n=%d orthogonality %g", n, orth);
    }
    {
        bool threw = false;
        DeviceTensor<double, 2> A("A", 3, 4);
        DeviceTensor<double, 1> W("W", 3);
        try {
            linear_algebra::syev(&A, &W);
        } catch (dimension_error const &) { threw = true; }
        EXPECT(threw, "syev of a 3x4 matrix did not throw dimension_error");
    }
    for (int n : {1, 9, 64}) {
        int const nrhs = 3;
        std::vector<double> a(size_t(n) * n), x(size_t(n) * nrhs), b(size_t(n) * nrhs, 0.0);
        for (auto &v : a) v = u(g);
        for (int i = 0; i < n; ++i) a[size_t(i) * n + i] += 4.0;  // well conditioned
        for (auto &v : x) v = u(g);
        for (int i = 0; i < n; ++i)
            for (int r = 0; r < nrhs; ++r)
                for (int k = 0; k < n; ++k) b[size_t(i) * nrhs + r] += a[size_t(i) * n + k] * x[size_t(k) * nrhs + r];
        DeviceTensor<double, 2> A("A", std::array<int64_t, 2>{n, n}, a);
        DeviceTensor<double, 2> B("B", std::array<int64_t, 2>{n, nrhs}, b);
        int info = linear_algebra::gesv(&A, &B);
        EXPECT(info == 0, "gesv n=%d info %d", n, info);
        auto got = B.to_host();
        double err = 0;
        for (size_t i = 0; i < got.size(); ++i) err = std::max(err, std::fabs(got[i] - x[i]));
        EXPECT(err <= 1e-12, "gesv n=%d error %g", n, err);
        // vector right-hand side
        std::vector<double> b1(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) b1[size_t(i)] = b[size_t(i) * nrhs];
        DeviceTensor<double, 2> A2("A2", std::array<int64_t, 2>{n, n}, a);
        DeviceTensor<double, 1> b1d("b1", std::array<int64_t, 1>{n}, b1);
        info = linear_algebra::gesv(&A2, &b1d);
        auto g1 = b1d.to_host();
        double e1 = 0;
        for (int i = 0; i < n; ++i) e1 = std::max(e1, std::fabs(g1[size_t(i)] - x[size_t(i) * nrhs]));
        EXPECT(info == 0 && e1 <= 1e-12, "gesv vector n=%d info %d error %g", n, info, e1);
    }
    {
        // column 2 of a 4x4 matrix is zero: U(3,3) is exactly zero, info = 3, B untouched
        std::vector<double> a(16);
        for (auto &v : a) v = u(g);
        for (int i = 0; i < 4; ++i) a[size_t(i) * 4 + 2] = 0.0;
        std::vector<double> b = {1.0, 2.0, 3.0, 4.0};
        DeviceTensor<double, 2> A("A", std::array<int64_t, 2>{4, 4}, a);
        DeviceTensor<double, 1> B("B", std::array<int64_t, 1>{4}, b);
        int const info = linear_algebra::gesv(&A, &B);
        EXPECT(info == 3, "singular gesv info %d", info);
        EXPECT(B.to_host() == b, "singular gesv changed B");
    }
    if (failures == 0) std::printf("PASS test_device_linalg\n");
    return failures == 0 ? 0 : 1;
}
