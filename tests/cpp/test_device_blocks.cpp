// C++ block API on the GPU: BlockDeviceTensor and TiledDeviceTensor einsums as grouped
// plans, checked against host loop nests. Exit status 0 = all checks passed.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "einsums/BlockTensor.hpp"
#include "einsums/Einsums.hpp"

using namespace einsums;
using namespace einsums::index;
using einsums::tensor_algebra::einsum;

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

int main() {
    check(ein_init(), "init");
    // BlockDeviceTensor: C_ij = 0.5 C_ij - A_ik B_kj per block (8 uneven blocks)
    std::vector<int64_t> dims{8, 77, 320, 33, 150, 9, 256, 41};
    BlockDeviceTensor<double, 2> A("A", dims), B("B", dims), C("C", dims);
    std::vector<std::vector<double>> hA, hB, hC;
    for (size_t i = 0; i < dims.size(); ++i) {
        size_t n = size_t(dims[i]);
        hA.push_back(randv(n * n, unsigned(10 + i)));
        hB.push_back(randv(n * n, unsigned(20 + i)));
        hC.push_back(randv(n * n, unsigned(30 + i)));
        A[i].from_host(hA.back());
        B[i].from_host(hB.back());
        C[i].from_host(hC.back());
    }
    auto plan = tensor_algebra::make_block_plan(0.5, Indices{i, j}, &C, -1.0, Indices{i, k}, A, Indices{k, j}, B);
    EXPECT(plan.waves() == 1, "block plan waves %d", plan.waves());
    plan.execute();
    for (size_t b = 0; b < dims.size(); ++b) {
        size_t n = size_t(dims[b]);
        std::vector<double> ref = hC[b];
        for (size_t r = 0; r < n; ++r)
            for (size_t c = 0; c < n; ++c) {
                double s = 0;
                for (size_t q = 0; q < n; ++q) s += hA[b][r * n + q] * hB[b][q * n + c];
                ref[r * n + c] = 0.5 * ref[r * n + c] - s;
            }
        EXPECT(rel(C[b].to_host(), ref) < 1e-13, "block %zu rel err %.3e", b, rel(C[b].to_host(), ref));
    }

    // TiledDeviceTensor: C_ij (tiles 2x2) = A_ik B_kj with k cut into 3 tiles
    std::array<std::vector<int64_t>, 2> ti{std::vector<int64_t>{5, 7}, std::vector<int64_t>{4, 6, 3}};
    std::array<std::vector<int64_t>, 2> tj{std::vector<int64_t>{4, 6, 3}, std::vector<int64_t>{8, 2}};
    std::array<std::vector<int64_t>, 2> tc{std::vector<int64_t>{5, 7}, std::vector<int64_t>{8, 2}};
    TiledDeviceTensor<double, 2> TA("TA", ti), TB("TB", tj), TC("TC", tc);
    int64_t const I = 12, K = 13, J = 10;
    std::vector<double> dA(size_t(I * K), 0.0), dB(size_t(K * J), 0.0);
    int64_t offI[3] = {0, 5, 12}, offK[4] = {0, 4, 10, 13}, offJ[3] = {0, 8, 10};
    unsigned seed = 100;
    for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 3; ++c) {
            if ((a + c) % 2 == 1) continue; // some tiles absent
            auto v = randv(size_t((offI[a + 1] - offI[a]) * (offK[c + 1] - offK[c])), seed++);
            TA.tile({a, c}).from_host(v);
            for (int64_t r = 0; r < offI[a + 1] - offI[a]; ++r)
                for (int64_t q = 0; q < offK[c + 1] - offK[c]; ++q)
                    dA[size_t((offI[a] + r) * K + offK[c] + q)] = v[size_t(r * (offK[c + 1] - offK[c]) + q)];
        }
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 2; ++b) {
            auto v = randv(size_t((offK[c + 1] - offK[c]) * (offJ[b + 1] - offJ[b])), seed++);
            TB.tile({c, b}).from_host(v);
            for (int64_t r = 0; r < offK[c + 1] - offK[c]; ++r)
                for (int64_t q = 0; q < offJ[b + 1] - offJ[b]; ++q)
                    dB[size_t((offK[c] + r) * J + offJ[b] + q)] = v[size_t(r * (offJ[b + 1] - offJ[b]) + q)];
        }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) TC.tile({a, b}).zero();
    einsum(0.0, Indices{i, j}, &TC, 1.0, Indices{i, k}, TA, Indices{k, j}, TB);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            auto got = TC.tile({a, b}).to_host();
            int64_t ni = offI[a + 1] - offI[a], nj = offJ[b + 1] - offJ[b];
            std::vector<double> ref(size_t(ni * nj), 0.0);
            for (int64_t r = 0; r < ni; ++r)
                for (int64_t c = 0; c < nj; ++c)
                    for (int64_t q = 0; q < K; ++q)
                        ref[size_t(r * nj + c)] += dA[size_t((offI[a] + r) * K + q)] * dB[size_t(q * J + offJ[b] + c)];
            EXPECT(rel(got, ref) < 1e-13, "tile (%d,%d) rel err %.3e", a, b, rel(got, ref));
        }
    if (failures == 0) std::printf("PASS test_device_blocks\n");
    return failures == 0 ? 0 : 1;
}
