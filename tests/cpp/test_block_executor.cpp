// The C++ block/tiled einsum beyond block x block -> block, on the GPU, against host loop
// nests: block x block -> dense C (diagonal blocks only) and -> scalar C; block x dense
// vector -> dense vector; dense vector x block -> block; tiled x tiled -> tiled with an absent
// C tile created, and -> scalar; block x tiled -> tiled (BlockAlgebra.hpp:23-428,
// TileAlgebra.hpp:18-158, BlockTileAlgebra.hpp:18-173). Exit status 0 = all checks passed.
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

static std::mt19937_64 rng(77);
static std::vector<double> randv(size_t n) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> v(n);
    for (auto &x : v) x = u(rng);
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
    std::vector<int64_t> dims{5, 0, 17, 33, 1, 64};
    int64_t N = 0;
    std::vector<int64_t> off;
    for (auto n : dims) {
        off.push_back(N);
        N += n;
    }
    BlockDeviceTensor<double, 2> A("A", dims), B("B", dims);
    std::vector<std::vector<double>> hA, hB;
    for (size_t i = 0; i < dims.size(); ++i) {
        size_t n = size_t(dims[i]);
        hA.push_back(randv(n * n));
        hB.push_back(randv(n * n));
        if (n) {
            A[i].from_host(hA.back());
            B[i].from_host(hB.back());
        }
    }
    // block x block -> dense C: C = 0.5 C + 2 A B on the diagonal blocks, the rest untouched
    std::vector<double> hC = randv(size_t(N * N));
    DeviceTensor<double, 2> C("C", {N, N}, hC);
    einsum(0.5, Indices{i, j}, &C, 2.0, Indices{i, k}, A, Indices{k, j}, B);
    std::vector<double> ref = hC;
    for (size_t b = 0; b < dims.size(); ++b) {
        int64_t n = dims[b], o = off[b];
        for (int64_t r = 0; r < n; ++r)
            for (int64_t c = 0; c < n; ++c) {
                double s = 0;
                for (int64_t q = 0; q < n; ++q) s += hA[b][size_t(r * n + q)] * hB[b][size_t(q * n + c)];
                ref[size_t((o + r) * N + o + c)] = 0.5 * hC[size_t((o + r) * N + o + c)] + 2.0 * s;
            }
    }
    EXPECT(rel(C.to_host(), ref) < 1e-13, "block x block -> dense rel err %.3e", rel(C.to_host(), ref));
    // block x block -> scalar: E = 0.25 sum_b A_b . B_b (C_prefactor 0 clears a NaN)
    double E = NAN, ER = 0;
    einsum(0.0, Indices{}, &E, 0.25, Indices{i, j}, A, Indices{i, j}, B);
    for (size_t b = 0; b < dims.size(); ++b)
        for (size_t e = 0; e < hA[b].size(); ++e) ER += 0.25 * hA[b][e] * hB[b][e];
    EXPECT(std::fabs(E - ER) <= 1e-13 * std::fabs(ER), "block scalar %.17g vs %.17g", E, ER);
    // block x dense vector -> dense vector: y = A x (block-diagonal)
    std::vector<double> hx = randv(size_t(N)), hy(size_t(N), 0.0);
    DeviceTensor<double, 1> x("x", {N}, hx), y("y", {N}, hy);
    einsum(0.0, Indices{i}, &y, 1.0, Indices{i, j}, A, Indices{j}, x);
    for (size_t b = 0; b < dims.size(); ++b)
        for (int64_t r = 0; r < dims[b]; ++r) {
            double s = 0;
            for (int64_t c = 0; c < dims[b]; ++c) s += hA[b][size_t(r * dims[b] + c)] * hx[size_t(off[b] + c)];
            hy[size_t(off[b] + r)] = s;
        }
    EXPECT(rel(y.to_host(), hy) < 1e-13, "block x vector rel err %.3e", rel(y.to_host(), hy));
    // dense vector x block -> block: D_ij = x_i B_ij
    BlockDeviceTensor<double, 2> D("D", dims);
    einsum(0.0, Indices{i, j}, &D, 1.0, Indices{i}, x, Indices{i, j}, B);
    double worst = 0;
    for (size_t b = 0; b < dims.size(); ++b) {
        if (!dims[b]) continue;
        std::vector<double> r(hB[b].size());
        for (int64_t p = 0; p < dims[b]; ++p)
            for (int64_t q = 0; q < dims[b]; ++q) r[size_t(p * dims[b] + q)] = hx[size_t(off[b] + p)] * hB[b][size_t(p * dims[b] + q)];
        worst = std::max(worst, rel(D[b].to_host(), r));
    }
    EXPECT(worst < 1e-13, "vector x block rel err %.3e", worst);
    // tiled x tiled -> tiled with an absent C tile created; then -> scalar
    std::array<std::vector<int64_t>, 2> td{std::vector<int64_t>{7, 0, 19}, std::vector<int64_t>{5, 16}};
    TiledDeviceTensor<double, 2> TA("TA", td), TB("TB", {td[1], td[0]}), TC("TC", {td[0], td[0]});
    auto fill = [&](TiledDeviceTensor<double, 2> &T, std::array<int, 2> c) {
        auto s = T.tile_shape(c);
        T.tile(c).from_host(randv(size_t(s[0] * s[1])));
    };
    fill(TA, {0, 0});
    fill(TA, {2, 1});
    fill(TB, {0, 2});
    fill(TB, {1, 0});
    fill(TC, {0, 0});
    auto dense = [&](TiledDeviceTensor<double, 2> &T, std::vector<int64_t> const &r, std::vector<int64_t> const &c) {
        int64_t R = 0, Cc = 0;
        for (auto v : r) R += v;
        for (auto v : c) Cc += v;
        std::vector<double> out(size_t(R * Cc), 0.0);
        for (auto &kv : T.tiles()) {
            int64_t r0 = 0, c0 = 0;
            for (int q = 0; q < kv.first[0]; ++q) r0 += r[size_t(q)];
            for (int q = 0; q < kv.first[1]; ++q) c0 += c[size_t(q)];
            auto h = kv.second.to_host();
            for (int64_t p = 0; p < kv.second.dim(0); ++p)
                for (int64_t q = 0; q < kv.second.dim(1); ++q) out[size_t((r0 + p) * Cc + c0 + q)] = h[size_t(p * kv.second.dim(1) + q)];
        }
        return out;
    };
    auto dA = dense(TA, td[0], td[1]), dB = dense(TB, td[1], td[0]), dC0 = dense(TC, td[0], td[0]);
    einsum(0.5, Indices{i, j}, &TC, 1.0, Indices{i, k}, TA, Indices{k, j}, TB);
    int64_t n0 = 26, n1 = 21;
    std::vector<double> tref(size_t(n0 * n0));
    for (int64_t p = 0; p < n0; ++p)
        for (int64_t q = 0; q < n0; ++q) {
            double s = 0;
            for (int64_t r = 0; r < n1; ++r) s += dA[size_t(p * n1 + r)] * dB[size_t(r * n0 + q)];
            tref[size_t(p * n0 + q)] = 0.5 * dC0[size_t(p * n0 + q)] + s;
        }
    EXPECT(rel(dense(TC, td[0], td[0]), tref) < 1e-13, "tiled rel err %.3e", rel(dense(TC, td[0], td[0]), tref));
    EXPECT(TC.has_tile({2, 0}) && TC.has_tile({0, 2}) && !TC.has_tile({2, 2}),
           "absent C tiles receiving a contribution must be created, the others not");
    double S = 0, SR = 0;
    einsum(0.0, Indices{}, &S, 1.0, Indices{i, k}, TA, Indices{k, i}, TB);
    for (int64_t p = 0; p < n0; ++p)
        for (int64_t r = 0; r < n1; ++r) SR += dA[size_t(p * n1 + r)] * dB[size_t(r * n0 + p)];
    EXPECT(std::fabs(S - SR) <= 1e-13 * std::fabs(SR), "tiled scalar %.17g vs %.17g", S, SR);
    // block x tiled -> tiled: A's diagonal blocks as tiles
    std::array<std::vector<int64_t>, 2> bd{dims, dims};
    TiledDeviceTensor<double, 2> TT("TT", bd), TO("TO", bd);
    fill(TT, {0, 2});
    fill(TT, {3, 5});
    fill(TT, {5, 0});
    auto dT = dense(TT, dims, dims);
    einsum(0.0, Indices{i, j}, &TO, 1.0, Indices{i, k}, A, Indices{k, j}, TT);
    std::vector<double> dAd(size_t(N * N), 0.0), mref(size_t(N * N), 0.0);
    for (size_t b = 0; b < dims.size(); ++b)
        for (int64_t p = 0; p < dims[b]; ++p)
            for (int64_t q = 0; q < dims[b]; ++q) dAd[size_t((off[b] + p) * N + off[b] + q)] = hA[b][size_t(p * dims[b] + q)];
    for (int64_t p = 0; p < N; ++p)
        for (int64_t q = 0; q < N; ++q) {
            double s = 0;
            for (int64_t r = 0; r < N; ++r) s += dAd[size_t(p * N + r)] * dT[size_t(r * N + q)];
            mref[size_t(p * N + q)] = s;
        }
    EXPECT(rel(dense(TO, dims, dims), mref) < 1e-13, "block x tiled rel err %.3e", rel(dense(TO, dims, dims), mref));
    if (failures == 0) std::printf("PASS test_block_executor\n");
    return failures == 0 ? 0 : 1;
}
