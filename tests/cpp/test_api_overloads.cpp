// The reference's einsum overload set spelled as its users spell it, run on the GPU and
// checked against the CPU oracle (oracle/liboracle.so, linked into this test program):
//   MAKE_INDEX labels; host Tensor / TensorView operands (column- and row-major, views);
//   a scalar double* C (DOT and a full contraction); std::shared_ptr operands; the batched
//   form over std::vector collections (and its size check); host permute; default
//   prefactors. Exit status 0 = all checks passed. Needs a HIP device.
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <vector>

#include "einsums/Einsums.hpp"

MAKE_INDEX(kk)
MAKE_INDEX(ll)

extern "C" {
int oracle_einsum_f64(double cp, ein_tensor_desc const *C, double *c, double abp, ein_tensor_desc const *A,
                      double const *a, ein_tensor_desc const *B, double const *b, int *alg);
char const *oracle_blas_name();
int oracle_permute_f64(int rank, int64_t const *dims_A, int const *perm, double alpha, double const *A,
                       int64_t const *strides_A, double beta, double *C, int64_t const *strides_C);
}

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

static std::mt19937_64 rng(20261203);

template <typename Tn>
static void fill(Tn &t) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    t.for_each_offset([&](int64_t off) { t.data()[off] = u(rng); });
}

// Host descriptor with single-letter labels for the oracle.
template <typename Tn>
static ein_tensor_desc hdesc(Tn const &t, char const *labels) {
    ein_tensor_desc d{};
    d.rank = int32_t(Tn::Rank);
    for (size_t i = 0; i < Tn::Rank; ++i) {
        d.labels[i]  = labels[i];
        d.dims[i]    = t.dim(i);
        d.strides[i] = t.stride(i);
    }
    return d;
}

// Relative Frobenius difference over the elements of two equally shaped host views.
template <typename Tn>
static double rel(Tn const &x, Tn const &ref) {
    std::vector<double> a, b;
    x.for_each_offset([&](int64_t o) { a.push_back(x.data()[o]); });
    ref.for_each_offset([&](int64_t o) { b.push_back(ref.data()[o]); });
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

static std::vector<double> host_of(DeviceTensor<double, 2> const &t) { return t.to_host(); }

static void host_tensors() {
    // column-major (the reference's default layout): C_ij = 0.5 C_ij + 2 A_ikk B_kkj
    Tensor<double, 2> A("A", 40, 30), B("B", 30, 20), C("C", 40, 20);
    fill(A);
    fill(B);
    fill(C);
    EXPECT(!A.is_row_major() && A.stride(0) == 1, "default layout is column-major");
    Tensor<double, 2> R = C;
    ta::AlgorithmChoice alg = ta::INDETERMINATE;
    einsum(0.5, Indices{i, j}, &C, 2.0, Indices{i, kk}, A, Indices{kk, j}, B, &alg);
    auto cd = hdesc(R, "ij"), ad = hdesc(A, "ik"), bd = hdesc(B, "kj");
    int ralg = -1;
    oracle_einsum_f64(0.5, &cd, R.data(), 2.0, &ad, A.data(), &bd, B.data(), &ralg);
    EXPECT(alg == ta::AlgorithmChoice(ralg) && alg == ta::GEMM, "host GEMM algorithm %s", ta::algorithm_name(alg));
    EXPECT(rel(C, R) < 1e-13, "host column-major GEMM rel err %.3e", rel(C, R));

    // row-major rank 4, C a view of a larger tensor (generic, non-GEMM layout):
    // W_mbej += -0.5 t_jnfb g_mnef with m = 6, b = 4, e = 7, j = 6, n = 5, f = 7
    Tensor<double, 4> big(true, "W", 8, 9, 7, 6);
    fill(big);
    Tensor<double, 4> bigR = big;
    auto W  = big(Range{1, 7}, Range{2, 6}, All, Range{0, 6});
    auto WR = bigR(Range{1, 7}, Range{2, 6}, All, Range{0, 6});
    Tensor<double, 4> t2(true, "t2", 6, 5, 7, 4), g2(true, "g2", 6, 5, 7, 7);
    fill(t2);
    fill(g2);
    einsum(1.0, Indices{m, b, e, j}, &W, -0.5, Indices{j, n, f, b}, t2, Indices{m, n, e, f}, g2, &alg);
    auto wd = hdesc(WR, "mbej"), td = hdesc(t2, "jnfb"), gd = hdesc(g2, "mnef");
    oracle_einsum_f64(1.0, &wd, WR.data(), -0.5, &td, t2.data(), &gd, g2.data(), &ralg);
    EXPECT(alg == ta::AlgorithmChoice(ralg) && alg == ta::GENERIC, "host view generic algorithm %s", ta::algorithm_name(alg));
    EXPECT(rel(big, bigR) < 1e-13, "host view generic rel err %.3e (whole tensor: outside the view unchanged)",
           rel(big, bigR));

    // default prefactors on host tensors: C = A B
    Tensor<double, 2> C2("C2", 40, 20), R2("R2", 40, 20);
    C2.set_all(NAN); // C_prefactor 0 clears NaN (BaseAlgebra.hpp:118-133)
    einsum(Indices{i, j}, &C2, Indices{i, kk}, A, Indices{kk, j}, B);
    cd = hdesc(R2, "ij");
    oracle_einsum_f64(0.0, &cd, R2.data(), 1.0, &ad, A.data(), &bd, B.data(), &ralg);
    EXPECT(rel(C2, R2) < 1e-13, "default-prefactor host GEMM rel err %.3e", rel(C2, R2));
}

static void scalar_c() {
    // E = 0.5 E + 0.25 sum_ij A_ij B_ij (DOT) with host operands and a plain double target
    Tensor<double, 2> A(true, "A", 33, 17), B(true, "B", 33, 17);
    fill(A);
    fill(B);
    double E = 0.75, ER = 0.75;
    ta::AlgorithmChoice alg = ta::INDETERMINATE;
    einsum(0.5, Indices{}, &E, 0.25, Indices{i, j}, A, Indices{i, j}, B, &alg);
    ein_tensor_desc e0{};
    auto ad = hdesc(A, "ij"), bd = hdesc(B, "ij");
    int ralg = -1;
    oracle_einsum_f64(0.5, &e0, &ER, 0.25, &ad, A.data(), &bd, B.data(), &ralg);
    EXPECT(alg == ta::DOT && ralg == int(ta::DOT), "scalar DOT algorithm %s", ta::algorithm_name(alg));
    EXPECT(std::fabs(E - ER) <= 1e-13 * std::fabs(ER), "scalar DOT %.17g vs %.17g", E, ER);

    // device operands, full contraction with a transposed B (generic rank-0 target)
    std::vector<double> ha(33 * 17), hb(17 * 33);
    std::uniform_real_distribution<double> u(-1, 1);
    for (auto &x : ha) x = u(rng);
    for (auto &x : hb) x = u(rng);
    DeviceTensor<double, 2> dA("A", {33, 17}, ha), dB("B", {17, 33}, hb);
    double F = 3.0;
    einsum(Indices{}, &F, Indices{i, j}, dA, Indices{j, i}, dB, &alg);
    double FR = 0.0;
    for (int r = 0; r < 33; ++r)
        for (int c = 0; c < 17; ++c) FR += ha[size_t(r * 17 + c)] * hb[size_t(c * 33 + r)];
    EXPECT(std::fabs(F - FR) <= 1e-13 * std::fabs(FR), "scalar full contraction %.17g vs %.17g", F, FR);
}

static void smart_pointers() {
    auto pA = std::make_shared<Tensor<double, 2>>(true, "A", 24, 36);
    auto pB = std::make_shared<Tensor<double, 2>>(true, "B", 36, 28);
    auto pC = std::make_shared<Tensor<double, 2>>(true, "C", 24, 28);
    fill(*pA);
    fill(*pB);
    fill(*pC);
    Tensor<double, 2> R = *pC;
    ta::AlgorithmChoice alg = ta::INDETERMINATE;
    einsum(1.0, Indices{i, j}, &pC, -1.0, Indices{i, kk}, pA, Indices{kk, j}, pB, &alg);
    auto cd = hdesc(R, "ij"), ad = hdesc(*pA, "ik"), bd = hdesc(*pB, "kj");
    int ralg = -1;
    oracle_einsum_f64(1.0, &cd, R.data(), -1.0, &ad, pA->data(), &bd, pB->data(), &ralg);
    EXPECT(alg == ta::GEMM, "smart-pointer algorithm %s", ta::algorithm_name(alg));
    EXPECT(rel(*pC, R) < 1e-13, "smart-pointer GEMM rel err %.3e", rel(*pC, R));
    // mixed: smart-pointer A, plain B, default prefactors
    einsum(Indices{i, j}, &pC, Indices{i, kk}, pA, Indices{kk, j}, *pB);
    R.zero();
    oracle_einsum_f64(0.0, &cd, R.data(), 1.0, &ad, pA->data(), &bd, pB->data(), &ralg);
    EXPECT(rel(*pC, R) < 1e-13, "mixed smart-pointer GEMM rel err %.3e", rel(*pC, R));
}

static void batched() {
    std::vector<DeviceTensor<double, 2>> Cs, As, Bs;
    std::vector<std::vector<double>> hA, hB;
    int const ms[] = {5, 64, 130, 17, 96}, ks[] = {7, 32, 100, 64, 1}, ns[] = {3, 48, 90, 17, 200};
    std::uniform_real_distribution<double> u(-1, 1);
    for (int q = 0; q < 5; ++q) {
        std::vector<double> a(size_t(ms[q] * ks[q])), b(size_t(ks[q] * ns[q]));
        for (auto &x : a) x = u(rng);
        for (auto &x : b) x = u(rng);
        hA.push_back(a);
        hB.push_back(b);
        As.emplace_back("A", std::array<int64_t, 2>{ms[q], ks[q]}, a);
        Bs.emplace_back("B", std::array<int64_t, 2>{ks[q], ns[q]}, b);
        Cs.emplace_back("C", ms[q], ns[q]);
    }
    ta::AlgorithmChoice alg = ta::INDETERMINATE;
    einsum(0.0, Indices{i, j}, &Cs, 1.0, Indices{i, ll}, As, Indices{ll, j}, Bs, &alg);
    EXPECT(alg == ta::GEMM, "batched algorithm %s", ta::algorithm_name(alg));
    for (int q = 0; q < 5; ++q) {
        std::vector<double> ref(size_t(ms[q] * ns[q]), 0.0);
        ein_tensor_desc cd{2, {'i', 'j'}, {ms[q], ns[q]}, {ns[q], 1}};
        ein_tensor_desc ad{2, {'i', 'l'}, {ms[q], ks[q]}, {ks[q], 1}};
        ein_tensor_desc bd{2, {'l', 'j'}, {ks[q], ns[q]}, {ns[q], 1}};
        int ralg = -1;
        oracle_einsum_f64(0.0, &cd, ref.data(), 1.0, &ad, hA[size_t(q)].data(), &bd, hB[size_t(q)].data(), &ralg);
        auto got     = host_of(Cs[size_t(q)]);
        double num = 0, den = 0;
        for (size_t e = 0; e < ref.size(); ++e) {
            num += (got[e] - ref[e]) * (got[e] - ref[e]);
            den += ref[e] * ref[e];
        }
        EXPECT(std::sqrt(num / den) < 1e-13, "batched entry %d rel err %.3e", q, std::sqrt(num / den));
    }
    bool threw = false;
    As.pop_back();
    try {
        einsum(0.0, Indices{i, j}, &Cs, 1.0, Indices{i, ll}, As, Indices{ll, j}, Bs);
    } catch (bad_logic const &) {
        threw = true;
    }
    EXPECT(threw, "batched einsum with lists of different lengths must throw bad_logic");
}

static void host_permute() {
    Tensor<double, 3> A(true, "A", 7, 9, 11), C("C", 11, 7, 9); // C column-major
    fill(A);
    fill(C);
    Tensor<double, 3> R = C;
    permute(Indices{k, i, j}, &C, Indices{i, j, k}, A);
    int64_t dA[3] = {7, 9, 11}, sA[3] = {A.stride(0), A.stride(1), A.stride(2)};
    int64_t sC[3] = {R.stride(0), R.stride(1), R.stride(2)};
    int perm[3]   = {2, 0, 1};
    oracle_permute_f64(3, dA, perm, 1.0, A.data(), sA, 0.0, R.data(), sC);
    bool same = true;
    C.for_each_offset([&](int64_t o) { same = same && C.data()[o] == R.data()[o]; });
    EXPECT(same, "host permute kij <- ijk is not bit-exact");
}

int main() {
    host_tensors();
    scalar_c();
    smart_pointers();
    batched();
    host_permute();
    std::printf("checker host BLAS: %s\n", oracle_blas_name());
    if (failures == 0) std::printf("PASS test_api_overloads\n");
    return failures == 0 ? 0 : 1;
}
