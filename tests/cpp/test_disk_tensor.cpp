// C++ API: DiskTensor (HDF5) and linear_algebra::gemm with C a DiskTensor.
//   test_disk_tensor cpu  -- create / put / get round trip and refusals (no GPU)
//   test_disk_tensor gpu  -- Wabef-style update of a (v, v, v, v) disk tensor in several
//                            panels, against a long-double triple loop on the host
// Files go to the directory in $TMPDIR (default /tmp). Exit status 0 = all checks passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <unistd.h>
#include <vector>

#include "einsums/DiskTensor.hpp"
#include "einsums/Einsums.hpp"

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

static std::string scratch_file(char const *tag) {
    char const *dir = std::getenv("TMPDIR");
    std::string p   = std::string(dir && *dir ? dir : "/tmp") + "/einsums_disk_" + tag + "_" + std::to_string(getpid()) + ".h5";
    std::remove(p.c_str());
    return p;
}

static void cpu_checks() {
    std::string const path = scratch_file("cpu");
    DiskTensor<double, 3> D(path, "D", 3, 4, 5);
    auto z = D.get();
    bool zero = true;
    for (int64_t i = 0; i < z.size(); ++i) zero = zero && z.data()[i] == 0.0;
    EXPECT(zero, "a new disk tensor reads as zeros");
    Tensor<double, 3> t(true, "t", 3, 4, 5);
    for (int64_t i = 0; i < t.size(); ++i) t.data()[i] = double(i) / 7.0;
    D.put(t);
    auto back = DiskTensor<double, 3>(path, "D", 3, 4, 5).get();
    bool same = true;
    for (int64_t i = 0; i < t.size(); ++i) same = same && back.data()[i] == t.data()[i];
    EXPECT(same, "put / get round trip is bit-exact");
    EXPECT(back(2, 3, 4) == t(2, 3, 4), "element (2,3,4)");
    bool threw = false;
    try {
        DiskTensor<double, 3> bad(path, "D", 3, 4, 6);
    } catch (dimension_error const &) { threw = true; }
    EXPECT(threw, "reopening with other extents throws dimension_error");
    threw = false;
    try {
        DiskTensor<float, 3> bad(path, "D", 3, 4, 5);
    } catch (std::domain_error const &) { threw = true; }
    EXPECT(threw, "reopening FP64 data as float throws");
    threw = false;
    try {
        Tensor<double, 3> cm(false, "cm", 3, 4, 5); // column-major
        D.put(cm);
    } catch (tensor_compat_error const &) { threw = true; }
    EXPECT(threw, "put of a column-major tensor throws");
    std::remove(path.c_str());
}

static void gpu_checks() {
    // Wabef(a,b,e,f) = 0.5 W0 + 0.25 sum_mn t(mn, ab) g(mn, ef) with W a (v, v, v, v) disk tensor
    int const o = 3, v = 10, V2 = v * v, O2 = o * o;
    std::mt19937_64 gen(9);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> t(size_t(O2) * V2), g(size_t(O2) * V2);
    for (auto &x : t) x = u(gen);
    for (auto &x : g) x = u(gen);
    std::string const path = scratch_file("gpu");
    DiskTensor<double, 4> W(path, "Wabef", v, v, v, v);
    Tensor<double, 4> W0(true, "W0", v, v, v, v);
    for (int64_t i = 0; i < W0.size(); ++i) W0.data()[i] = u(gen);
    W.put(W0);
    DeviceTensor<double, 2> T("t", std::array<int64_t, 2>{O2, V2}, t);
    DeviceTensor<double, 2> G("g", std::array<int64_t, 2>{O2, V2}, g);
    // one leading slice is v^3 doubles = 8000 B: 20 000 B -> 2 slices a panel, 5 panels
    linear_algebra::gemm<true, false>(0.25, T, G, 0.5, &W, 20000);
    auto got = W.get();
    double err = 0, ref_norm = 0;
    for (int ab = 0; ab < V2; ++ab)
        for (int ef = 0; ef < V2; ++ef) {
            long double s = 0;
            for (int mn = 0; mn < O2; ++mn) s += (long double)t[size_t(mn) * V2 + ab] * g[size_t(mn) * V2 + ef];
            double const ref = double(0.5L * W0.data()[size_t(ab) * V2 + ef] + 0.25L * s);
            err += (got.data()[size_t(ab) * V2 + ef] - ref) * (got.data()[size_t(ab) * V2 + ef] - ref);
            ref_norm += ref * ref;
        }
    EXPECT(std::sqrt(err / ref_norm) < 1e-14, "Wabef on disk: relative error %g", std::sqrt(err / ref_norm));
    bool threw = false;
    try {
        DiskTensor<double, 2> wrong(path, "wrong", V2 + 1, V2);
        linear_algebra::gemm<true, false>(1.0, T, G, 0.0, &wrong);
    } catch (dimension_error const &) { threw = true; }
    EXPECT(threw, "gemm into a disk tensor of other size throws dimension_error");
    std::remove(path.c_str());
}

int main(int argc, char **argv) {
    std::string const mode = argc > 1 ? argv[1] : "cpu";
    try {
        if (mode == "cpu") cpu_checks();
        else gpu_checks();
    } catch (std::exception const &e) {
        std::printf("FAIL exception: %s\n", e.what());
        ++failures;
    }
    if (failures == 0) std::printf("PASS test_disk_tensor %s\n", mode.c_str());
    return failures ? 1 : 0;
}
