// Host check of the direct-to-LDS slab layouts (einsums_amd/csrc/tile_engine.hpp): for
// every realisable (element type, rows, layout, DMA width), the dwords written by the
// 4 waves x INS_PER_WAVE instructions x 64 lanes cover every (row, k, half) of the slab
// exactly once, each at the LDS dword the MFMA fragment reads expect (pos(row, k)), inside
// the buffer; and the fragment reads of one instruction are free of LDS bank conflicts
// (64 banks of 4 bytes; 64-bit reads are served 32 lanes at a time).
#include <cstdio>
#include <set>
#include <vector>

#include "tile_engine.hpp"

using namespace ein;

static int failures = 0;
#define CHECK(c, ...)                                                                                                  \
    do {                                                                                                               \
        if (!(c)) {                                                                                                    \
            ++failures;                                                                                                \
            std::printf("FAIL: " __VA_ARGS__);                                                                         \
            std::printf("\n");                                                                                         \
            return;                                                                                                    \
        }                                                                                                              \
    } while (0)

template <typename T, int ROWS, int BK, bool KROW, bool WIDE>
void check_layout(char const *name) {
    using L = SlabLayout<T, ROWS, BK, KROW, WIDE>;
    if constexpr (!L::VALID) {
        std::printf("  %-34s not realisable (skipped by the launcher)\n", name);
        return;
    } else {
        constexpr int EW = L::EW;
        std::vector<int> seen(size_t(ROWS) * BK * EW, 0);
        std::vector<int> lds_owner(size_t(L::SIZE) * EW, -1);
        for (int w = 0; w < 4; ++w)
            for (int x = 0; x < L::INS_PER_WAVE; ++x)
                for (int l = 0; l < 64; ++l) {
                    int row, k, h;
                    L::lane_elem(w, x, l, row, k, h);
                    for (int j = 0; j < L::VW; ++j) {
                        // dword j of the lane's chunk
                        int r2 = row, k2 = k, h2 = h + j;
                        if (WIDE) {
                            int e = j / EW;
                            h2    = j % EW;
                            if (KROW) k2 = k + e; else r2 = row + e;
                        }
                        CHECK(r2 >= 0 && r2 < ROWS && k2 >= 0 && k2 < BK && h2 < EW, "%s: element out of the slab", name);
                        int dw = L::ins_base_dw(w, x) + l * L::VW + j;
                        CHECK(dw >= 0 && dw < L::SIZE * EW, "%s: LDS dword %d outside the buffer", name, dw);
                        CHECK(dw == L::pos(r2, k2) * EW + h2, "%s: lane writes (%d,%d,%d) at %d, reads expect %d", name,
                              r2, k2, h2, dw, L::pos(r2, k2) * EW + h2);
                        CHECK(lds_owner[size_t(dw)] < 0, "%s: LDS dword %d written twice", name, dw);
                        lds_owner[size_t(dw)] = 1;
                        seen[(size_t(r2) * BK + size_t(k2)) * EW + size_t(h2)]++;
                    }
                }
        for (size_t i = 0; i < seen.size(); ++i) CHECK(seen[i] == 1, "%s: slab dword %zu copied %d times", name, i, seen[i]);
        // fragment reads: lane (r, q) reads element (base + r, kk + q)
        for (int base = 0; base < ROWS; base += 16)
            for (int kk = 0; kk < BK; kk += 4) {
                int const group = EW == 2 ? 32 : 64; // lanes served together
                for (int g0 = 0; g0 < 64; g0 += group) {
                    std::set<int> banks;
                    for (int l = g0; l < g0 + group; ++l) {
                        int r = l & 15, q = l >> 4;
                        int dw = L::pos(base + r, kk + q) * EW;
                        for (int h = 0; h < EW; ++h) {
                            bool fresh = banks.insert((dw + h) % 64).second;
                            CHECK(fresh, "%s: bank conflict in the fragment read (rows %d.., k %d..)", name, base, kk);
                        }
                    }
                }
            }
        std::printf("  %-34s ok (%d instructions per wave)\n", name, L::INS_PER_WAVE);
    }
}

int main() {
    check_layout<double, 128, 16, false, false>("f64 128 kmajor dword");
    check_layout<double, 128, 16, false, true>("f64 128 kmajor x4");
    check_layout<double, 128, 16, true, false>("f64 128 krow dword");
    check_layout<double, 128, 16, true, true>("f64 128 krow x4");
    check_layout<double, 64, 16, false, false>("f64 64 kmajor dword");
    check_layout<double, 64, 16, false, true>("f64 64 kmajor x4");
    check_layout<double, 64, 16, true, false>("f64 64 krow dword");
    check_layout<double, 64, 16, true, true>("f64 64 krow x4");
    check_layout<double, 96, 16, true, false>("f64 96 krow dword");
    check_layout<double, 96, 16, true, true>("f64 96 krow x4");
    check_layout<float, 96, 16, true, false>("f32 96 krow dword");
    check_layout<float, 96, 16, true, true>("f32 96 krow x4");
    check_layout<float, 128, 16, false, false>("f32 128 kmajor dword");
    check_layout<float, 128, 16, false, true>("f32 128 kmajor x4");
    check_layout<float, 128, 16, true, false>("f32 128 krow dword");
    check_layout<float, 128, 16, true, true>("f32 128 krow x4");
    check_layout<float, 64, 16, false, false>("f32 64 kmajor dword");
    check_layout<float, 64, 16, true, false>("f32 64 krow dword");
    check_layout<float, 64, 16, true, true>("f32 64 krow x4");
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
