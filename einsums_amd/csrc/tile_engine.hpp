// Direct-to-LDS tile staging for the MFMA GEMM kernels (gfx950).
//
// A K-slab of one operand (ROWS rows x BK contraction steps) is copied from global memory
// straight into LDS by `global_load_lds_dword{,x4}`: lane l of one wave instruction writes
// LDS dwords [M0 + VW*l, M0 + VW*(l+1)), so an instruction fills 64*VW contiguous LDS
// dwords and no VGPRs or ds_write instructions are spent on staging. Each lane chooses
// which global dwords land in its LDS slot, which is what the two layouts below are
// built on. A lane whose element lies outside the problem reads a zero page instead.
//
// Layouts (positions in elements inside one slab buffer):
//  - KMAJOR (source contiguous along the rows, e.g. column-major A and i):
//      pos(row, k) = k * PITCH + row,  PITCH = ROWS + 16.
//    The 16-element pad puts consecutive k-rows 32 banks (FP64) / 16 banks (FP32) apart,
//    so the 16 rows x 2 (FP64) or 4 (FP32) k values one MFMA fragment read touches are
//    on distinct banks.
//  - KROW (source contiguous along k): each tile row holds its BK elements as 16-byte
//    chunks (CE = 16 bytes / element elements each), the chunk index XOR-swizzled by
//    SW(row) = (row / RB) % NC, RB = rows per 64 banks, NC = chunks per row:
//      pos(row, k) = row * BK + ((k / CE) ^ SW(row)) * CE + k % CE.
//    One DMA instruction fills whole rows, so every lane group reads contiguous 128-byte
//    (FP64) or 64-byte (FP32) row segments; the swizzle keeps the fragment reads (16 rows,
//    2 or 4 consecutive k) on distinct banks.
//
// Instructions per slab are spread over the 4 waves of the workgroup (n = wave + 4x).
// The index maps are checked on the host by tests/cpp/test_lds_layout.cpp.
#pragma once

#include "gemm_core.hpp"

namespace ein {

#define EIN_LDS_PTR(p) ((__attribute__((address_space(3))) void *)(p))

// Zero dwords read by lanes whose element is outside the problem (never written).
static __device__ __attribute__((aligned(16))) uint32_t ein_zero_page[4];

template <typename T, int ROWS, int BK, bool KROW, bool WIDE>
struct SlabLayout {
    static constexpr int EW    = int(sizeof(T)) / 4;     // dwords per element
    static constexpr int VW    = WIDE ? 4 : 1;           // dwords per lane per instruction
    static constexpr int CE    = 4 / EW;                 // elements per 16-byte chunk
    static constexpr int NC    = BK / CE;                // chunks per KROW row
    static constexpr int RB    = 64 / (BK * EW);         // KROW rows per 64 banks
    static constexpr int PITCH = ROWS + 16;              // KMAJOR k-row pitch (elements)
    static constexpr int SIZE  = KROW ? ROWS * BK : BK * PITCH; // elements per buffer
    static constexpr int SLAB_DW = ROWS * BK * EW;       // dwords moved per slab
    static constexpr int INS   = SLAB_DW / (64 * VW);    // wave instructions per slab
    static constexpr int INS_PER_WAVE = INS / 4;
    // KMAJOR: instructions per k-row (a wide instruction must not span two padded k-rows)
    static constexpr int IPR = KROW ? 1 : (ROWS * EW) / (64 * VW);
    static constexpr bool VALID = IPR >= 1 && INS % 4 == 0 && BK % CE == 0 && RB >= 1 && NC >= 1;

    static __host__ __device__ constexpr int sw(int row) { return (row / RB) % NC; }
    static __host__ __device__ constexpr int pos(int row, int k) {
        return KROW ? row * BK + ((k / CE) ^ sw(row)) * CE + k % CE : k * PITCH + row;
    }
    // KMAJOR: instruction n = w + 4x (wave w, x-th instruction of the wave) covers part
    // n % IPR of k-row n / IPR. Written per case so that the compiler sees that a lane
    // touches only IPR/4 (or one) distinct rows.
    static __host__ __device__ constexpr int part_of(int w, int x) {
        return IPR % 4 == 0 ? w + 4 * (x % (IPR / 4)) : w % IPR;
    }
    static __host__ __device__ constexpr int group_of(int w, int x) {
        return IPR % 4 == 0 ? x / (IPR / 4) : w / IPR + (4 / IPR) * x;
    }
    // First LDS dword written by instruction (w, x).
    static __host__ __device__ constexpr int ins_base_dw(int w, int x) {
        return KROW ? (w + 4 * x) * 64 * VW : group_of(w, x) * PITCH * EW + part_of(w, x) * 64 * VW;
    }
    // Element (row, k) and dword half h that lane l of instruction (w, x) copies (for WIDE,
    // the first element of its 16-byte chunk; the chunk runs along rows for KMAJOR, along
    // k for KROW).
    static __host__ __device__ constexpr void lane_elem(int w, int x, int l, int &row, int &k, int &h) {
        if (KROW) {
            int const d  = (w + 4 * x) * 64 * VW + l * VW; // dword inside the slab buffer
            int const e  = d / EW;
            row          = e / BK;
            int const cs = (e % BK) / CE; // chunk slot in the row
            k            = (cs ^ sw(row)) * CE + e % CE;
            h            = d % EW;
        } else {
            int const o = part_of(w, x) * 64 * VW + l * VW; // dword inside the k-row
            row         = o / EW;
            k           = group_of(w, x);
            h           = o % EW;
        }
    }
};

// Issues the DMA of the slab starting at contraction step k0. `off(row, k0, k, check, valid)`
// returns the element offset of (tile row, contraction step k0 + k) from `base` and sets
// `valid` to whether that element is inside the problem. With CHECK false the caller knows
// the whole slab is inside (an interior tile and slab), the checks drop out and `check`
// tells the functor it may skip its own guards. For a 16-byte chunk, (row, k) is its first
// element and the host guarantees the chunk is entirely inside or entirely outside.
// k0 enters separately from k so that k0 * stride is formed once per slab (scalar) and
// k * stride once per kernel.
template <typename T, int ROWS, int BK, bool KROW, bool WIDE, bool CHECK = true, typename OffsetFn>
__device__ __forceinline__ void dma_slab(T const *__restrict__ base, T *lds, int wave, int lane, int64_t k0,
                                         OffsetFn const &off) {
    using L = SlabLayout<T, ROWS, BK, KROW, WIDE>;
    static_assert(L::VALID, "slab layout not realisable");
    uint32_t const *bdw = reinterpret_cast<uint32_t const *>(base);
#pragma unroll
    for (int x = 0; x < L::INS_PER_WAVE; ++x) {
        int row, k, h;
        L::lane_elem(wave, x, lane, row, k, h);
        // a KROW lane visits a different row per instruction: keep the row offsets from
        // being hoisted out of the slab loop (they would pin 2*INS_PER_WAVE registers)
        if constexpr (KROW) asm volatile("" : "+v"(row));
        bool valid      = true;
        int64_t const e = off(row, k0, k, CHECK, valid);
        void const *src = (!CHECK || valid) ? static_cast<void const *>(bdw + e * L::EW + h)
                                            : static_cast<void const *>(ein_zero_page);
        uint32_t *dst   = reinterpret_cast<uint32_t *>(lds) + L::ins_base_dw(wave, x);
        __builtin_amdgcn_global_load_lds(src, EIN_LDS_PTR(dst), 4 * L::VW, 0, 0);
    }
}

// MFMA work of one slab from LDS, with the fragments of step kk+4 read while the MFMAs of
// step kk run. As/Bs are the slab buffers of A (rows along i) and B (rows along j), An/Bn
// the buffers the next slab's DMA fills: `dma(An, Bn)` runs after the first fragments are
// requested and before the first MFMA, so the fragment reads are not queued behind it.
// All four are restrict parameters of this (inlined) function: the compiler then knows
// that the fragment reads cannot alias the in-flight DMA and does not wait for it before
// every LDS read of the slab.
template <typename T, int BM, int BN, int BK, bool AKG, bool AW, bool BKG, bool BW, typename Dma>
__device__ __forceinline__ void mfma_slab_lds(T const *__restrict__ As, T const *__restrict__ Bs, T *__restrict__ An,
                                              T *__restrict__ Bn,
                                              WaveAcc<T, TileCfg<T, BM, BN, BK>::FM, TileCfg<T, BM, BN, BK>::FN> &acc,
                                              int wm, int wn, int lane, Dma const &dma) {
    using Cfg = TileCfg<T, BM, BN, BK>;
    using LA  = SlabLayout<T, BM, BK, AKG, AW>;
    using LB  = SlabLayout<T, BN, BK, BKG, BW>;
    int const r = lane & 15;
    int const q = lane >> 4;
    T af[2][Cfg::FM], bf[2][Cfg::FN];
    auto frag = [&](int kk, int slot) {
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a) af[slot][a] = As[LA::pos(wm * Cfg::WM + a * 16 + r, kk + q)];
#pragma unroll
        for (int b = 0; b < Cfg::FN; ++b) bf[slot][b] = Bs[LB::pos(wn * Cfg::WN + b * 16 + r, kk + q)];
    };
    frag(0, 0);
    dma(An, Bn);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
        int const cur = (kk / 4) & 1;
        if (kk + 4 < BK) frag(kk + 4, cur ^ 1);
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
            for (int b = 0; b < Cfg::FN; ++b) acc.c[a][b] = Mfma<T>::mma(bf[cur][b], af[cur][a], acc.c[a][b]);
        // LLVM's MFMA/LDS interleaving strategy for this region (measured +1% on FP64
        // 4096^3..16384^2; the wave-priority alternative measured -2%)
        __builtin_amdgcn_iglp_opt(0);
    }
}

// Staging choice for one operand (host side): KROW layout when the source is contiguous
// along k, 16-byte DMA when every chunk is contiguous, 16-byte aligned and entirely
// inside or outside the problem.
struct OperandStaging {
    bool krow;
    bool wide;
};

template <typename T>
inline OperandStaging choose_staging(int rows_tile, int64_t s_row, int64_t s_k, int64_t s_batch, int64_t batch,
                                     int64_t rows, int64_t K, void const *ptr) {
    constexpr int EW = int(sizeof(T)) / 4;
    constexpr int KG = 4 / EW;
    auto mag         = [](int64_t x) { return x < 0 ? -x : x; };
    OperandStaging st{};
    bool const row_fast = s_row == 1 || (s_k != 1 && mag(s_row) <= mag(s_k));
    st.krow             = !row_fast;
    bool const aligned  = reinterpret_cast<uintptr_t>(ptr) % 16 == 0 && (batch <= 1 || s_batch % KG == 0);
    if (!st.krow)
        st.wide = aligned && s_row == 1 && s_k % KG == 0 && rows % KG == 0 && rows_tile * EW >= 256;
    else
        st.wide = aligned && s_k == 1 && s_row % KG == 0 && K % KG == 0;
    return st;
}

// Waits for this wave's outstanding DMA, then for the whole workgroup.
__device__ __forceinline__ void dma_wait_all() {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
}

} // namespace ein
