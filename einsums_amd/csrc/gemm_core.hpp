// MFMA GEMM building blocks for gfx950 shared by the dense, strided-batched,
// gather (generic-contraction) and grouped kernels.
//
// Tile scheme: a 256-thread workgroup (4 wave64s as 2x2) owns a BM x BN tile of C and
// walks K in BK-deep slabs staged through double-buffered LDS; each wave owns a
// (BM/2) x (BN/2) sub-tile made of 16x16 MFMA blocks (v_mfma_f64_16x16x4f64 or
// v_mfma_f32_16x16x4f32).
//
// MFMA operand placement (16x16x4, wave64): lane l supplies A[r=l%16][k=l/16] and
// B[k=l/16][c=l%16] and receives four results D[r][c=l%16], v=0..3, with
// r = 4*(l/16)+v for the FP32 instruction and r = 4*v+(l/16) for the FP64 one
// (Mfma<T>::row gives the mapping; both were checked against known answers). The kernels feed
// the B side of the contraction as the MFMA "A" operand and the A side as the MFMA
// "B" operand, so lane%16 runs along C's row index i -- C's unit-stride direction
// after host-side normalisation -- and every C store is a 128-byte contiguous run.
#pragma once

#include "ein_common.hpp"

namespace ein {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    using acc_t = v4d;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of the 16x16 result held in accumulator element v by lane group q = lane/16
    static __host__ __device__ constexpr int row(int q, int v) { return 4 * v + q; }
};

template <>
struct Mfma<float> {
    using acc_t = v4f;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __host__ __device__ constexpr int row(int q, int v) { return 4 * q + v; }
};

constexpr int GEMM_THREADS = 256;

template <typename T, int BM, int BN, int BK>
struct TileCfg {
    static constexpr int WM  = BM / 2;
    static constexpr int WN  = BN / 2;
    static constexpr int FM  = WM / 16;
    static constexpr int FN  = WN / 16;
    static constexpr int PAD = 1;
    static constexpr int LDA = BM + PAD; // LDS row length of the A slab (indexed [k][i])
    static constexpr int LDB = BN + PAD; // LDS row length of the B slab (indexed [k][j])
    static constexpr int LA  = BM * BK / GEMM_THREADS;
    static constexpr int LB  = BN * BK / GEMM_THREADS;
    static_assert(BM % 32 == 0 && BN % 32 == 0, "tile must split into 2x2 waves of 16x16 blocks");
    static_assert((BM * BK) % GEMM_THREADS == 0 && (BN * BK) % GEMM_THREADS == 0, "slab must split over threads");
    static_assert(BK % 4 == 0, "BK is a multiple of the MFMA depth");
};

// Accumulator block of one wave.
template <typename T, int FM, int FN>
struct WaveAcc {
    typename Mfma<T>::acc_t c[FM][FN];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
            for (int b = 0; b < FN; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) c[a][b][v] = T(0);
    }
};

// One BK-deep slab of MFMAs out of LDS. As is [BK][LDA] (k-major, i fastest),
// Bs is [BK][LDB] (k-major, j fastest).
template <typename T, int BM, int BN, int BK>
__device__ __forceinline__ void mfma_slab(T const *__restrict__ As, T const *__restrict__ Bs,
                                          WaveAcc<T, TileCfg<T, BM, BN, BK>::FM, TileCfg<T, BM, BN, BK>::FN> &acc,
                                          int wm, int wn, int lane) {
    using Cfg  = TileCfg<T, BM, BN, BK>;
    int const r = lane & 15;
    int const q = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
        T af[Cfg::FM], bf[Cfg::FN];
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a) af[a] = As[(kk + q) * Cfg::LDA + wm * Cfg::WM + a * 16 + r];
#pragma unroll
        for (int b = 0; b < Cfg::FN; ++b) bf[b] = Bs[(kk + q) * Cfg::LDB + wn * Cfg::WN + b * 16 + r];
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
            for (int b = 0; b < Cfg::FN; ++b) acc.c[a][b] = Mfma<T>::mma(bf[b], af[a], acc.c[a][b]);
    }
}

// Grouped (L2-friendly) ordering of output tiles: walks GROUP rows of tiles
// column by column so neighbouring workgroups share A rows and B columns.
__host__ __device__ __forceinline__ void tile_coords(uint32_t logical, uint32_t tiles_m, uint32_t tiles_n, uint32_t &tm,
                                            uint32_t &tn) {
    constexpr uint32_t GROUP = 8;
    uint32_t per_group = GROUP * tiles_n;
    uint32_t g         = logical / per_group;
    uint32_t first_m   = g * GROUP;
    uint32_t gsize     = tiles_m - first_m < GROUP ? tiles_m - first_m : GROUP;
    uint32_t in_g      = logical - g * per_group;
    tm                 = first_m + in_g % gsize;
    tn                 = in_g / gsize;
}

// ------------------------------------------------------- stream-K schedule
// The work of a GEMM-like launch is (tile, BK-slab) units, tile-major. A hybrid schedule
// keeps L2 locality and still fills the last round:
// - tiles [0, dp_tiles) are data-parallel: workgroup b < dp_tiles owns one whole tile
//   (XCD-remapped, so tiles running at the same time are neighbours on one XCD);
// - the remaining tiles (those of the last, partial round) are shared stream-K style by
//   the last grid - dp_tiles workgroups: stream-K workgroup w owns the contiguous unit
//   range base + [w*U, min(W, (w+1)*U)) with base = dp_tiles*slabs. Whole tiles inside it
//   are finished in place; the (at most two) tiles cut by its range ends are written as
//   partial accumulators to piece slots 2w (first tile it touches) and 2w+1 (last tile).
// A fix-up pass over the stream-K tiles (sk_fixup_grid workgroups per tile) sums the pieces
// of every cut tile in workgroup order, so results are deterministic and no atomics are used.
struct StreamK {
    uint32_t tiles;    // tiles in the launch (batch included)
    uint32_t slabs;    // BK-slabs per tile
    uint32_t dp_tiles; // tiles owned whole by one workgroup each
    uint32_t grid;     // workgroups launched: dp_tiles + stream-K workgroups
    uint64_t units;    // units per stream-K workgroup; 0 = no stream-K part
    uint32_t ksplit;   // > 1: split-K schedule (plan_split_k), 0 otherwise
    uint32_t kchunk;   // split-K: slabs per contraction chunk
};

// Decides the schedule for `tiles` tiles of `slabs` slabs with `slots` workgroups
// resident at once.
inline StreamK plan_stream_k(uint64_t tiles, uint64_t slabs, uint64_t slots) {
    StreamK sk{};
    sk.tiles = uint32_t(tiles);
    sk.slabs = uint32_t(slabs);
    uint64_t rounds = slots ? (tiles + slots - 1) / slots : 1;
    double eff      = slots ? double(tiles) / double(rounds * slots) : 1.0;
    if (tiles == 0 || slabs == 0 || slots == 0 || eff >= 0.97 || tiles * slabs < 2 * slots) {
        sk.dp_tiles = uint32_t(tiles);
        sk.grid     = uint32_t(tiles);
        sk.units    = 0;
        return sk;
    }
    // every full round runs data-parallel; only the tiles of the last, partial round are
    // shared (each by several workgroups), which keeps the fix-up small
    uint64_t full = tiles / slots;
    uint64_t dp   = full * slots;
    dp -= dp % 8; // keep the stream-K workgroups on the same XCD pattern
    uint64_t W    = (tiles - dp) * slabs;
    uint64_t G    = slots < W ? slots : W;
    sk.dp_tiles   = uint32_t(dp);
    sk.units      = (W + G - 1) / G;
    sk.grid       = uint32_t(dp + (W + sk.units - 1) / sk.units);
    return sk;
}

// Split-K schedule for launches with less than half a round of tiles and a long
// contraction (the 576 x 576 x 9216 Wmnij GEMM: 81 tiles of 64 x 64): the contraction is cut
// into P chunks of equal length and workgroup w (XCD-remapped) owns chunk w / tiles of tile
// w % tiles. Consecutive workgroups, which share an XCD and its L2, thus work on all tiles
// of one chunk at the same contraction offsets, so each A and B panel slab is fetched from
// HBM / MALL about once per XCD; a stream-K range instead spans every chunk of a few tiles.
// Every workgroup writes its partial tile to piece slot w; the fix-up sums the P pieces of a
// tile in chunk order (deterministic). Falls back to plan_stream_k where it does not apply.
inline StreamK plan_split_k(uint64_t tiles, uint64_t slabs, uint64_t slots) {
    if (tiles == 0 || slabs == 0 || 2 * tiles > slots) return plan_stream_k(tiles, slabs, slots);
    uint64_t p = slots / tiles;
    if (p > slabs / 8) p = slabs / 8; // chunks of at least 8 slabs
    if (p < 2) return plan_stream_k(tiles, slabs, slots);
    uint64_t const chunk = (slabs + p - 1) / p;
    p                    = (slabs + chunk - 1) / chunk; // every chunk non-empty
    StreamK sk{};
    sk.tiles    = uint32_t(tiles);
    sk.slabs    = uint32_t(slabs);
    sk.dp_tiles = 0;
    sk.grid     = uint32_t(tiles * p);
    sk.units    = chunk; // non-zero: pieces and a fix-up pass
    sk.ksplit   = uint32_t(p);
    sk.kchunk   = uint32_t(chunk);
    return sk;
}

// Unit range [u0, u1) of launched workgroup `block`, and its stream-K index w (0 for a
// data-parallel workgroup; the logical index under split-K).
__host__ __device__ __forceinline__ void sk_range_of(StreamK const &sk, uint32_t block, uint64_t &u0, uint64_t &u1,
                                                     uint64_t &w) {
    if (sk.ksplit) {
        w                   = xcd_remap(block, sk.grid);
        uint64_t const c    = w / sk.tiles;
        uint64_t const t    = w - c * sk.tiles;
        uint64_t const end  = (c + 1) * sk.kchunk;
        u0                  = t * sk.slabs + c * sk.kchunk;
        u1                  = t * sk.slabs + (end < sk.slabs ? end : sk.slabs);
        return;
    }
    if (block < sk.dp_tiles) {
        uint64_t const t = xcd_remap(block, sk.dp_tiles);
        u0               = t * sk.slabs;
        u1               = u0 + sk.slabs;
        w                = 0;
        return;
    }
    uint64_t const W    = uint64_t(sk.tiles) * sk.slabs;
    uint64_t const base = uint64_t(sk.dp_tiles) * sk.slabs;
    w                   = xcd_remap(block - sk.dp_tiles, sk.grid - sk.dp_tiles);
    u0                  = base + w * sk.units;
    u1                  = u0 + sk.units < W ? u0 + sk.units : W;
}

__device__ __forceinline__ void sk_range(StreamK const &sk, uint64_t &u0, uint64_t &u1, uint64_t &w) {
    sk_range_of(sk, blockIdx.x, u0, u1, w);
}

// Piece slot of tile t for stream-K workgroup w, where w's unit range meets t: slot 2w when
// the range starts inside t (t is w's first tile), 2w+1 otherwise. No division, so the
// fix-up can evaluate it per piece.
__host__ __device__ __forceinline__ uint64_t sk_slot(StreamK const &sk, uint64_t w, uint64_t t) {
    if (sk.ksplit) return w; // split-K: one piece per workgroup, slot c * tiles + t
    uint64_t const start = uint64_t(sk.dp_tiles) * sk.slabs + w * sk.units;
    return 2 * w + (start >= t * sk.slabs ? 0 : 1);
}

// Stream-K workgroups [g0, g1] whose unit ranges meet tile t (t >= dp_tiles); the tile was
// finished in place when g0 == g1.
__host__ __device__ __forceinline__ void sk_owners(StreamK const &sk, uint64_t t, uint64_t &g0, uint64_t &g1) {
    uint64_t const off = (t - sk.dp_tiles) * sk.slabs;
    g0                 = off / sk.units;
    g1                 = (off + sk.slabs - 1) / sk.units;
}

// Fix-up launch shape: each thread sums FIXUP_EPT elements of one cut tile, so a BM x BN
// tile takes BM*BN / (256*FIXUP_EPT) workgroups (grid x) and the grid's y runs over the
// stream-K tiles; every thread issues FIXUP_EPT independent loads per piece.
constexpr int FIXUP_EPT = 4;
inline dim3 sk_fixup_grid(StreamK const &sk, int tile_elems) {
    return dim3(uint32_t(tile_elems / (256 * FIXUP_EPT)), sk.tiles - sk.dp_tiles, 1);
}

// Number of piece slots a launch needs.
inline uint64_t sk_piece_slots(StreamK const &sk) {
    return sk.ksplit ? uint64_t(sk.grid) : sk.units ? 2ull * (sk.grid - sk.dp_tiles) : 0;
}

} // namespace ein
