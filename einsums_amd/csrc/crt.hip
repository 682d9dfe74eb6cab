// Split-precision FP64 GEMM by the Chinese remainder theorem (the integer modular scheme of
// H. Ozaki, Y. Uchino, T. Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of
// floating-point matrix multiplication using an integer modular technique", 2024),
// on the int8 MFMA of gfx950:
//
//   1. Scaling: A'(i,p) = rint(A(i,p) 2^(kb - e_i)) with max_p |A(i,p)| < 2^e_i, and
//      B'(p,j) = rint(B(p,j) 2^(kb - f_j)) per column; |A'|, |B'| <= 2^kb are integers held
//      exactly in FP64 (kb <= 53).
//   2. Residues: for each of the L pairwise coprime moduli m_l <= 256, a_l = A' mod m_l and
//      b_l = B' mod m_l as symmetric representatives, which fit int8.
//   3. One int8 GEMM per modulus, exact in int32 (|a_l b_l| <= 2^14 over at most 65536
//      contraction steps per launch; longer contractions add the residues of several
//      launches): R_l = (a_l b_l) mod m_l in [0, m_l), stored as bytes.
//   4. Reconstruction: S = sum_l R_l W_l with W_l the CRT basis of M = prod m_l, and
//      C' = S - M rint(S / M), the representative in (-M/2, M/2). kb is the largest value
//      with 4 K 2^(2 kb) <= M, so |A'B'| <= M/4 and C' = A'B' exactly. W_l and M are kept
//      as four 32-bit limbs (M < 2^128 for 16 moduli), so every limb sum of S and every limb
//      of S - qM is an exact FP64 integer below 2^45; the limbs of C' are then added as a
//      double-double (error-free two-sums), which rounds C' once even where its limbs cancel.
//      C = alpha C' 2^(e_i + f_j - 2 kb) + beta C.
//
// Every product of A' and B' is formed exactly, so the only error is the rounding of A
// and B to kb bits relative to their row (column) maximum. The number of moduli sets kb:
// 13 moduli give kb = 44 at K = 2304 and 42 at K = 65536.
//
// Against the 21 products of six int8 slices, L moduli take L int8 GEMMs, each with one
// int32 accumulator per output: the GEMM here can use 256 x 256 tiles and a four-deep
// LDS ring, where six slices leave room for 128 x 64 tiles and two stages.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <string>

#include "ozaki.hpp"
#include "tile_engine.hpp"

namespace ein {

namespace {

constexpr int CRT_MAX_MODS = 16;
// pairwise coprime, descending from 256; the first L are used
constexpr int32_t CRT_MODS[CRT_MAX_MODS] = {256, 255, 253, 251, 247, 241, 239, 233,
                                            229, 227, 223, 211, 199, 197, 193, 191};
constexpr int CRT_FOLD_STAGES = 1024; // stages of one launch, 65536 contraction steps: |int32 sum| <= 2^30

typedef int v4i __attribute__((ext_vector_type(4)));

struct CrtMods {
    int32_t nmod;
    int32_t kbits;
    int32_t mod[CRT_MAX_MODS];
    uint32_t magic[CRT_MAX_MODS]; // floor(2^32 / m)
    uint32_t bias[CRT_MAX_MODS];  // m * ceil(2^30 / m): makes an int32 in [-2^30, 2^30 + 2^8) non-negative
    double rinv[CRT_MAX_MODS];    // 1 / m
};

struct CrtBasis {
    int32_t nmod;
    int32_t shift;             // 2 kb
    double w[CRT_MAX_MODS][4]; // W_l = w0 + w1 2^32 + w2 2^64 + w3 2^96
    double mlimb[4];           // M in the same limbs
    double inv_m;              // 1 / M (rounded)
};

// ------------------------------------------------------------------ residues
// 64 x 64 tiles of X staged through LDS so that the read follows X's unit-stride axis;
// residues written along p: out[l][r][kp]. Each thread takes four consecutive p of four rows
// and writes one 32-bit word (four residues) per row and modulus.
template <bool ROWS_FAST>
__global__ __launch_bounds__(256) void crt_residue_kernel(SliceSrc S, int32_t const *__restrict__ e, CrtMods K,
                                                          int8_t *__restrict__ out) {
    __shared__ double tile[64][65];
    int64_t const r0 = int64_t(blockIdx.y) * 64;
    int64_t const p0 = int64_t(blockIdx.x) * 64;
    int const t      = threadIdx.x;
    if constexpr (ROWS_FAST) {
        int const lr = t & 63;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            int lp       = (t >> 6) + 4 * q;
            int64_t r    = r0 + lr, p = p0 + lp;
            tile[lr][lp] = (r < S.rows && p < S.k) ? S.X[r * S.s_row + p * S.s_k] : 0.0;
        }
    } else {
        int const lp = t & 63;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            int lr       = (t >> 6) + 4 * q;
            int64_t r    = r0 + lr, p = p0 + lp;
            tile[lr][lp] = (r < S.rows && p < S.k) ? S.X[r * S.s_row + p * S.s_k] : 0.0;
        }
    }
    __syncthreads();
    int const p4    = (t & 15) * 4;
    int64_t const p = p0 + p4; // kp is a multiple of 64: p .. p+3 < kp whenever p < kp
    if (p >= S.kp) return;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        int const lr    = (t >> 4) + 16 * q;
        int64_t const r = r0 + lr;
        if (r >= S.rows) continue;
        int const sh = K.kbits - e[r];
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = rint(ldexp(tile[lr][p4 + u], sh)); // |x| <= 2^kb, exact
        int8_t *__restrict__ dst = out + r * S.kp + p;
#pragma unroll 1
        for (int l = 0; l < K.nmod; ++l) {
            double const m  = double(K.mod[l]);
            double const hi = K.mod[l] == 256 ? 127.0 : 0.5 * (m - 1.0);
            double const lo = K.mod[l] == 256 ? -128.0 : -hi;
            uint32_t word   = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double v = fma(-rint(x[u] * K.rinv[l]), m, x[u]); // x - m*q: an exact small integer
                v        = v > hi ? v - m : v;
                v        = v < lo ? v + m : v;
                word |= uint32_t(uint8_t(int8_t(v))) << (8 * u);
            }
            *reinterpret_cast<uint32_t *>(dst + int64_t(l) * S.rows * S.kp) = word;
        }
    }
}

// ------------------------------------------------------------- residue GEMM
// R_l(j, i) = (sum_p a_l(i,p) b_l(j,p)) mod m_l for all moduli in one launch. Workgroup
// tile 256 x 256 (i x j), eight waves (two per SIMD) of 128 x 64: 8 x 4 blocks of the
// 16 x 16 x 64 int8 MFMA, 128 int32 accumulators per lane. Operand stages of 64
// contraction steps go global -> LDS by DMA through a ring of four buffers.
constexpr int CG_TILE = 256, CG_KS = 64, CG_NBUF = 4, CG_WAVES = 8;
constexpr int CG_SLOTS = CG_TILE * 4;                 // 16-byte chunks per operand stage
constexpr int CG_DMA   = CG_SLOTS / 64 / CG_WAVES;    // DMA instructions per operand, wave and stage

struct CrtGemm {
    int64_t m, n, kp;         // rows of a (i), rows of b (j, this panel), padded contraction length
    int64_t k0, kc;           // contraction steps [k0, k0 + kc) of this launch (multiples of 64)
    int32_t accumulate;       // add to the stored residues instead of overwriting them
    int64_t a_plane, b_plane; // elements between the moduli of the residue operands
    int8_t const *A;
    int8_t const *B;
    uint8_t *R; // [l][j][i], i fastest
    int64_t ldr, r_plane;
    uint32_t tiles_m, tiles_n;
    CrtMods K;
};

// chunk c of row `row` sits at slot row*4 + (c ^ ((row >> 2) & 3)): the 16 rows of one
// fragment read fall on distinct 16-byte bank groups
__device__ __forceinline__ int cg_slot(int row, int c) { return row * 4 + (c ^ ((row >> 2) & 3)); }

// One operand stage: lane l of instruction n (n = wave, wave + 8) fills slot 64n + l.
__device__ __forceinline__ void cg_dma(int8_t const *__restrict__ X, int64_t rows, int64_t kp, int64_t row0, int64_t k0,
                                       int8_t *lds, int wave, int lane) {
#pragma unroll
    for (int x = 0; x < CG_DMA; ++x) {
        int const n     = wave + CG_WAVES * x;
        int const sl    = n * 64 + lane;
        int const row   = sl >> 2;
        int const c     = (sl & 3) ^ ((row >> 2) & 3);
        int64_t const g = row0 + row;
        void const *src = g < rows ? static_cast<void const *>(X + g * kp + k0 + 16 * c)
                                   : static_cast<void const *>(ein_zero_page);
        __builtin_amdgcn_global_load_lds(src, EIN_LDS_PTR(lds + n * 64 * 16), 16, 0, 0);
    }
}

// x mod m in [0, m) for x in [-2^30, 2^30 + 2^8): Barrett with floor(2^32/m), which is low
// by at most two.
__device__ __forceinline__ uint32_t cg_mod(int32_t x, uint32_t m, uint32_t magic, uint32_t bias) {
    uint32_t const u = uint32_t(x) + bias;
    uint32_t r       = u - __umulhi(u, magic) * m;
    r                = r >= m ? r - m : r;
    r                = r >= m ? r - m : r;
    return r;
}

__device__ __forceinline__ void cg_load_b(int8_t const *__restrict__ stage, v4i (&bf)[4], int wn, int r, int q) {
    v4i const *Bs = reinterpret_cast<v4i const *>(stage) + CG_SLOTS;
#pragma unroll
    for (int b = 0; b < 4; ++b) bf[b] = Bs[cg_slot(wn * 64 + b * 16 + r, q)];
}

__device__ __forceinline__ v4i cg_load_a(int8_t const *__restrict__ stage, int a, int wm, int r, int q) {
    return reinterpret_cast<v4i const *>(stage)[cg_slot(wm * 128 + a * 16 + r, q)];
}

// Waits until at most `pending` later stages of this wave's DMA are outstanding, then for
// the whole workgroup.
__device__ __forceinline__ void cg_wait(int pending) {
    static_assert(2 * CG_DMA == 4, "vmcnt counts below assume 4 DMA instructions per wave and stage");
    if (pending >= 3)
        asm volatile("s_waitcnt vmcnt(12)" ::: "memory");
    else if (pending == 2)
        asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
    else if (pending == 1)
        asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
    else
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
}

// Residue GEMM of one (modulus, tile) per workgroup. Stage s sits in LDS buffer s % 4, and
// four stages are in flight when the loop starts. Within stage s each wave runs its eight
// row blocks of a against the four b fragments held in registers, reading the next a
// fragment ahead of each block's MFMAs. Halfway through, it waits for stage s+1 and reads
// that stage's b fragments into the second register set; after the last block a barrier
// (every wave has read stage s) frees buffer s % 4 for the DMA of stage s+4.
// The contraction is at most CRT_FOLD_STAGES stages (|int32 sum| <= 2^30); longer ones run
// as several launches with `accumulate` set, each adding its residues to the stored ones.
__global__ __launch_bounds__(64 * CG_WAVES, 1) void crt_gemm_kernel(CrtGemm P) {
    __shared__ __attribute__((aligned(16))) int8_t lds[CG_NBUF * 2 * CG_SLOTS * 16];

    int const tid  = threadIdx.x;
    int const lane = tid & 63;
    int const wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const wm   = wave & 1;
    int const wn   = wave >> 1;
    int const r    = lane & 15;
    int const q    = lane >> 4;

    uint32_t const ntiles  = P.tiles_m * P.tiles_n;
    uint32_t const logical = xcd_remap(blockIdx.x, gridDim.x);
    uint32_t const l       = logical / ntiles;
    uint32_t const tile    = logical - l * ntiles;
    uint32_t tm, tn;
    tile_coords(tile, P.tiles_m, P.tiles_n, tm, tn);
    int64_t const i0 = int64_t(tm) * CG_TILE;
    int64_t const j0 = int64_t(tn) * CG_TILE;
    int8_t const *__restrict__ Ag = P.A + int64_t(l) * P.a_plane + P.k0;
    int8_t const *__restrict__ Bg = P.B + int64_t(l) * P.b_plane + P.k0;

    v4i acc[8][4];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = v4i{0, 0, 0, 0};

    auto buf = [&](int64_t s) { return lds + (s & (CG_NBUF - 1)) * (2 * CG_SLOTS * 16); };
    auto dma = [&](int64_t s) {
        int8_t *dst = buf(s);
        cg_dma(Ag, P.m, P.kp, i0, s * CG_KS, dst, wave, lane);
        cg_dma(Bg, P.n, P.kp, j0, s * CG_KS, dst + CG_SLOTS * 16, wave, lane);
    };
    int64_t const nst = P.kc / CG_KS;
    for (int64_t s = 0; s < nst && s < CG_NBUF; ++s) dma(s);
    cg_wait(int(nst - 1 < 3 ? nst - 1 : 3));
    v4i b0[4], b1[4];
    cg_load_b(buf(0), b0, wn, r, q);
    v4i an = cg_load_a(buf(0), 0, wm, r, q);

#define CG_STAGE(BC, BN, ST)                                                                                           \
    do {                                                                                                               \
        int8_t const *cur_ = buf(ST);                                                                                  \
        bool const more_   = (ST) + 1 < nst;                                                                           \
        _Pragma("unroll") for (int a = 0; a < 8; ++a) {                                                                \
            v4i const af = an;                                                                                         \
            if (a < 7)                                                                                                 \
                an = cg_load_a(cur_, a + 1, wm, r, q);                                                                 \
            else if (more_)                                                                                            \
                an = cg_load_a(buf((ST) + 1), 0, wm, r, q);                                                            \
            _Pragma("unroll") for (int b = 0; b < 4; ++b) acc[a][b] =                                                  \
                __builtin_amdgcn_mfma_i32_16x16x64_i8(af, BC[b], acc[a][b], 0, 0, 0);                                   \
            if (a == 3 && more_) {                                                                                     \
                int64_t const after_ = nst - 2 - (ST); /* stages issued beyond ST+1 */                                 \
                cg_wait(int(after_ < 2 ? after_ : 2));                                                                 \
                cg_load_b(buf((ST) + 1), BN, wn, r, q);                                                                \
            }                                                                                                          \
        }                                                                                                              \
        if ((ST) + CG_NBUF < nst) {                                                                                    \
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");                                                          \
            __builtin_amdgcn_s_barrier();                                                                              \
            dma((ST) + CG_NBUF);                                                                                       \
        }                                                                                                              \
    } while (0)

    for (int64_t st = 0; st < nst; st += 2) {
        CG_STAGE(b0, b1, st);
        if (st + 1 < nst) CG_STAGE(b1, b0, st + 1);
    }
#undef CG_STAGE

    // D[4*(lane/16)+v][lane%16] of block (a, b): i = ... + 4q + v, j = ... + r
    uint32_t const m     = uint32_t(P.K.mod[l]);
    uint32_t const magic = P.K.magic[l];
    uint32_t const bias  = P.K.bias[l];
    uint8_t *__restrict__ Rl = P.R + int64_t(l) * P.r_plane;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        int64_t const j = j0 + wn * 64 + b * 16 + r;
        if (j >= P.n) continue;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            int64_t const i    = i0 + wm * 128 + a * 16 + 4 * q; // < ldr (padded to whole tiles)
            uint32_t *dst      = reinterpret_cast<uint32_t *>(Rl + j * P.ldr + i);
            uint32_t const old = P.accumulate ? *dst : 0u;
            uint32_t word      = 0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                uint32_t x = cg_mod(acc[a][b][v], m, magic, bias) + ((old >> (8 * v)) & 0xffu);
                x          = x >= m ? x - m : x;
                word |= x << (8 * v);
            }
            *dst = word;
        }
    }
}

// ------------------------------------------------------------ reconstruction
// C(i,j) = alpha 2^(e_i + f_j - 2kb) C'(i,j) + beta C(i,j) for the panel's columns; each
// thread takes four consecutive i of one column (one 32-bit word of every residue plane).
struct CrtRec {
    int64_t m, n;                     // rows, panel columns
    uint8_t const *R;                 // [l][j][i]
    int64_t ldr, r_plane;
    int32_t const *ea, *eb;           // eb already offset to the panel
    double *C;                        // offset to the panel
    int64_t c_sm, c_sn;
    double alpha, beta;
    CrtBasis W;
};

// a + b = s + e exactly (Knuth's two-sum)
__device__ __forceinline__ double two_sum(double a, double b, double &e) {
    double const s  = a + b;
    double const bb = s - a;
    e               = (a - (s - bb)) + (b - bb);
    return s;
}

__global__ __launch_bounds__(256) void crt_reconstruct_kernel(CrtRec P) {
    int64_t const words = P.ldr / 4;
    int64_t const g     = int64_t(blockIdx.x) * 256 + threadIdx.x;
    int64_t const j     = g / words;
    int64_t const i     = (g - j * words) * 4;
    if (j >= P.n || i >= P.m) return;
    double s[4][4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int t = 0; t < 4; ++t) s[v][t] = 0.0;
    uint8_t const *__restrict__ src = P.R + j * P.ldr + i;
#pragma unroll 4
    for (int l = 0; l < P.W.nmod; ++l) {
        uint32_t const word = *reinterpret_cast<uint32_t const *>(src + int64_t(l) * P.r_plane);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double const rv = double((word >> (8 * v)) & 0xffu);
#pragma unroll
            for (int t = 0; t < 4; ++t) s[v][t] = fma(rv, P.W.w[l][t], s[v][t]); // exact: < 2^44
        }
    }
    int const eb = P.eb[j];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        int64_t const ii = i + v;
        if (ii >= P.m) break;
        double const approx = fma(s[v][3], 0x1p96, fma(s[v][2], 0x1p64, fma(s[v][1], 0x1p32, s[v][0])));
        double const qm     = rint(approx * P.W.inv_m); // S / M rounded: |C'| <= M/4 leaves margin
        double lim[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) lim[t] = fma(-qm, P.W.mlimb[t], s[v][t]); // exact limbs of S - qM
        double e0, e1, e2;
        double h = two_sum(lim[1] * 0x1p32, lim[0], e0);
        h        = two_sum(lim[2] * 0x1p64, h, e1);
        h        = two_sum(lim[3] * 0x1p96, h, e2);
        double const cp = h + ((e0 + e1) + e2);
        double *c       = P.C + ii * P.c_sm + j * P.c_sn;
        double val      = P.alpha * ldexp(cp, P.ea[ii] + eb - P.W.shift);
        if (P.beta != 0.0) val = fma(P.beta, *c, val);
        *c = val;
    }
}

// ------------------------------------------------------------------- host
typedef unsigned __int128 u128;

int crt_kbits(int nmod, int64_t k) {
    u128 M = 1;
    for (int l = 0; l < nmod; ++l) M *= u128(CRT_MODS[l]);
    int kb = 0;
    // largest kb with 4 k 2^(2 kb) <= M, at most 53 (A' must stay an exact FP64 integer)
    while (kb < 53) {
        int const sh = 2 * (kb + 1) + 2;
        u128 const need = u128(uint64_t(k)) << sh;
        if (sh >= 126 || (need >> sh) != u128(uint64_t(k)) || need > M) break;
        ++kb;
    }
    return kb;
}

void crt_constants(int nmod, int kbits, CrtMods &K, CrtBasis &W) {
    K       = CrtMods{};
    W       = CrtBasis{};
    K.nmod  = nmod;
    K.kbits = kbits;
    W.nmod  = nmod;
    W.shift = 2 * kbits;
    u128 M  = 1;
    for (int l = 0; l < nmod; ++l) M *= u128(CRT_MODS[l]);
    auto limbs = [](u128 x, double *out) {
        for (int t = 0; t < 4; ++t) out[t] = double(uint64_t(x >> (32 * t)) & 0xffffffffull);
    };
    for (int l = 0; l < nmod; ++l) {
        uint32_t const m = uint32_t(CRT_MODS[l]);
        K.mod[l]         = int32_t(m);
        K.magic[l]       = uint32_t((uint64_t(1) << 32) / m);
        K.bias[l]        = uint32_t(((uint64_t(1) << 30) + m - 1) / m * m);
        K.rinv[l]        = 1.0 / double(m);
        u128 const Mi    = M / m;
        // inverse of Mi mod m by search (m <= 256)
        uint32_t const mi = uint32_t(Mi % m);
        uint32_t inv      = 0;
        for (uint32_t x = 1; x < m; ++x)
            if ((mi * x) % m == 1) {
                inv = x;
                break;
            }
        limbs(Mi * inv % M, W.w[l]); // W_l = Mi * (Mi^-1 mod m) < M
    }
    limbs(M, W.mlimb);
    W.inv_m = 1.0 / (double(uint64_t(M >> 64)) * 0x1p64 + double(uint64_t(M)));
}

int crt_residues(SliceSrc const &S, CrtMods const &K, int32_t *e, int8_t *out, double *scratch, hipStream_t stream) {
    bool const rows_fast = S.s_row == 1 || (S.s_k != 1 && std::llabs(S.s_row) < std::llabs(S.s_k));
    EIN_TRY(row_exponents(S, e, false, scratch, stream));
    dim3 grid{uint32_t(ceil_div(S.kp, 64)), uint32_t(ceil_div(S.rows, 64))};
    if (grid.y > 65535) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: more than 4M rows");
    if (rows_fast)
        crt_residue_kernel<true><<<grid, 256, 0, stream>>>(S, e, K, out);
    else
        crt_residue_kernel<false><<<grid, 256, 0, stream>>>(S, e, K, out);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

// The side stream of the current device (created once, default priority) and four events
// of this host thread on it.
struct CrtSide {
    hipStream_t aux = nullptr;
    hipEvent_t gemm_done[2]  = {nullptr, nullptr};
    hipEvent_t recon_done[2] = {nullptr, nullptr};
};

int crt_side(CrtSide &out) {
    int dev = 0;
    EIN_HIP_CHECK(hipGetDevice(&dev));
    {
        static std::mutex mu;
        static std::map<int, hipStream_t> aux;
        std::lock_guard<std::mutex> lock(mu);
        auto it = aux.find(dev);
        if (it == aux.end()) {
            hipStream_t s = nullptr;
            EIN_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            it = aux.emplace(dev, s).first;
        }
        out.aux = it->second;
    }
    thread_local std::map<int, std::array<hipEvent_t, 4>> events;
    auto it = events.find(dev);
    if (it == events.end()) {
        std::array<hipEvent_t, 4> e{};
        for (auto &x : e) EIN_HIP_CHECK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
        it = events.emplace(dev, e).first;
    }
    out.gemm_done[0]  = it->second[0];
    out.gemm_done[1]  = it->second[1];
    out.recon_done[0] = it->second[2];
    out.recon_done[1] = it->second[3];
    return EIN_SUCCESS;
}

} // namespace

// Residue planes are written for column panels of at most this many bytes.
constexpr size_t CRT_PANEL_BYTES = size_t(8) << 30;

int gemm_split_crt(GemmProblem<double> P, int nmod, hipStream_t stream) {
    if (nmod < CRT_MIN_MODULI || nmod > CRT_MAX_MODS)
        return fail(EIN_ERROR_VALUE, "split gemm: the CRT scheme takes " + std::to_string(CRT_MIN_MODULI) + ".." +
                                         std::to_string(CRT_MAX_MODS) + " moduli");
    int64_t const kp     = ceil_div(P.k, CG_KS) * CG_KS;
    int const kbits      = crt_kbits(nmod, P.k);
    if (kbits < 8) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: contraction too long for the CRT scheme");
    CrtMods K;
    CrtBasis W;
    crt_constants(nmod, kbits, K, W);

    uint32_t const tiles_m = uint32_t(ceil_div(P.m, CG_TILE));
    int64_t const ldr      = int64_t(tiles_m) * CG_TILE;
    // column panels of whole 256-column tiles whose residue planes take at most
    // CRT_PANEL_BYTES; with more than one panel, two plane buffers of half that so the
    // reconstruction of panel p (HBM-bound) runs on a side stream while the residue GEMM of
    // panel p+1 (MFMA-bound) runs on the caller's stream
    auto panel_cols = [&](size_t budget) {
        int64_t c = int64_t(budget / (size_t(nmod) * size_t(ldr))) / CG_TILE * CG_TILE;
        return std::max<int64_t>(CG_TILE, std::min<int64_t>(c, ceil_div(P.n, CG_TILE) * CG_TILE));
    };
    int64_t pn      = panel_cols(CRT_PANEL_BYTES);
    bool const fork = pn < P.n && !stream_capturing(stream);
    if (fork) pn = panel_cols(CRT_PANEL_BYTES / 2);
    int const nbuf = fork ? 2 : 1;

    size_t const a_bytes = size_t(nmod) * size_t(P.m) * size_t(kp);
    size_t const b_bytes = size_t(nmod) * size_t(P.n) * size_t(kp);
    size_t const r_bytes = size_t(nmod) * size_t(pn) * size_t(ldr);
    auto up              = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t const b_off   = up(a_bytes);
    size_t const r_off   = b_off + up(b_bytes);
    size_t const ea_off  = r_off + size_t(nbuf) * up(r_bytes);
    size_t const eb_off  = ea_off + up(size_t(P.m) * 4);
    size_t const x_off   = eb_off + up(size_t(P.n) * 4);
    size_t const total   = x_off + row_exp_scratch(std::max(P.m, P.n)) * sizeof(double);
    void *ws             = nullptr;
    EIN_TRY(workspace(stream, total, &ws));
    char *base      = static_cast<char *>(ws);
    int8_t *As      = reinterpret_cast<int8_t *>(base);
    int8_t *Bs      = reinterpret_cast<int8_t *>(base + b_off);
    uint8_t *R      = reinterpret_cast<uint8_t *>(base + r_off);
    int32_t *ea     = reinterpret_cast<int32_t *>(base + ea_off);
    int32_t *eb     = reinterpret_cast<int32_t *>(base + eb_off);
    double *scratch = reinterpret_cast<double *>(base + x_off);

    SliceSrc SA{P.m, P.k, kp, P.A, P.a_sm, P.a_sk};
    SliceSrc SB{P.n, P.k, kp, P.B, P.b_sn, P.b_sk};
    EIN_TRY(crt_residues(SA, K, ea, As, scratch, stream));
    EIN_TRY(crt_residues(SB, K, eb, Bs, scratch, stream));

    CrtSide side;
    if (fork) EIN_TRY(crt_side(side));
    int last    = -1; // plane buffer of the last reconstruction queued on the side stream
    auto panels = [&]() -> int {
        for (int64_t j0 = 0, panel = 0; j0 < P.n; j0 += pn, ++panel) {
            int64_t const nj = std::min(pn, P.n - j0);
            int const b      = fork ? int(panel & 1) : 0;
            uint8_t *Rb      = R + size_t(b) * up(r_bytes);
            if (fork && panel >= 2) EIN_HIP_CHECK(hipStreamWaitEvent(stream, side.recon_done[b], 0)); // buffer b free
            CrtGemm G{};
            G.m       = P.m;
            G.n       = nj;
            G.kp      = kp;
            G.a_plane = P.m * kp;
            G.b_plane = P.n * kp;
            G.A       = As;
            G.B       = Bs + j0 * kp;
            G.R       = Rb;
            G.ldr     = ldr;
            G.r_plane = nj * ldr;
            G.tiles_m = tiles_m;
            G.tiles_n = uint32_t(ceil_div(nj, CG_TILE));
            G.K       = K;
            uint64_t const grid = uint64_t(G.tiles_m) * G.tiles_n * uint64_t(nmod);
            if (grid > 0x7fffffffu) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: tile grid too large");
            for (int64_t k0 = 0; k0 < kp; k0 += int64_t(CRT_FOLD_STAGES) * CG_KS) {
                G.k0         = k0;
                G.kc         = std::min<int64_t>(int64_t(CRT_FOLD_STAGES) * CG_KS, kp - k0);
                G.accumulate = k0 > 0;
                crt_gemm_kernel<<<uint32_t(grid), 64 * CG_WAVES, 0, stream>>>(G);
                EIN_LAUNCH_CHECK();
            }

            hipStream_t rs = stream;
            if (fork) {
                EIN_HIP_CHECK(hipEventRecord(side.gemm_done[b], stream));
                EIN_HIP_CHECK(hipStreamWaitEvent(side.aux, side.gemm_done[b], 0));
                rs = side.aux;
            }
            CrtRec Q{};
            Q.m       = P.m;
            Q.n       = nj;
            Q.R       = Rb;
            Q.ldr     = ldr;
            Q.r_plane = nj * ldr;
            Q.ea      = ea;
            Q.eb      = eb + j0;
            Q.C       = P.C + j0 * P.c_sn;
            Q.c_sm    = P.c_sm;
            Q.c_sn    = P.c_sn;
            Q.alpha   = P.alpha;
            Q.beta    = P.beta;
            Q.W       = W;
            int64_t const threads = nj * (ldr / 4);
            if (ceil_div(threads, 256) > 0x7fffffff) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: panel too large");
            crt_reconstruct_kernel<<<uint32_t(ceil_div(threads, 256)), 256, 0, rs>>>(Q);
            if (fork) {
                last = b;
                EIN_HIP_CHECK(hipEventRecord(side.recon_done[b], side.aux));
            }
            EIN_LAUNCH_CHECK();
        }
        return EIN_SUCCESS;
    };
    int const status = panels();
    // join, also after a failure, so no queued reconstruction outlives the call's workspace:
    // the side stream runs the reconstructions in order, so the last one covers all
    if (last >= 0) EIN_HIP_CHECK(hipStreamWaitEvent(stream, side.recon_done[last], 0));
    return status;
}

int crt_bits(int nmod, int64_t k) {
    if (nmod < CRT_MIN_MODULI || nmod > CRT_MAX_MODS || k < 0) return -1;
    return crt_kbits(nmod, std::max<int64_t>(k, 1));
}

int crt_auto_moduli(int64_t k) {
    for (int L = CRT_MIN_MODULI; L < CRT_MAX_MODS; ++L)
        if (crt_kbits(L, std::max<int64_t>(k, 1)) >= CRT_AUTO_BITS) return L;
    return CRT_MAX_MODS;
}

} // namespace ein

extern "C" {

int ein_split_crt_bits(int32_t nmoduli, int64_t k) {
    if (nmoduli == 0 && k >= 0) nmoduli = ein::crt_auto_moduli(k);
    return ein::crt_bits(nmoduli, k);
}

int ein_split_crt_moduli(int64_t k) { return k < 0 ? -1 : ein::crt_auto_moduli(k); }

} // extern "C"
