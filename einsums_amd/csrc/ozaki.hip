// Split-precision FP64 GEMM for gfx950 (Ozaki scheme on BF16/FP16 MFMA).
//
// Each row i of op(A) is scaled by 2^-e_i (|A(i,:)| < 2^e_i) and cut into NS integer
// slices of BETA = 8 bits:  A(i,p) = 2^e_i * sum_t a_t(i,p) 2^(-8(t+1)),  |a_0| <= 256,
// |a_t| <= 128 for t > 0 -- integers that BF16 and FP16 hold exactly. B is cut per
// column j the same way (exponent f_j). Then
//     (A B)(i,j) = 2^(e_i+f_j) sum_{t,u} 2^(-8(t+u+2)) sum_p a_t(i,p) b_u(p,j)
// and the pairs with t+u >= NS are dropped (their weight is below 2^(-8(NS+1))).
// Every inner sum is an exact integer while it stays below 2^24, so the MFMA FP32
// accumulators for one diagonal d = t+u (at most d+1 pairs, each product <= 2^16 for d=0,
// 2^15 / 2^14 otherwise) are exact over chunks of KC = 128 contraction steps:
// (d+3) * 2^14 * 128 < 2^24 for d <= 4, and 2^16 * 128 < 2^24 for d = 0. After every chunk
// the diagonals are converted to FP64, scaled by 2^(-8(d+2)) and added to FP64
// accumulators; the epilogue applies 2^(e_i+f_j), alpha and beta in FP64.
// With NS = 5 the representation error is 2^-40 relative to the row/column maximum.
#include <algorithm>
#include <cmath>
#include <tuple>

#include "gemm.hpp"
#include "ozaki.hpp"
#include "tile_engine.hpp"

namespace ein {

namespace {

constexpr int OZ_BETA = 8;
constexpr int OZ_KC   = 128; // exact-accumulation chunk (multiple of 32)
constexpr int OZ_LDS_ROW = 40; // 32 used + 8 padding elements per LDS row (80 bytes)

template <typename H>
struct HalfMfma;
template <>
struct HalfMfma<__bf16> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ v4f mma(v8 a, v8 b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct HalfMfma<_Float16> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ v4f mma(v8 a, v8 b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// ----------------------------------------------------------- row exponents
// e[r] = exponent with max_p |X(r,p)| < 2^e (0 for an all-zero row).
// With `guard`, the exponent is raised by one when the scaled maximum would exceed 127/128,
// so that a 7-bit slice rounded to nearest stays within int8 (|slice| <= 127).
__device__ __forceinline__ int row_exponent(double mx, bool guard) {
    int ex = 0;
    if (mx > 0.0) {
        double const f = frexp(mx, &ex);
        if (guard && f > 127.0 / 128.0) ++ex;
    }
    return ex;
}


// Rows-fast operands: the maximum of each row over chunk `blockIdx.y` of kch contraction
// steps (64 rows x 4 k-lanes per workgroup), then a second pass over the chunks. A single
// thread per row walking all of k would leave the GPU idle for long k.
__global__ __launch_bounds__(256) void row_max_partial(SliceSrc S, int64_t kch, double *__restrict__ part) {
    __shared__ double red[4][64];
    int const t      = threadIdx.x;
    int64_t const r  = int64_t(blockIdx.x) * 64 + (t & 63);
    int64_t const p0 = int64_t(blockIdx.y) * kch;
    int64_t const p1 = p0 + kch < S.k ? p0 + kch : S.k;
    double mx        = 0.0;
    if (r < S.rows)
        for (int64_t p = p0 + (t >> 6); p < p1; p += 4) mx = fmax(mx, fabs(S.X[r * S.s_row + p * S.s_k]));
    red[t >> 6][t & 63] = mx;
    __syncthreads();
    if (t < 64 && r < S.rows)
        part[int64_t(blockIdx.y) * S.rows + r] = fmax(fmax(red[0][t], red[1][t]), fmax(red[2][t], red[3][t]));
}

__global__ __launch_bounds__(256) void row_exp_finish(SliceSrc S, double const *__restrict__ part, int nck,
                                                      int32_t *__restrict__ e, bool guard) {
    int64_t const r = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r >= S.rows) return;
    double mx = 0.0;
    for (int c = 0; c < nck; ++c) mx = fmax(mx, part[int64_t(c) * S.rows + r]);
    e[r] = row_exponent(mx, guard);
}

__global__ __launch_bounds__(256) void row_exp_k_fast(SliceSrc S, int32_t *__restrict__ e, bool guard) {
    int const lane = threadIdx.x & 63;
    int64_t r      = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (r >= S.rows) return;
    double mx = 0.0;
    for (int64_t p = lane; p < S.k; p += 64) mx = fmax(mx, fabs(S.X[r * S.s_row + p * S.s_k]));
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if (lane == 0) e[r] = row_exponent(mx, guard);
}

// ----------------------------------------------------------------- slicing
// 64 x 64 tiles of X staged through LDS so that the read follows X's unit-stride axis
// and the slice writes run along p (slice layout: out[t][r][kp], p fastest).
template <typename H, bool ROWS_FAST>
__global__ __launch_bounds__(256) void slice_kernel(SliceSrc S, int32_t const *__restrict__ e, int ns,
                                                    H *__restrict__ out) {
    __shared__ double tile[64][65];
    int64_t const r0 = int64_t(blockIdx.y) * 64;
    int64_t const p0 = int64_t(blockIdx.x) * 64;
    int const t      = threadIdx.x;
    if constexpr (ROWS_FAST) {
        int const lr = t & 63;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            int lp     = (t >> 6) + 4 * q;
            int64_t r  = r0 + lr, p = p0 + lp;
            tile[lr][lp] = (r < S.rows && p < S.k) ? S.X[r * S.s_row + p * S.s_k] : 0.0;
        }
    } else {
        int const lp = t & 63;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            int lr     = (t >> 6) + 4 * q;
            int64_t r  = r0 + lr, p = p0 + lp;
            tile[lr][lp] = (r < S.rows && p < S.k) ? S.X[r * S.s_row + p * S.s_k] : 0.0;
        }
    }
    __syncthreads();
    int const lp = t & 63;
    int64_t const p = p0 + lp;
    if (p >= S.kp) return;
#pragma unroll 1
    for (int q = 0; q < 16; ++q) {
        int lr    = (t >> 6) + 4 * q;
        int64_t r = r0 + lr;
        if (r >= S.rows) continue;
        double x = ldexp(tile[lr][lp], -e[r]);
        for (int s = 0; s < ns; ++s) {
            double y = ldexp(x, OZ_BETA);
            double d = rint(y);
            x        = y - d;
            out[(int64_t(s) * S.rows + r) * S.kp + p] = static_cast<H>(static_cast<float>(d));
        }
    }
}

// ------------------------------------------------------------- slice GEMM
// C(i,j) = alpha * 2^(e_i+f_j) * sum_d 2^(-8(d+2)) D_d(i,j) + beta * C(i,j);
// workgroup tile BM x BN, waves of 32 x 32 (2 x 2 MFMA blocks of 16 x 16 x 32).
template <typename H, int BM, int BN, int NS>
__global__ __launch_bounds__((BM / 32) * (BN / 32) * 64) void ozaki_gemm_kernel(OzProblem P) {
    using M8                 = typename HalfMfma<H>::v8;
    constexpr int WAVES_M    = BM / 32;
    constexpr int NT         = WAVES_M * (BN / 32) * 64;
    constexpr int A_CHUNKS   = NS * BM * 4; // 16-byte chunks per stage
    constexpr int B_CHUNKS   = NS * BN * 4;
    constexpr int A_PER_T    = (A_CHUNKS + NT - 1) / NT;
    constexpr int B_PER_T    = (B_CHUNKS + NT - 1) / NT;
    constexpr int A_STAGE    = NS * BM * OZ_LDS_ROW;
    constexpr int B_STAGE    = NS * BN * OZ_LDS_ROW;
    __shared__ __attribute__((aligned(16))) H lds[2 * (A_STAGE + B_STAGE)];

    int const tid  = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const wm   = wave % WAVES_M;
    int const wn   = wave / WAVES_M;
    int const r    = lane & 15;
    int const q    = lane >> 4;

    uint32_t const ntiles  = P.tiles_m * P.tiles_n;
    uint32_t const logical = xcd_remap(blockIdx.x, ntiles);
    uint32_t tm, tn;
    tile_coords(logical, P.tiles_m, P.tiles_n, tm, tn);
    int64_t const i0 = int64_t(tm) * BM;
    int64_t const j0 = int64_t(tn) * BN;

    H const *__restrict__ Ag = static_cast<H const *>(P.As_raw);
    H const *__restrict__ Bg = static_cast<H const *>(P.Bs_raw);
    uint4 ra[A_PER_T], rb[B_PER_T];
    auto load_stage = [&](int64_t k0) {
#pragma unroll
        for (int c = 0; c < A_PER_T; ++c) {
            int ch = tid + c * NT;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (ch < A_CHUNKS) {
                int s = ch / (BM * 4), rr = (ch / 4) % BM, kc = ch % 4;
                int64_t i = i0 + rr;
                if (i < P.m) v = *reinterpret_cast<uint4 const *>(Ag + (int64_t(s) * P.m + i) * P.kp + k0 + kc * 8);
            }
            ra[c] = v;
        }
#pragma unroll
        for (int c = 0; c < B_PER_T; ++c) {
            int ch = tid + c * NT;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (ch < B_CHUNKS) {
                int s = ch / (BN * 4), rr = (ch / 4) % BN, kc = ch % 4;
                int64_t j = j0 + rr;
                if (j < P.n) v = *reinterpret_cast<uint4 const *>(Bg + (int64_t(s) * P.n + j) * P.kp + k0 + kc * 8);
            }
            rb[c] = v;
        }
    };
    auto store_stage = [&](int buf) {
        H *As = lds + buf * (A_STAGE + B_STAGE);
        H *Bs = As + A_STAGE;
#pragma unroll
        for (int c = 0; c < A_PER_T; ++c) {
            int ch = tid + c * NT;
            if (ch < A_CHUNKS) {
                int s = ch / (BM * 4), rr = (ch / 4) % BM, kc = ch % 4;
                *reinterpret_cast<uint4 *>(As + (s * BM + rr) * OZ_LDS_ROW + kc * 8) = ra[c];
            }
        }
#pragma unroll
        for (int c = 0; c < B_PER_T; ++c) {
            int ch = tid + c * NT;
            if (ch < B_CHUNKS) {
                int s = ch / (BN * 4), rr = (ch / 4) % BN, kc = ch % 4;
                *reinterpret_cast<uint4 *>(Bs + (s * BN + rr) * OZ_LDS_ROW + kc * 8) = rb[c];
            }
        }
    };

    v4f accf[NS][2][2];
    v4d accd[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
#pragma unroll
            for (int v = 0; v < 4; ++v) accd[a][b][v] = 0.0;
#pragma unroll
            for (int d = 0; d < NS; ++d)
#pragma unroll
                for (int v = 0; v < 4; ++v) accf[d][a][b][v] = 0.f;
        }

    auto flush = [&]() {
#pragma unroll
        for (int d = 0; d < NS; ++d) {
            double const sc = ldexp(1.0, -OZ_BETA * (d + 2));
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        accd[a][b][v] += double(accf[d][a][b][v]) * sc;
                        accf[d][a][b][v] = 0.f;
                    }
        }
    };

    int64_t const nstages = P.kp / 32;
    if (nstages > 0) {
        load_stage(0);
        store_stage(0);
        __syncthreads();
    }
    for (int64_t st = 0; st < nstages; ++st) {
        int const buf = int(st & 1);
        if (st + 1 < nstages) load_stage((st + 1) * 32);
        H const *As = lds + buf * (A_STAGE + B_STAGE);
        H const *Bs = As + A_STAGE;
        M8 af[NS][2], bfr[NS][2];
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                af[s][a]  = *reinterpret_cast<M8 const *>(As + (s * BM + wm * 32 + a * 16 + r) * OZ_LDS_ROW + q * 8);
                bfr[s][a] = *reinterpret_cast<M8 const *>(Bs + (s * BN + wn * 32 + a * 16 + r) * OZ_LDS_ROW + q * 8);
            }
#pragma unroll
        for (int t = 0; t < NS; ++t)
#pragma unroll
            for (int u = 0; u < NS - t; ++u)
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) accf[t + u][a][b] = HalfMfma<H>::mma(bfr[u][b], af[t][a], accf[t + u][a][b]);
        if (((st + 1) * 32) % OZ_KC == 0 || st + 1 == nstages) flush();
        if (st + 1 < nstages) store_stage(buf ^ 1);
        __syncthreads();
    }

    // epilogue: FP32-MFMA result layout D[4*(lane/16)+v][lane%16]; rows of D run along j
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        int64_t i = i0 + wm * 32 + a * 16 + r;
        if (i >= P.m) continue;
        int const ei = P.ea[i];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int64_t j = j0 + wn * 32 + b * 16 + q * 4 + v;
                if (j >= P.n) continue;
                double *cp = P.C + i * P.c_sm + j * P.c_sn;
                double val = P.alpha * ldexp(accd[a][b][v], ei + P.eb[j]);
                if (P.beta != 0.0) val += P.beta * *cp;
                *cp = val;
            }
    }
}

// ------------------------------------------------------------ int8 slices
// Integer variant on v_mfma_i32_16x16x64_i8. The row exponent is chosen so that
// |A(i,:)| * 2^-e_i <= 127/128; slices carry 7 bits each, rounded to nearest:
//     a_0 = rint(128 x) in [-127, 127], then a_t = rint(128 r_t) in [-64, 64], r in [-1/2, 1/2],
//     A(i,p) = 2^e_i * sum_t a_t(i,p) 2^(-7(t+1)) + r,  |r| <= 2^(e_i - 7 NS - 1).
// A diagonal d = t+u holds at most one product of two leading slices (<= 127^2, d = 0 only)
// and otherwise products <= 127*64, so its int32 accumulator is exact while
// (d+1) * 127 * 64 * K < 2^31: over I8_KC = 32768 contraction steps for NS <= 6. Longer
// contractions are split (split-K) and the partial results summed in FP64.
constexpr int I8_BETA = 7;
constexpr int I8_KC   = 32768;
constexpr int I8_KS   = 64; // contraction steps per stage (one MFMA depth)

typedef int v4i __attribute__((ext_vector_type(4)));

// 64 x 64 tiles of X staged through LDS; slices written along p (out[t][r][kp]).
template <bool ROWS_FAST>
__global__ __launch_bounds__(256) void slice_i8_kernel(SliceSrc S, int32_t const *__restrict__ e, int ns,
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
    // four consecutive p of four rows per thread: one 32-bit store per row and slice
    int const p4    = (t & 15) * 4;
    int64_t const p = p0 + p4; // kp is a multiple of 64: p .. p+3 < kp whenever p < kp
    if (p >= S.kp) return;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        int const lr    = (t >> 4) + 16 * q;
        int64_t const r = r0 + lr;
        if (r >= S.rows) continue;
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = ldexp(tile[lr][p4 + u], -e[r]);
        int8_t *__restrict__ dst = out + r * S.kp + p;
        for (int s = 0; s < ns; ++s) {
            uint32_t word = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double const y = ldexp(x[u], I8_BETA);
                double const d = rint(y);
                x[u]           = y - d;
                word |= uint32_t(uint8_t(static_cast<int8_t>(d))) << (8 * u);
            }
            *reinterpret_cast<uint32_t *>(dst + int64_t(s) * S.rows * S.kp) = word;
        }
    }
}

// LDS stage of one operand: NS slices x R rows x 64 int8 as 16-byte chunks; chunk c of
// (slice s, row) sits at slot (s*R + row)*4 + (c ^ ((row >> 2) & 3)), so the 16 rows of
// one fragment read fall on 16 distinct 16-byte bank groups.
template <int R, int NS>
struct I8Stage {
    static constexpr int SLOTS = NS * R * 4;
    static constexpr int INS   = SLOTS / 64;
    static_assert(INS % 4 == 0, "stage instructions must split over four waves");
    static __device__ __forceinline__ int slot(int s, int row, int c) { return (s * R + row) * 4 + (c ^ ((row >> 2) & 3)); }
};

// Stage DMA: lane l of instruction n fills slot 64n + l.
template <int R, int NS>
__device__ __forceinline__ void i8_dma(int8_t const *__restrict__ X, int64_t rows_total, int64_t kp, int64_t row0,
                                       int64_t k0, int8_t *lds, int wave, int lane) {
    using St = I8Stage<R, NS>;
#pragma unroll
    for (int x = 0; x < St::INS / 4; ++x) {
        int const n    = wave + 4 * x;
        int const sl   = n * 64 + lane;
        int const srow = sl / 4; // s*R + row
        int const s    = srow / R;
        int const row  = srow % R;
        int const c    = (sl % 4) ^ ((row >> 2) & 3);
        int64_t const g = row0 + row;
        void const *src = g < rows_total ? static_cast<void const *>(X + (int64_t(s) * rows_total + g) * kp + k0 + 16 * c)
                                         : static_cast<void const *>(ein_zero_page);
        __builtin_amdgcn_global_load_lds(src, EIN_LDS_PTR(lds + n * 64 * 16), 16, 0, 0);
    }
}

// MFMA work of one contraction stage from the LDS stage `cur` (A slots, then B slots);
// `dma(nxt)` fills the other stage buffer after the B fragments are requested. cur and nxt
// are restrict parameters so that the compiler does not wait for that DMA before the A
// fragment reads of this stage.
template <int NS, typename Dma>
__device__ __forceinline__ void i8_stage(int8_t const *__restrict__ cur, int8_t *__restrict__ nxt, v4i (&acc)[NS][4][2],
                                         int wm, int wn, int r, int q, Dma const &dma) {
    constexpr int FM = 4, FN = 2;
    using SA = I8Stage<128, NS>;
    using SB = I8Stage<64, NS>;
    v4i const *As = reinterpret_cast<v4i const *>(cur);
    v4i const *Bs = As + SA::SLOTS;
    v4i bfr[NS][FN];
#pragma unroll
    for (int u = 0; u < NS; ++u)
#pragma unroll
        for (int b = 0; b < FN; ++b) bfr[u][b] = Bs[SB::slot(u, wn * 32 + b * 16 + r, q)];
    dma(nxt);
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        v4i af[FM];
#pragma unroll
        for (int a = 0; a < FM; ++a) af[a] = As[SA::slot(t, wm * 64 + a * 16 + r, q)];
#pragma unroll
        for (int u = 0; u < NS - t; ++u)
#pragma unroll
            for (int a = 0; a < FM; ++a)
#pragma unroll
                for (int b = 0; b < FN; ++b)
                    acc[t + u][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bfr[u][b], af[a], acc[t + u][a][b], 0, 0, 0);
    }
}

// C(i,j) = alpha * 2^(e_i+f_j) * sum_d 2^(-7(d+2)) D_d(i,j) + beta * C(i,j); workgroup tile
// 128 x 64, waves of 64 x 32 (4 x 2 blocks of 16 x 16 x 64).
// Split-K: workgroup (tile, split) accumulates contraction steps [split*kchunk, ...) with
// kchunk <= I8_KC, so every int32 diagonal stays exact without intermediate folding.
template <int NS>
__global__ __launch_bounds__(256, 1) void ozaki_i8_kernel(OzProblem P) {
    constexpr int BM = 128, BN = 64, FM = 4, FN = 2;
    using SA = I8Stage<BM, NS>;
    using SB = I8Stage<BN, NS>;
    __shared__ __attribute__((aligned(16))) int8_t lds[2 * (SA::SLOTS + SB::SLOTS) * 16];

    int const tid  = threadIdx.x;
    int const lane = tid & 63;
    int const wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const wm   = wave & 1;
    int const wn   = wave >> 1;
    int const r    = lane & 15;
    int const q    = lane >> 4;

    uint32_t const ntiles  = P.tiles_m * P.tiles_n;
    uint32_t const logical = xcd_remap(blockIdx.x, gridDim.x);
    uint32_t const split   = logical / ntiles;
    uint32_t const tile    = logical - split * ntiles;
    uint32_t tm, tn;
    tile_coords(tile, P.tiles_m, P.tiles_n, tm, tn);
    int64_t const i0 = int64_t(tm) * BM;
    int64_t const j0 = int64_t(tn) * BN;
    int64_t const kb = int64_t(split) * P.kchunk;
    int64_t const ke = kb + P.kchunk < P.kp ? kb + P.kchunk : P.kp;
    int8_t const *__restrict__ Ag = static_cast<int8_t const *>(P.As_raw);
    int8_t const *__restrict__ Bg = static_cast<int8_t const *>(P.Bs_raw);

    v4i acc[NS][FM][FN];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int d = 0; d < NS; ++d) acc[d][a][b] = v4i{0, 0, 0, 0};

    int64_t const nst = (ke - kb) / I8_KS;
    auto stage_ptr    = [&](int buf) { return lds + buf * (SA::SLOTS + SB::SLOTS) * 16; };
    if (nst > 0) {
        i8_dma<BM, NS>(Ag, P.m, P.kp, i0, kb, stage_ptr(0), wave, lane);
        i8_dma<BN, NS>(Bg, P.n, P.kp, j0, kb, stage_ptr(0) + SA::SLOTS * 16, wave, lane);
    }
    for (int64_t st = 0; st < nst; ++st) {
        int const buf = int(st & 1);
        dma_wait_all();
        auto dma = [&](int8_t *nx) {
            if (st + 1 < nst) {
                i8_dma<BM, NS>(Ag, P.m, P.kp, i0, kb + (st + 1) * I8_KS, nx, wave, lane);
                i8_dma<BN, NS>(Bg, P.n, P.kp, j0, kb + (st + 1) * I8_KS, nx + SA::SLOTS * 16, wave, lane);
            }
        };
        i8_stage<NS>(stage_ptr(buf), stage_ptr(buf ^ 1), acc, wm, wn, r, q, dma);
    }

    // weighted sum of the diagonals of block (a, b), element v
    auto diag_sum = [&](int a, int b, int v) __attribute__((always_inline)) {
        double x = 0.0;
#pragma unroll
        for (int d = NS - 1; d >= 0; --d) x += double(acc[d][a][b][v]) * ldexp(1.0, -I8_BETA * (d + 2));
        return x;
    };

    // result layout D[4*(lane/16)+v][lane%16]; rows of D run along j
    if (P.splits > 1) {
        double *__restrict__ part = P.partial + (int64_t(split) * ntiles + tile) * (BM * BN);
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
            for (int b = 0; b < FN; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    part[(wn * 32 + b * 16 + q * 4 + v) * BM + wm * 64 + a * 16 + r] = diag_sum(a, b, v);
        return;
    }
#pragma unroll
    for (int a = 0; a < FM; ++a) {
        int64_t i = i0 + wm * 64 + a * 16 + r;
        if (i >= P.m) continue;
        int const ei = P.ea[i];
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int64_t j = j0 + wn * 32 + b * 16 + q * 4 + v;
                if (j >= P.n) continue;
                double *cp = P.C + i * P.c_sm + j * P.c_sn;
                double val = P.alpha * ldexp(diag_sum(a, b, v), ei + P.eb[j]);
                if (P.beta != 0.0) val += P.beta * *cp;
                *cp = val;
            }
    }
}

// Sums the split partial tiles in split order and applies the epilogue; 256 elements of
// one 128 x 64 tile per workgroup.
__global__ __launch_bounds__(256) void ozaki_i8_reduce(OzProblem P) {
    constexpr int BM = 128, BN = 64;
    uint32_t const ntiles = P.tiles_m * P.tiles_n;
    uint32_t const tile   = blockIdx.x / (BM * BN / 256);
    int const e           = int(blockIdx.x % (BM * BN / 256)) * 256 + int(threadIdx.x);
    int const ii = e % BM, jj = e / BM;
    uint32_t tm, tn;
    tile_coords(tile, P.tiles_m, P.tiles_n, tm, tn);
    int64_t const i = int64_t(tm) * BM + ii, j = int64_t(tn) * BN + jj;
    if (i >= P.m || j >= P.n) return;
    double sum = 0.0;
    for (int s = 0; s < P.splits; ++s) sum += P.partial[(int64_t(s) * ntiles + tile) * (BM * BN) + e];
    double *cp = P.C + i * P.c_sm + j * P.c_sn;
    double val = P.alpha * ldexp(sum, P.ea[i] + P.eb[j]);
    if (P.beta != 0.0) val += P.beta * *cp;
    *cp = val;
}

int gemm_split_i8(GemmProblem<double> P, int ns, hipStream_t stream) {
    int64_t const kp        = ceil_div(P.k, I8_KS) * I8_KS;
    uint32_t const tiles_m  = uint32_t(ceil_div(P.m, 128));
    uint32_t const tiles_n  = uint32_t(ceil_div(P.n, 64));
    uint64_t const ntiles   = uint64_t(tiles_m) * tiles_n;
    // splits: enough for exact int32 accumulation (chunk <= I8_KC), and enough workgroups
    // to fill the GPU (one per CU resident) while every chunk keeps >= 1024 steps
    int64_t splits = ceil_div(kp, I8_KC);
    int64_t const want = ceil_div(int64_t(num_cus()) * 2, int64_t(ntiles));
    int64_t const cap  = std::max<int64_t>(1, kp / 1024);
    splits             = std::max(splits, std::min(want, cap));
    int64_t const kchunk = ceil_div(ceil_div(kp, splits), I8_KS) * I8_KS;
    splits               = ceil_div(kp, kchunk);
    if (ntiles * uint64_t(splits) > 0x7fffffffu) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: tile grid too large");

    size_t const a_bytes  = size_t(ns) * size_t(P.m) * size_t(kp);
    size_t const b_bytes  = size_t(ns) * size_t(P.n) * size_t(kp);
    size_t const b_off    = (a_bytes + 255) / 256 * 256;
    size_t const ea_off   = b_off + (b_bytes + 255) / 256 * 256;
    size_t const eb_off   = ea_off + (size_t(P.m) * 4 + 255) / 256 * 256;
    size_t const x_off    = eb_off + (size_t(P.n) * 4 + 255) / 256 * 256;
    size_t const p_off    = x_off + (row_exp_scratch(std::max(P.m, P.n)) * sizeof(double) + 255) / 256 * 256;
    size_t const p_bytes  = splits > 1 ? size_t(splits) * size_t(ntiles) * 128 * 64 * sizeof(double) : 0;
    void *ws = nullptr;
    EIN_TRY(workspace(stream, p_off + p_bytes, &ws));
    char *base  = static_cast<char *>(ws);
    int8_t *As  = reinterpret_cast<int8_t *>(base);
    int8_t *Bs  = reinterpret_cast<int8_t *>(base + b_off);
    int32_t *ea = reinterpret_cast<int32_t *>(base + ea_off);
    int32_t *eb = reinterpret_cast<int32_t *>(base + eb_off);
    SliceSrc SA{P.m, P.k, kp, P.A, P.a_sm, P.a_sk};
    SliceSrc SB{P.n, P.k, kp, P.B, P.b_sn, P.b_sk};
    double *scratch = reinterpret_cast<double *>(base + x_off);
    for (auto const &[S, e, out] : {std::tuple<SliceSrc, int32_t *, int8_t *>{SA, ea, As}, {SB, eb, Bs}}) {
        bool const rows_fast = S.s_row == 1 || (S.s_k != 1 && std::llabs(S.s_row) < std::llabs(S.s_k));
        EIN_TRY(row_exponents(S, e, true, scratch, stream));
        dim3 grid{uint32_t(ceil_div(S.kp, 64)), uint32_t(ceil_div(S.rows, 64))};
        if (grid.y > 65535) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: more than 4M rows");
        if (rows_fast)
            slice_i8_kernel<true><<<grid, 256, 0, stream>>>(S, e, ns, out);
        else
            slice_i8_kernel<false><<<grid, 256, 0, stream>>>(S, e, ns, out);
        EIN_LAUNCH_CHECK();
    }
    OzProblem Q{};
    Q.m       = P.m;
    Q.n       = P.n;
    Q.kp      = kp;
    Q.As_raw  = As;
    Q.Bs_raw  = Bs;
    Q.ea      = ea;
    Q.eb      = eb;
    Q.C       = P.C;
    Q.c_sm    = P.c_sm;
    Q.c_sn    = P.c_sn;
    Q.alpha   = P.alpha;
    Q.beta    = P.beta;
    Q.tiles_m = tiles_m;
    Q.tiles_n = tiles_n;
    Q.kchunk  = kchunk;
    Q.splits  = int32_t(splits);
    Q.partial = splits > 1 ? reinterpret_cast<double *>(base + p_off) : nullptr;
    dim3 grid{uint32_t(ntiles * uint64_t(splits))};
    switch (ns) {
    case 3: ozaki_i8_kernel<3><<<grid, 256, 0, stream>>>(Q); break;
    case 4: ozaki_i8_kernel<4><<<grid, 256, 0, stream>>>(Q); break;
    case 5: ozaki_i8_kernel<5><<<grid, 256, 0, stream>>>(Q); break;
    case 6: ozaki_i8_kernel<6><<<grid, 256, 0, stream>>>(Q); break;
    default: return fail(EIN_ERROR_VALUE, "split gemm: int8 slices must number 3..6");
    }
    EIN_LAUNCH_CHECK();
    if (splits > 1) {
        ozaki_i8_reduce<<<uint32_t(ntiles * (128 * 64 / 256)), 256, 0, stream>>>(Q);
        EIN_LAUNCH_CHECK();
    }
    return EIN_SUCCESS;
}

template <typename H, int NS>
void launch_oz(OzProblem const &P, int bm, hipStream_t stream) {
    if constexpr (NS <= 5) {
        if (bm == 128) {
            dim3 grid{P.tiles_m * P.tiles_n};
            ozaki_gemm_kernel<H, 128, 64, NS><<<grid, 512, 0, stream>>>(P);
            return;
        }
    }
    {
        dim3 grid{P.tiles_m * P.tiles_n};
        ozaki_gemm_kernel<H, 64, 64, NS><<<grid, 256, 0, stream>>>(P);
    }
}

template <typename H>
int slice_operand(SliceSrc const &S, int ns, int32_t *e, H *out, double *scratch, hipStream_t stream) {
    bool const rows_fast = S.s_row == 1 || (S.s_k != 1 && std::llabs(S.s_row) < std::llabs(S.s_k));
    EIN_TRY(row_exponents(S, e, false, scratch, stream));
    dim3 grid{uint32_t(ceil_div(S.kp, 64)), uint32_t(ceil_div(S.rows, 64))};
    if (grid.y > 65535) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: more than 4M rows");
    if (rows_fast)
        slice_kernel<H, true><<<grid, 256, 0, stream>>>(S, e, ns, out);
    else
        slice_kernel<H, false><<<grid, 256, 0, stream>>>(S, e, ns, out);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

template <typename H>
int gemm_split_typed(GemmProblem<double> P, int ns, int bm, hipStream_t stream) {
    int64_t const kp = ceil_div(P.k, 32) * 32;
    size_t const a_bytes = size_t(ns) * size_t(P.m) * size_t(kp) * sizeof(H);
    size_t const b_bytes = size_t(ns) * size_t(P.n) * size_t(kp) * sizeof(H);
    size_t const a_off = 0;
    size_t const b_off = (a_bytes + 255) / 256 * 256;
    size_t const ea_off = b_off + (b_bytes + 255) / 256 * 256;
    size_t const eb_off = ea_off + (size_t(P.m) * 4 + 255) / 256 * 256;
    size_t const x_off = eb_off + (size_t(P.n) * 4 + 255) / 256 * 256;
    size_t const total = x_off + row_exp_scratch(std::max(P.m, P.n)) * sizeof(double);
    void *ws = nullptr;
    EIN_TRY(workspace(stream, total, &ws));
    char *base = static_cast<char *>(ws);
    H *As = reinterpret_cast<H *>(base + a_off);
    H *Bs = reinterpret_cast<H *>(base + b_off);
    int32_t *ea = reinterpret_cast<int32_t *>(base + ea_off);
    int32_t *eb = reinterpret_cast<int32_t *>(base + eb_off);
    double *scratch = reinterpret_cast<double *>(base + x_off);

    SliceSrc SA{P.m, P.k, kp, P.A, P.a_sm, P.a_sk};
    SliceSrc SB{P.n, P.k, kp, P.B, P.b_sn, P.b_sk};
    EIN_TRY(slice_operand<H>(SA, ns, ea, As, scratch, stream));
    EIN_TRY(slice_operand<H>(SB, ns, eb, Bs, scratch, stream));

    OzProblem Q{};
    Q.m = P.m;
    Q.n = P.n;
    Q.kp = kp;
    Q.As_raw = As;
    Q.Bs_raw = Bs;
    Q.ea = ea;
    Q.eb = eb;
    Q.C = P.C;
    Q.c_sm = P.c_sm;
    Q.c_sn = P.c_sn;
    Q.alpha = P.alpha;
    Q.beta = P.beta;
    Q.tiles_m = uint32_t(ceil_div(P.m, bm));
    Q.tiles_n = uint32_t(ceil_div(P.n, 64));
    switch (ns) {
    case 3: launch_oz<H, 3>(Q, bm, stream); break;
    case 4: launch_oz<H, 4>(Q, bm, stream); break;
    case 5: launch_oz<H, 5>(Q, bm, stream); break;
    case 6:
        Q.tiles_m = uint32_t(ceil_div(P.m, 64));
        launch_oz<H, 6>(Q, 64, stream);
        break;
    default: return fail(EIN_ERROR_VALUE, "split gemm: nslices must be 3..6");
    }
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

} // namespace

// Scratch doubles row_exponents needs for `rows` rows.
size_t row_exp_scratch(int64_t rows) { return size_t(rows) + 65536; }

// Per-row exponents of a strided operand (e[r] with max_p |X(r,p)| < 2^e[r]).
int row_exponents(SliceSrc const &S, int32_t *e, bool guard, double *scratch, hipStream_t stream) {
    bool const rows_fast = S.s_row == 1 || (S.s_k != 1 && std::llabs(S.s_row) < std::llabs(S.s_k));
    if (!rows_fast) {
        row_exp_k_fast<<<uint32_t(ceil_div(S.rows, 4)), 256, 0, stream>>>(S, e, guard);
        EIN_LAUNCH_CHECK();
        return EIN_SUCCESS;
    }
    int64_t const rb = ceil_div(S.rows, 64);
    int64_t nck      = std::max<int64_t>(1, std::min<int64_t>(ceil_div(int64_t(num_cus()) * 4, rb), 64));
    nck              = std::min<int64_t>(nck, std::max<int64_t>(1, ceil_div(S.k, 256)));
    while (nck > 1 && size_t(nck) * size_t(S.rows) > row_exp_scratch(S.rows)) --nck;
    int64_t const kch = ceil_div(std::max<int64_t>(S.k, 1), nck);
    if (rb > 0x7fffffff) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: too many rows");
    row_max_partial<<<dim3(uint32_t(rb), uint32_t(nck)), 256, 0, stream>>>(S, kch, scratch);
    EIN_LAUNCH_CHECK();
    row_exp_finish<<<uint32_t(ceil_div(S.rows, 256)), 256, 0, stream>>>(S, scratch, int(nck), e, guard);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

namespace {

// Residual of the split representation of a strided operand, written along k:
// out[r*k + p] = X(r,p) - rep(X(r,p)). rep is the representation the schemes multiply:
// x = X 2^-e_r, then `steps` times y = x 2^bits, d = rint(y), x = y - d; the remainder is
// x 2^(e_r - bits*steps) (slices: bits 7 or 8, steps = slices; CRT: bits = kb, one step).
template <bool ROWS_FAST>
__global__ __launch_bounds__(256) void split_residual_kernel(SliceSrc S, int32_t const *__restrict__ e, int bits,
                                                             int steps, double *__restrict__ out) {
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
    int const lp    = t & 63;
    int64_t const p = p0 + lp;
    if (p >= S.k) return;
#pragma unroll 1
    for (int q = 0; q < 16; ++q) {
        int lr    = (t >> 6) + 4 * q;
        int64_t r = r0 + lr;
        if (r >= S.rows) continue;
        double x = ldexp(tile[lr][lp], -e[r]);
        for (int s = 0; s < steps; ++s) {
            double const y = ldexp(x, bits);
            x              = y - rint(y);
        }
        out[r * S.k + p] = ldexp(x, e[r] - bits * steps);
    }
}

int split_residual(SliceSrc const &S, bool guard, int bits, int steps, int32_t *e, double *out, double *scratch,
                   hipStream_t stream) {
    bool const rows_fast = S.s_row == 1 || (S.s_k != 1 && std::llabs(S.s_row) < std::llabs(S.s_k));
    EIN_TRY(row_exponents(S, e, guard, scratch, stream));
    dim3 grid{uint32_t(ceil_div(S.k, 64)), uint32_t(ceil_div(S.rows, 64))};
    if (grid.y > 65535) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: more than 4M rows");
    if (rows_fast)
        split_residual_kernel<true><<<grid, 256, 0, stream>>>(S, e, bits, steps, out);
    else
        split_residual_kernel<false><<<grid, 256, 0, stream>>>(S, e, bits, steps, out);
    EIN_LAUNCH_CHECK();
    return EIN_SUCCESS;
}

// One split GEMM of a normalised problem (C's unit stride along i, k > 0, alpha != 0).
int gemm_split_once(GemmProblem<double> const &P, int nslices, int type, hipStream_t stream) {
    int64_t const t128 = ceil_div(P.m, 128) * ceil_div(P.n, 64);
    int const bm       = t128 >= int64_t(num_cus()) * 2 ? 128 : 64;
    if (type == EIN_SLICE_CRT) return gemm_split_crt(P, nslices, stream);
    if (type == EIN_SLICE_INT8) return gemm_split_i8(P, nslices, stream);
    if (type == EIN_SLICE_BF16) return gemm_split_typed<__bf16>(P, nslices, bm, stream);
    return gemm_split_typed<_Float16>(P, nslices, bm, stream);
}

} // namespace

int gemm_split(GemmProblem<double> P, int nslices, int slice_type, hipStream_t stream) {
    bool const refine = (slice_type & EIN_SPLIT_REFINE) != 0;
    int const type    = slice_type & ~EIN_SPLIT_REFINE;
    if (type != EIN_SLICE_BF16 && type != EIN_SLICE_FP16 && type != EIN_SLICE_INT8 && type != EIN_SLICE_CRT)
        return fail(EIN_ERROR_VALUE,
                    "split gemm: slice_type must be EIN_SLICE_BF16, EIN_SLICE_FP16, EIN_SLICE_INT8 or EIN_SLICE_CRT");
    if (type == EIN_SLICE_CRT) {
        if (nslices != 0 && (nslices < CRT_MIN_MODULI || nslices > 16))
            return fail(EIN_ERROR_VALUE, "split gemm: EIN_SLICE_CRT takes 8..16 moduli (0: automatic)");
        if (nslices == 0) nslices = crt_auto_moduli(P.k);
    } else if (nslices < 3 || nslices > 6) {
        return fail(EIN_ERROR_VALUE, "split gemm: nslices must be 3..6");
    }
    if (refine && type != EIN_SLICE_CRT)
        return fail(EIN_ERROR_VALUE, "split gemm: EIN_SPLIT_REFINE applies to EIN_SLICE_CRT only (the slice schemes "
                                     "drop the cross products of slices t + u >= nslices, which residuals of A and B "
                                     "do not correct)");
    if (P.m <= 0 || P.n <= 0) return EIN_SUCCESS;
    if (P.batch != 1) return fail(EIN_ERROR_UNSUPPORTED, "split gemm: batched problems are not supported");
    // C's unit-stride direction must be i (lane-contiguous stores): transpose the problem otherwise
    bool const c_row_fast = (P.c_sn == 1 && P.c_sm != 1) || (P.c_sm != 1 && P.c_sn != 1 && P.c_sn < P.c_sm);
    if (c_row_fast) {
        GemmProblem<double> Q = P;
        Q.m = P.n;
        Q.n = P.m;
        Q.A = P.B;
        Q.a_sm = P.b_sn;
        Q.a_sk = P.b_sk;
        Q.B = P.A;
        Q.b_sk = P.a_sk;
        Q.b_sn = P.a_sm;
        Q.c_sm = P.c_sn;
        Q.c_sn = P.c_sm;
        P = Q;
    }
    if (P.k <= 0 || P.alpha == 0.0) {
        GemmConfig cfg;
        return gemm_strided<double>(P, &cfg, stream); // scales C by beta
    }
    EIN_TRY(gemm_split_once(P, nslices, type, stream));
    if (!refine) return EIN_SUCCESS;

    // One refinement correction: with dA = A - rep(A) and dB = B - rep(B) formed in FP64,
    // A B = rep(A) rep(B) + dA B + A dB - dA dB. The CRT GEMM above gives the first term
    // exactly (one rounding); the next two run as CRT GEMMs with beta = 1 and 8 moduli (their
    // magnitude is 2^-kb of C's, so their own 2^-24..2^-30 error is below FP64 rounding);
    // dA dB (2^-2kb) is dropped. The residual kernel also covers the slice representations
    // (bits 7 or 8 per slice), which the schemes above do not use: their error is dominated
    // by the dropped slice cross products.
    int const bits = crt_bits(nslices, P.k), steps = 1;
    bool const guard = false;
    int const corr_slices = CRT_MIN_MODULI;
    size_t const da = size_t(P.m) * size_t(P.k), db = size_t(P.n) * size_t(P.k);
    auto up         = [](size_t x) { return (x * sizeof(double) + 255) / 256 * 256; };
    size_t const ea_off = up(da) + up(db);
    size_t const eb_off = ea_off + (size_t(P.m) * 4 + 255) / 256 * 256;
    size_t const x_off  = eb_off + (size_t(P.n) * 4 + 255) / 256 * 256;
    size_t const total  = x_off + row_exp_scratch(std::max(P.m, P.n)) * sizeof(double);
    void *ws            = nullptr;
    EIN_TRY(workspace_slot(stream, 2, total, &ws)); // slot 0 is the split GEMM's, slot 1 the packed route's
    char *base = static_cast<char *>(ws);
    double *dA = reinterpret_cast<double *>(base);
    double *dB = reinterpret_cast<double *>(base + up(da));
    int32_t *ea = reinterpret_cast<int32_t *>(base + ea_off);
    int32_t *eb = reinterpret_cast<int32_t *>(base + eb_off);
    double *scratch = reinterpret_cast<double *>(base + x_off);
    EIN_TRY(split_residual(SliceSrc{P.m, P.k, P.k, P.A, P.a_sm, P.a_sk}, guard, bits, steps, ea, dA, scratch, stream));
    EIN_TRY(split_residual(SliceSrc{P.n, P.k, P.k, P.B, P.b_sn, P.b_sk}, guard, bits, steps, eb, dB, scratch, stream));
    GemmProblem<double> Q = P;
    Q.beta = 1.0;
    Q.A    = dA; // dA(i,p) = dA[i*k + p]
    Q.a_sm = P.k;
    Q.a_sk = 1;
    EIN_TRY(gemm_split_once(Q, corr_slices, type, stream));
    Q      = P;
    Q.beta = 1.0;
    Q.B    = dB; // dB(p,j) = dB[j*k + p]
    Q.b_sn = P.k;
    Q.b_sk = 1;
    return gemm_split_once(Q, corr_slices, type, stream);
}

} // namespace ein

extern "C" {

int ein_gemm_split_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                       const double *B, int64_t ldb, double beta, double *C, int64_t ldc, int32_t nslices,
                       int32_t slice_type, ein_stream_t stream) {
    auto notrans = [](char t) { return t == 'N' || t == 'n'; };
    auto valid = [&](char t) { return notrans(t) || t == 'T' || t == 't' || t == 'C' || t == 'c'; };
    if (!valid(transa)) return ein::fail(-1, "gemm_split: invalid transa");
    if (!valid(transb)) return ein::fail(-2, "gemm_split: invalid transb");
    if (m < 0) return ein::fail(-3, "gemm_split: m < 0");
    if (n < 0) return ein::fail(-4, "gemm_split: n < 0");
    if (k < 0) return ein::fail(-5, "gemm_split: k < 0");
    if (lda < std::max<int64_t>(1, notrans(transa) ? m : k)) return ein::fail(-8, "gemm_split: lda too small");
    if (ldb < std::max<int64_t>(1, notrans(transb) ? k : n)) return ein::fail(-10, "gemm_split: ldb too small");
    if (ldc < std::max<int64_t>(1, m)) return ein::fail(-13, "gemm_split: ldc too small");
    ein::GemmProblem<double> P{};
    P.m = m;
    P.n = n;
    P.k = k;
    P.batch = 1;
    P.A = A;
    P.B = B;
    P.C = C;
    P.alpha = alpha;
    P.beta = beta;
    P.a_sm = notrans(transa) ? 1 : lda;
    P.a_sk = notrans(transa) ? lda : 1;
    P.b_sk = notrans(transb) ? 1 : ldb;
    P.b_sn = notrans(transb) ? ldb : 1;
    P.c_sm = 1;
    P.c_sn = ldc;
    return ein::gemm_split(P, nslices, slice_type, ein::as_stream(stream));
}

} // extern "C"
