// Dense strided GEMM on MFMA for gfx950 (FP64 and FP32), strided batching, and a
// deterministic stream-K schedule for launches whose tile count does not fill the GPU.
//
// C(i,j) = alpha * sum_p A(i,p) B(p,j) + beta * C(i,j) with
//   A(i,p) = A[i*a_sm + p*a_sk + b*a_sb], B(p,j) = B[p*b_sk + j*b_sn + b*b_sb],
//   C(i,j) = C[i*c_sm + j*c_sn + b*c_sb].
// Column-major BLAS calls (ein_gemm_*) and the folded 2-D views of gemm-mapped einsums
// both land here. Reference leaf replaced: blas::gpu::gemm -> hipblasDgemm_64
// (libs/Einsums/hipBLAS/include/Einsums/hipBLAS.hpp:64, hipBLASVendor/src/gemm.hip:82-155).
#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "gemm.hpp"
#include "tile_engine.hpp"

namespace ein {

// Operand slabs go global -> LDS by DMA (tile_engine.hpp); AKG / BKG select the
// row-major (KROW) LDS layout for an operand contiguous along k, AW / BW the 16-byte DMA.
template <typename T, int BM, int BN, int BK, bool AKG, bool AW, bool BKG, bool BW>
__global__ __launch_bounds__(GEMM_THREADS, 2) void gemm_dense_kernel(GemmProblem<T> P) {
    using Cfg = TileCfg<T, BM, BN, BK>;
    using LA  = SlabLayout<T, BM, BK, AKG, AW>;
    using LB  = SlabLayout<T, BN, BK, BKG, BW>;
    __shared__ __attribute__((aligned(16))) T As[2 * LA::SIZE];
    __shared__ __attribute__((aligned(16))) T Bs[2 * LB::SIZE];

    int const tid  = threadIdx.x;
    int const lane = tid & 63;
    // wave-uniform, so per-instruction contraction offsets of the DMA stay scalar (not for
    // the dword row-major staging, where it makes the compiler keep more lane offsets live)
    constexpr bool dword_krow = (AKG && !AW) || (BKG && !BW);
    int const wave = dword_krow ? tid >> 6 : __builtin_amdgcn_readfirstlane(tid >> 6);
    int const wm   = wave & 1;
    int const wn   = wave >> 1;
    int const r    = lane & 15;
    int const q    = lane >> 4;

    StreamK const &sk       = P.sk;
    uint64_t const S        = sk.slabs;
    uint32_t const tiles_mn = P.tiles_m * P.tiles_n;
    uint64_t u0, u1, wlog;
    sk_range(sk, u0, u1, wlog);

    for (uint64_t u = u0; u < u1;) {
        uint64_t const t  = u / S;
        uint64_t const s0 = u - t * S;
        uint64_t const s1 = (S - s0 < u1 - u) ? S : s0 + (u1 - u);
        u += s1 - s0;

        int64_t const batch = int64_t(t / tiles_mn);
        uint32_t tm, tn;
        tile_coords(uint32_t(t % tiles_mn), P.tiles_m, P.tiles_n, tm, tn);
        int64_t const i0   = int64_t(tm) * BM;
        int64_t const j0   = int64_t(tn) * BN;
        int64_t const kend = int64_t(s1) * BK < P.k ? int64_t(s1) * BK : P.k;
        T const *__restrict__ Ab = P.A + batch * P.a_sb + i0 * P.a_sm;
        T const *__restrict__ Bb = P.B + batch * P.b_sb + j0 * P.b_sn;
        int64_t const mrem = P.m - i0, nrem = P.n - j0;
        auto offA = [&](int row, int64_t k0, int k, bool, bool &valid) -> int64_t {
            valid = row < mrem && k0 + k < kend;
            return row * P.a_sm + k0 * P.a_sk + int64_t(k) * P.a_sk;
        };
        auto offB = [&](int row, int64_t k0, int k, bool, bool &valid) -> int64_t {
            valid = row < nrem && k0 + k < kend;
            return row * P.b_sn + k0 * P.b_sk + int64_t(k) * P.b_sk;
        };
        // both operands' slabs at k0; without bounds checks when tile and slab are inside
        // (not for the dword row-major staging, whose registers a second copy would spill)
        bool const full_mn = !dword_krow && mrem >= BM && nrem >= BN;
        auto dma_both      = [&](T *Ad, T *Bd, int64_t k0) {
            if (full_mn && k0 + BK <= kend) {
                dma_slab<T, BM, BK, AKG, AW, false>(Ab, Ad, wave, lane, k0, offA);
                dma_slab<T, BN, BK, BKG, BW, false>(Bb, Bd, wave, lane, k0, offB);
            } else {
                dma_slab<T, BM, BK, AKG, AW>(Ab, Ad, wave, lane, k0, offA);
                dma_slab<T, BN, BK, BKG, BW>(Bb, Bd, wave, lane, k0, offB);
            }
        };

        WaveAcc<T, Cfg::FM, Cfg::FN> acc;
        acc.zero();
        int64_t const nslabs = int64_t(s1 - s0);
        int64_t const kbeg   = int64_t(s0) * BK;
        __syncthreads(); // LDS of the previous segment no longer read
        dma_both(As, Bs, kbeg);
        for (int64_t sl = 0; sl < nslabs; ++sl) {
            int const buf = int(sl & 1);
            dma_wait_all();
            auto next_dma = [&](T *An, T *Bn) {
                if (sl + 1 < nslabs) dma_both(An, Bn, kbeg + (sl + 1) * BK);
            };
            mfma_slab_lds<T, BM, BN, BK, AKG, AW, BKG, BW>(As + buf * LA::SIZE, Bs + buf * LB::SIZE,
                                                           As + (buf ^ 1) * LA::SIZE, Bs + (buf ^ 1) * LB::SIZE, acc, wm,
                                                           wn, lane, next_dma);
        }

        // Lane coordinates made opaque per segment so the compiler does not hoist the
        // 4*FM*FN store addresses out of the segment loop (they would pin registers).
        int ri = wm * Cfg::WM + r, rj = wn * Cfg::WN + q;
        asm volatile("" : "+v"(ri), "+v"(rj));
        if (s0 == 0 && s1 == S) {
            T *__restrict__ Cb = P.C + batch * P.c_sb;
            if (i0 + BM <= P.m && j0 + BN <= P.n) {
                // whole tile inside C: the reads of one column block are issued together
                // so their latency is paid once per block, not once per element
#pragma unroll
                for (int b = 0; b < Cfg::FN; ++b) {
                    T cv[Cfg::FM][4];
#pragma unroll
                    for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            int64_t i = i0 + ri + a * 16;
                            int64_t j = j0 + rj - q + b * 16 + Mfma<T>::row(q, v);
                            cv[a][v]  = P.beta != T(0) ? Cb[i * P.c_sm + j * P.c_sn] : T(0);
                        }
#pragma unroll
                    for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            int64_t i = i0 + ri + a * 16;
                            int64_t j = j0 + rj - q + b * 16 + Mfma<T>::row(q, v);
                            Cb[i * P.c_sm + j * P.c_sn] = P.alpha * acc.c[a][b][v] + P.beta * cv[a][v];
                        }
                }
            } else {
#pragma unroll
                for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
                    for (int b = 0; b < Cfg::FN; ++b)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            int64_t i = i0 + ri + a * 16;
                            int64_t j = j0 + rj - q + b * 16 + Mfma<T>::row(q, v);
                            if (i < P.m && j < P.n) {
                                T *cp = Cb + i * P.c_sm + j * P.c_sn;
                                T val = P.alpha * acc.c[a][b][v];
                                if (P.beta != T(0)) val += P.beta * *cp;
                                *cp = val;
                            }
                        }
            }
        } else {
            // tile cut by the range ends: raw accumulators to this workgroup's slot, [BN][BM]
            T *__restrict__ piece = P.partial + sk_slot(sk, wlog, t) * (BM * BN);
#pragma unroll
            for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
                for (int b = 0; b < Cfg::FN; ++b)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        int ii = ri + a * 16;
                        int jj = rj - q + b * 16 + Mfma<T>::row(q, v);
                        piece[jj * BM + ii] = acc.c[a][b][v];
                    }
        }
    }
}

// Stream-K fix-up: grid y runs over the stream-K tiles, grid x over parts of one tile;
// tiles finished in place return at once. Each thread sums the pieces of FIXUP_EPT
// elements in workgroup order (deterministic), then applies alpha and beta.
template <typename T, int BM, int BN>
__global__ __launch_bounds__(256) void gemm_streamk_fixup(GemmProblem<T> P) {
    StreamK const &sk = P.sk;
    uint64_t const t  = uint64_t(sk.dp_tiles) + blockIdx.y;
    int const e0      = int(blockIdx.x) * 256 + int(threadIdx.x);
    int const stride  = int(gridDim.x) * 256;
    T sum[FIXUP_EPT];
#pragma unroll
    for (int x = 0; x < FIXUP_EPT; ++x) sum[x] = T(0);
    if (sk.ksplit) { // split-K: the pieces of tile t in chunk order
        for (uint64_t c = 0; c < sk.ksplit; ++c) {
            T const *__restrict__ pc = P.partial + (c * sk.tiles + t) * (BM * BN);
#pragma unroll
            for (int x = 0; x < FIXUP_EPT; ++x) sum[x] += pc[e0 + x * stride];
        }
    } else {
        uint64_t g0, g1;
        sk_owners(sk, t, g0, g1);
        if (g0 == g1) return;
        for (uint64_t g = g0; g <= g1; ++g) {
            T const *__restrict__ pc = P.partial + sk_slot(sk, g, t) * (BM * BN);
#pragma unroll
            for (int x = 0; x < FIXUP_EPT; ++x) sum[x] += pc[e0 + x * stride];
        }
    }
    uint32_t const tiles_mn = P.tiles_m * P.tiles_n;
    int64_t const batch     = int64_t(t / tiles_mn);
    uint32_t tm, tn;
    tile_coords(uint32_t(t % tiles_mn), P.tiles_m, P.tiles_n, tm, tn);
    int64_t const i0 = int64_t(tm) * BM, j0 = int64_t(tn) * BN;
    T *__restrict__ Cb = P.C + batch * P.c_sb;
#pragma unroll
    for (int x = 0; x < FIXUP_EPT; ++x) {
        int const e = e0 + x * stride;
        int64_t i = i0 + e % BM, j = j0 + e / BM;
        if (i >= P.m || j >= P.n) continue;
        T *cp = Cb + i * P.c_sm + j * P.c_sn;
        T val = P.alpha * sum[x];
        if (P.beta != T(0)) val += P.beta * *cp;
        *cp = val;
    }
}

// C = beta*C over the m x n (x batch) view (k == 0 or alpha == 0 case); beta == 0 writes zeros.
template <typename T>
__global__ __launch_bounds__(256) void scale_2d_kernel(GemmProblem<T> P) {
    int64_t const mn    = P.m * P.n;
    int64_t const total = mn * P.batch;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        int64_t b   = e / mn;
        int64_t rem = e - b * mn;
        int64_t j   = rem / P.m;
        int64_t i   = rem - j * P.m;
        T *cp       = P.C + b * P.c_sb + i * P.c_sm + j * P.c_sn;
        *cp         = P.beta == T(0) ? T(0) : P.beta * *cp;
    }
}

namespace {

template <typename T, int BM, int BN, int BK, bool AKG, bool AW, bool BKG, bool BW>
void const *dense_kernel_ptr() {
    if constexpr (SlabLayout<T, BM, BK, AKG, AW>::VALID && SlabLayout<T, BN, BK, BKG, BW>::VALID)
        return reinterpret_cast<void const *>(&gemm_dense_kernel<T, BM, BN, BK, AKG, AW, BKG, BW>);
    else
        return nullptr;
}

// Kernel instance for staging code (A krow, A wide, B krow, B wide) as bits 3..0.
template <typename T, int BM, int BN, int BK>
void const *dense_kernel_for(int code) {
    switch (code) {
    case 0: return dense_kernel_ptr<T, BM, BN, BK, false, false, false, false>();
    case 1: return dense_kernel_ptr<T, BM, BN, BK, false, false, false, true>();
    case 2: return dense_kernel_ptr<T, BM, BN, BK, false, false, true, false>();
    case 3: return dense_kernel_ptr<T, BM, BN, BK, false, false, true, true>();
    case 4: return dense_kernel_ptr<T, BM, BN, BK, false, true, false, false>();
    case 5: return dense_kernel_ptr<T, BM, BN, BK, false, true, false, true>();
    case 6: return dense_kernel_ptr<T, BM, BN, BK, false, true, true, false>();
    case 7: return dense_kernel_ptr<T, BM, BN, BK, false, true, true, true>();
    case 8: return dense_kernel_ptr<T, BM, BN, BK, true, false, false, false>();
    case 9: return dense_kernel_ptr<T, BM, BN, BK, true, false, false, true>();
    case 10: return dense_kernel_ptr<T, BM, BN, BK, true, false, true, false>();
    case 11: return dense_kernel_ptr<T, BM, BN, BK, true, false, true, true>();
    case 12: return dense_kernel_ptr<T, BM, BN, BK, true, true, false, false>();
    case 13: return dense_kernel_ptr<T, BM, BN, BK, true, true, false, true>();
    case 14: return dense_kernel_ptr<T, BM, BN, BK, true, true, true, false>();
    default: return dense_kernel_ptr<T, BM, BN, BK, true, true, true, true>();
    }
}

// With `ws_bytes` set, only reports the workspace the launch would take (stream-K pieces).
template <typename T, int BM, int BN, int BK>
int launch_dense(GemmProblem<T> &P, int blocks_per_cu, hipStream_t stream, size_t *ws_bytes) {
    P.tiles_m      = uint32_t(ceil_div(P.m, BM));
    P.tiles_n      = uint32_t(ceil_div(P.n, BN));
    uint64_t tiles = uint64_t(P.tiles_m) * P.tiles_n * uint64_t(P.batch);
    uint64_t slabs = uint64_t(ceil_div(P.k, BK));
    if (tiles > 0x7fffffffu) return fail(EIN_ERROR_UNSUPPORTED, "gemm tile grid too large");
    OperandStaging const sa = choose_staging<T>(BM, P.a_sm, P.a_sk, P.a_sb, P.batch, P.m, P.k, P.A);
    OperandStaging const sb = choose_staging<T>(BN, P.b_sn, P.b_sk, P.b_sb, P.batch, P.n, P.k, P.B);
    int const code = (sa.krow ? 8 : 0) | (sa.wide ? 4 : 0) | (sb.krow ? 2 : 0) | (sb.wide ? 1 : 0);
    void const *kern = dense_kernel_for<T, BM, BN, BK>(code);
    if (kern == nullptr) return fail(EIN_ERROR_UNSUPPORTED, "gemm: operand staging not realisable for this tile");
    // resident workgroups: the stream-K schedule balances the last round over them
    int const fit  = resident_blocks(kern, GEMM_THREADS);
    int const bpc  = blocks_per_cu > 0 && blocks_per_cu < fit ? blocks_per_cu : fit;
    uint64_t slots = P.force_dp == 1 ? tiles : uint64_t(num_cus()) * uint64_t(bpc);
    P.sk           = P.force_dp == 2 ? plan_split_k(tiles, slabs, slots) : plan_stream_k(tiles, slabs, slots);
    P.partial      = nullptr;
    size_t const piece_bytes = size_t(sk_piece_slots(P.sk)) * BM * BN * sizeof(T);
    if (ws_bytes != nullptr) {
        *ws_bytes = piece_bytes;
        return EIN_SUCCESS;
    }
    if (P.sk.units != 0) {
        void *ws = nullptr;
        EIN_TRY(workspace(stream, piece_bytes, &ws));
        P.partial = static_cast<T *>(ws);
    }
    void *args[] = {&P};
    EIN_HIP_CHECK(hipLaunchKernel(kern, dim3(P.sk.grid), dim3(GEMM_THREADS), args, 0, stream));
    if (P.sk.units != 0) {
        gemm_streamk_fixup<T, BM, BN><<<sk_fixup_grid(P.sk, BM * BN), 256, 0, stream>>>(P);
        EIN_LAUNCH_CHECK();
    }
    return EIN_SUCCESS;
}

} // namespace

GemmConfig choose_gemm_config(int64_t m, int64_t n, int64_t k, int64_t batch, int elem_bytes) {
    GemmConfig cfg;
    if (m < 128 || n < 128) {
        cfg.bm = 64;
        cfg.bn = 64;
        return cfg;
    }
    // 128 x 128 tiles do four times the MFMA work of 64 x 64 per slab barrier. Measured on
    // MI355X (FP64): with fewer than a quarter round of them (576 x 576 x 9216) 128 x 64
    // tiles balance better and halve the stream-K pieces.
    (void)k;
    (void)elem_bytes;
    int64_t const t128 = ceil_div(m, 128) * ceil_div(n, 128) * batch;
    if (t128 < num_cus() / 4) cfg.bn = 64;
    // At least one round of 128 x 128 tiles: plan the schedule for one workgroup per CU. A
    // second resident workgroup per CU adds no MFMA throughput here (8192^3 and 16384^2 x
    // 4096 FP64 measure the same either way), and planning for one keeps whole tiles in the
    // data-parallel rounds where planning for two splits the last round, or with one to two
    // rounds every tile, into stream-K pieces. Measured FP64: 2304^3 56.2 -> 58.8 TFLOP/s,
    // 3072^3 65.8 -> 66.4, the Wabef update (9216 x 9216 x 576) 1.565 -> 1.555 ms.
    if (cfg.bn == 128 && t128 >= num_cus()) cfg.blocks_per_cu = 1;
    return cfg;
}

// ------------------------------------------------------------- tuning table
// Measured GEMM configurations, keyed by the normalised problem (C's unit-stride direction
// along i): element size, m, n, k, batch and whether A and B are contiguous along k (their
// staging class). A call whose key is in the table runs the stored configuration; with tuning
// mode on, a call whose key is absent first times every candidate on its own operands
// (writing a scratch C, beta = 0) and stores the fastest. The table can be saved to and
// loaded from a text file (one "elem m n k batch ak bk : bm bn dp bpc" line per entry).
struct TuneKey {
    int eb;
    int64_t m, n, k, batch;
    int ak, bk;
    bool operator<(TuneKey const &o) const {
        return std::tie(eb, m, n, k, batch, ak, bk) < std::tie(o.eb, o.m, o.n, o.k, o.batch, o.ak, o.bk);
    }
};

namespace {
std::mutex g_tune_mutex;
std::map<TuneKey, GemmConfig> g_tune;
bool g_tuning = false;

template <typename T>
TuneKey tune_key(GemmProblem<T> const &P) {
    return TuneKey{int(sizeof(T)), P.m, P.n, P.k, P.batch, P.a_sk == 1 ? 1 : 0, P.b_sk == 1 ? 1 : 0};
}

// Candidate configurations: tile shape x schedule (hybrid stream-K or data-parallel) x
// workgroups per CU the schedule plans for (all that fit, or one).
// krow_both: both operands contiguous along k (row-major LDS staging), the case the 96 x 96
// tile is built for (576 = 6 x 96: the Wmnij GEMM at o=24 has no ragged tiles).
std::vector<GemmConfig> tune_candidates(bool krow_both) {
    std::vector<GemmConfig> c;
    if (krow_both)
        for (int dp : {0, 1, 2})
            for (int bpc : {0, 1}) {
                if (dp == 1 && bpc == 1) continue;
                GemmConfig g;
                g.bm            = 96;
                g.bn            = 96;
                g.data_parallel = dp;
                g.blocks_per_cu = bpc;
                c.push_back(g);
            }
    for (int bm : {128, 64})
        for (int bn : {128, 64})
            for (int dp : {0, 1, 2})
                for (int bpc : {0, 1}) {
                    if (dp == 1 && bpc == 1) continue; // data-parallel ignores the planned occupancy
                    GemmConfig g;
                    g.bm            = bm;
                    g.bn            = bn;
                    g.data_parallel = dp;
                    g.blocks_per_cu = bpc;
                    c.push_back(g);
                }
    return c;
}
} // namespace

bool tuned_config(TuneKey const &key, GemmConfig *out) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    *out = it->second;
    return true;
}

namespace {

template <typename T>
int gemm_strided_impl(GemmProblem<T> P, GemmConfig const *force, hipStream_t stream, size_t *ws_bytes);

// Times every candidate on P's operands and returns the fastest. C is a scratch buffer with
// P's beta (its values do not matter for the timing, the C traffic does). Each candidate is
// timed in two passes, the candidate list in forward then reverse order (so clock drift over
// the tuning run does not favour the early or the late candidates), each pass running it
// for about 10 ms (3 to 200 launches) after a warm-up; its time is the smaller of the two.
// P is normalised.
// Whether the 96 x 96 tile applies: FP64 (its FP32 16-byte row-major staging is not
// realisable at 96 rows) with both operands contiguous along k.
template <typename T>
bool krow_both(GemmProblem<T> const &P) {
    return sizeof(T) == 8 && choose_staging<T>(96, P.a_sm, P.a_sk, P.a_sb, P.batch, P.m, P.k, P.A).krow &&
           choose_staging<T>(96, P.b_sn, P.b_sk, P.b_sb, P.batch, P.n, P.k, P.B).krow;
}

template <typename T>
int tune_gemm(GemmProblem<T> const &P, hipStream_t stream, GemmConfig *best) {
    size_t const c_elems = size_t(P.m) * size_t(P.n) * size_t(P.batch);
    void *scratch        = nullptr;
    EIN_TRY(ein_malloc(&scratch, c_elems * sizeof(T)));
    GemmProblem<T> Q = P;
    Q.C              = static_cast<T *>(scratch);
    Q.c_sm           = 1;
    Q.c_sn           = P.m;
    Q.c_sb           = P.m * P.n;
    Q.beta           = P.beta;
    std::vector<GemmConfig> const cands = tune_candidates(krow_both(P));
    std::vector<float> best_of(cands.size(), 0.f);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = EIN_SUCCESS;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = fail(EIN_ERROR_HIP, "tuning events");
    auto timed = [&](GemmConfig const &c, int reps, float *ms) {
        (void)hipEventRecord(e0, stream);
        for (int r = 0; r < reps && st == EIN_SUCCESS; ++r) st = gemm_strided_impl<T>(Q, &c, stream, nullptr);
        (void)hipEventRecord(e1, stream);
        if (st == EIN_SUCCESS && hipEventSynchronize(e1) != hipSuccess) st = fail(EIN_ERROR_HIP, "tuning synchronisation");
        if (st == EIN_SUCCESS) (void)hipEventElapsedTime(ms, e0, e1);
    };
    for (int pass = 0; pass < 2 && st == EIN_SUCCESS; ++pass)
        for (size_t q = 0; q < cands.size() && st == EIN_SUCCESS; ++q) {
            size_t const i = pass == 0 ? q : cands.size() - 1 - q;
            float first = 0.f, ms = 0.f;
            timed(cands[i], 1, &first); // warm-up, and the length of one launch
            int const reps = std::max(3, std::min(200, int(10.0f / std::max(first, 1e-3f))));
            timed(cands[i], reps, &ms);
            ms /= float(reps);
            if (best_of[i] == 0.f || ms < best_of[i]) best_of[i] = ms;
        }
    if (st == EIN_SUCCESS) {
        size_t arg = 0;
        for (size_t i = 1; i < cands.size(); ++i)
            if (best_of[i] < best_of[arg]) arg = i;
        *best = cands[arg];
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)ein_free_async(scratch, reinterpret_cast<ein_stream_t>(stream));
    return st;
}

template <typename T>
int gemm_strided_impl(GemmProblem<T> P, GemmConfig const *force, hipStream_t stream, size_t *ws_bytes) {
    if (ws_bytes != nullptr) *ws_bytes = 0;
    if (P.m <= 0 || P.n <= 0 || P.batch <= 0) return EIN_SUCCESS;
    // Normalise so that C's unit-stride direction is i (lane-contiguous stores):
    // C^T = B^T A^T swaps the roles of the operands.
    bool const c_row_fast = (P.c_sn == 1 && P.c_sm != 1) || (P.c_sm != 1 && P.c_sn != 1 && P.c_sn < P.c_sm);
    if (c_row_fast) {
        GemmProblem<T> Q = P;
        Q.m    = P.n;
        Q.n    = P.m;
        Q.A    = P.B;
        Q.a_sm = P.b_sn;
        Q.a_sk = P.b_sk;
        Q.a_sb = P.b_sb;
        Q.B    = P.A;
        Q.b_sk = P.a_sk;
        Q.b_sn = P.a_sm;
        Q.b_sb = P.a_sb;
        Q.c_sm = P.c_sn;
        Q.c_sn = P.c_sm;
        P      = Q;
    }
    int64_t const total = P.m * P.n * P.batch;
    if (P.k <= 0 || P.alpha == T(0)) {
        if (ws_bytes != nullptr) return EIN_SUCCESS;
        int blocks = int(std::min<int64_t>(ceil_div(total, 256), 65536));
        scale_2d_kernel<T><<<blocks, 256, 0, stream>>>(P);
        EIN_LAUNCH_CHECK();
        return EIN_SUCCESS;
    }
    GemmConfig tuned;
    bool have_tuned = false;
    if (!force) {
        TuneKey const key = tune_key(P);
        have_tuned        = tuned_config(key, &tuned);
        bool tuning;
        {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            tuning = g_tuning;
        }
        if (!have_tuned && tuning) {
            if (ws_bytes != nullptr) {
                // an untuned key in tuning mode may run any candidate: report the largest need
                size_t most = 0;
                for (GemmConfig const &c : tune_candidates(krow_both(P))) {
                    size_t b = 0;
                    EIN_TRY(gemm_strided_impl<T>(P, &c, stream, &b));
                    most = std::max(most, b);
                }
                *ws_bytes = most;
                return EIN_SUCCESS;
            }
            if (!stream_capturing(stream)) {
                EIN_TRY(tune_gemm<T>(P, stream, &tuned));
                std::lock_guard<std::mutex> lock(g_tune_mutex);
                g_tune[key] = tuned;
                have_tuned  = true;
            }
        }
    }
    if (have_tuned) force = &tuned;
    GemmConfig cfg = force ? *force : choose_gemm_config(P.m, P.n, P.k, P.batch, int(sizeof(T)));
    if (!force && cfg.bm == 128 && cfg.bn == 64) {
        // few tiles with both operands staged row-major (contiguous along k): 64 x 64
        // measured 4% faster than 128 x 64 on 576 x 576 x 9216
        bool const a_krow = choose_staging<T>(128, P.a_sm, P.a_sk, P.a_sb, P.batch, P.m, P.k, P.A).krow;
        bool const b_krow = choose_staging<T>(64, P.b_sn, P.b_sk, P.b_sb, P.batch, P.n, P.k, P.B).krow;
        if (a_krow && b_krow) cfg.bm = 64;
    }
    P.force_dp     = cfg.data_parallel;
    constexpr int BK = 16;
    if (cfg.bm == 96 && cfg.bn == 96) {
        if (!krow_both(P))
            return fail(EIN_ERROR_UNSUPPORTED, "gemm: 96 x 96 tiles need FP64 operands contiguous along k");
        return launch_dense<T, 96, 96, BK>(P, cfg.blocks_per_cu, stream, ws_bytes);
    }
    if (cfg.bm == 128 && cfg.bn == 128) return launch_dense<T, 128, 128, BK>(P, cfg.blocks_per_cu, stream, ws_bytes);
    if (cfg.bm == 128 && cfg.bn == 64) return launch_dense<T, 128, 64, BK>(P, cfg.blocks_per_cu, stream, ws_bytes);
    if (cfg.bm == 64 && cfg.bn == 128) return launch_dense<T, 64, 128, BK>(P, cfg.blocks_per_cu, stream, ws_bytes);
    return launch_dense<T, 64, 64, BK>(P, cfg.blocks_per_cu, stream, ws_bytes);
}

} // namespace

template <typename T>
int gemm_strided(GemmProblem<T> P, GemmConfig const *force, hipStream_t stream) {
    return gemm_strided_impl<T>(P, force, stream, nullptr);
}

template <typename T>
int gemm_workspace_bytes(GemmProblem<T> const &P, size_t *bytes) {
    return gemm_strided_impl<T>(P, nullptr, nullptr, bytes);
}

template int gemm_strided<double>(GemmProblem<double>, GemmConfig const *, hipStream_t);
template int gemm_strided<float>(GemmProblem<float>, GemmConfig const *, hipStream_t);
template int gemm_workspace_bytes<double>(GemmProblem<double> const &, size_t *);
template int gemm_workspace_bytes<float>(GemmProblem<float> const &, size_t *);

namespace {

bool valid_trans(char t) { return t == 'N' || t == 'n' || t == 'T' || t == 't' || t == 'C' || t == 'c'; }
bool is_notrans(char t) { return t == 'N' || t == 'n'; }

// Column-major BLAS argument checks (same order and codes as the reference's
// hipBLASVendor/src/gemm.hip:23-80 and reference BLAS xGEMM).
int check_gemm_args(char ta, char tb, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc) {
    if (!valid_trans(ta)) return fail(-1, "gemm: invalid transa");
    if (!valid_trans(tb)) return fail(-2, "gemm: invalid transb");
    if (m < 0) return fail(-3, "gemm: m < 0");
    if (n < 0) return fail(-4, "gemm: n < 0");
    if (k < 0) return fail(-5, "gemm: k < 0");
    int64_t nrowa = is_notrans(ta) ? m : k;
    int64_t nrowb = is_notrans(tb) ? k : n;
    if (lda < std::max<int64_t>(1, nrowa)) return fail(-8, "gemm: lda too small");
    if (ldb < std::max<int64_t>(1, nrowb)) return fail(-10, "gemm: ldb too small");
    if (ldc < std::max<int64_t>(1, m)) return fail(-13, "gemm: ldc too small");
    return EIN_SUCCESS;
}

template <typename T>
int gemm_colmajor(char ta, char tb, int64_t m, int64_t n, int64_t k, T alpha, T const *A, int64_t lda, int64_t sa,
                  T const *B, int64_t ldb, int64_t sb, T beta, T *C, int64_t ldc, int64_t sc, int64_t batch,
                  hipStream_t stream) {
    EIN_TRY(check_gemm_args(ta, tb, m, n, k, lda, ldb, ldc));
    if (batch < 0) return fail(-17, "gemm: batch < 0");
    GemmProblem<T> P{};
    P.m     = m;
    P.n     = n;
    P.k     = k;
    P.batch = batch;
    P.A     = A;
    P.B     = B;
    P.C     = C;
    P.alpha = alpha;
    P.beta  = beta;
    if (is_notrans(ta)) {
        P.a_sm = 1;
        P.a_sk = lda;
    } else {
        P.a_sm = lda;
        P.a_sk = 1;
    }
    if (is_notrans(tb)) {
        P.b_sk = 1;
        P.b_sn = ldb;
    } else {
        P.b_sk = ldb;
        P.b_sn = 1;
    }
    P.c_sm = 1;
    P.c_sn = ldc;
    P.a_sb = sa;
    P.b_sb = sb;
    P.c_sb = sc;
    return gemm_strided<T>(P, nullptr, stream);
}

// C = alpha*op(A)*op(B) + beta*C with C in host memory (column-major, ldc) and A, B on the
// device: the output for a C that does not fit in HBM next to its operands (the v^4
// intermediates of a large CCD). Column panels of C cycle through two device buffers. Panel
// i's upload (stream `up`), its GEMM (the caller's stream) and its download (stream `down`)
// are chained by events, so the upload of panel i+1 and the download of panel i-1 overlap the
// GEMM of panel i, and both PCIe directions run at once. A buffer is reused only after its
// previous download. Returns when C is back in host memory. Host C should be pinned
// (hipHostMalloc); pageable memory works but copies through a staging buffer.
template <typename T>
int gemm_host_c(char ta, char tb, int64_t m, int64_t n, int64_t k, T alpha, T const *A, int64_t lda, T const *B,
                int64_t ldb, T beta, T *C, int64_t ldc, size_t panel_bytes, int nslices, int slice_type,
                hipStream_t stream) {
    EIN_TRY(check_gemm_args(ta, tb, m, n, k, lda, ldb, ldc));
    if (C == nullptr && m > 0 && n > 0) return fail(-12, "gemm_host_c: C is null");
    if (m == 0 || n == 0) return EIN_SUCCESS;
    if (panel_bytes == 0) panel_bytes = size_t(256) << 20;
    int64_t const cols = std::max<int64_t>(1, std::min<int64_t>(n, int64_t(panel_bytes / (size_t(m) * sizeof(T)))));
    int64_t const npan = ceil_div(n, cols);
    static std::mutex mu; // the side streams and events below are shared by all calls
    std::lock_guard<std::mutex> lock(mu);
    static hipStream_t up = nullptr, down = nullptr;
    static hipEvent_t ev_start, ev_up[2], ev_comp[2], ev_down[2];
    if (up == nullptr) {
        EIN_HIP_CHECK(hipStreamCreateWithFlags(&up, hipStreamNonBlocking));
        EIN_HIP_CHECK(hipStreamCreateWithFlags(&down, hipStreamNonBlocking));
        EIN_HIP_CHECK(hipEventCreateWithFlags(&ev_start, hipEventDisableTiming));
        for (int b = 0; b < 2; ++b) {
            EIN_HIP_CHECK(hipEventCreateWithFlags(&ev_up[b], hipEventDisableTiming));
            EIN_HIP_CHECK(hipEventCreateWithFlags(&ev_comp[b], hipEventDisableTiming));
            EIN_HIP_CHECK(hipEventCreateWithFlags(&ev_down[b], hipEventDisableTiming));
        }
    }
    void *buf[2] = {nullptr, nullptr};
    size_t const pbytes = size_t(m) * size_t(cols) * sizeof(T);
    EIN_TRY(ein_malloc(&buf[0], pbytes));
    if (int const rc = ein_malloc(&buf[1], pbytes); rc != EIN_SUCCESS) {
        ein_free(buf[0]);
        return rc;
    }
    int status = EIN_SUCCESS;
    auto run = [&]() -> int {
        // the side streams start after the work already queued on the caller's stream
        EIN_HIP_CHECK(hipEventRecord(ev_start, stream));
        EIN_HIP_CHECK(hipStreamWaitEvent(up, ev_start, 0));
        EIN_HIP_CHECK(hipStreamWaitEvent(down, ev_start, 0));
        for (int64_t p = 0; p < npan; ++p) {
            int const b     = int(p & 1);
            int64_t const j0 = p * cols;
            int64_t const nc = std::min<int64_t>(cols, n - j0);
            T *Cd            = static_cast<T *>(buf[b]);
            T *Ch            = C + j0 * ldc;
            if (p >= 2) EIN_HIP_CHECK(hipStreamWaitEvent(up, ev_down[b], 0)); // buffer free again
            if (beta != T(0))
                EIN_HIP_CHECK(hipMemcpy2DAsync(Cd, size_t(m) * sizeof(T), Ch, size_t(ldc) * sizeof(T),
                                               size_t(m) * sizeof(T), size_t(nc), hipMemcpyHostToDevice, up));
            EIN_HIP_CHECK(hipEventRecord(ev_up[b], up));
            EIN_HIP_CHECK(hipStreamWaitEvent(stream, ev_up[b], 0));
            T const *Bp = is_notrans(tb) ? B + j0 * ldb : B + j0;
            if (nslices > 0) {
                if constexpr (sizeof(T) == 8)
                    EIN_TRY(ein_gemm_split_f64(ta, tb, m, nc, k, alpha, A, lda, Bp, ldb, beta, Cd, m, nslices, slice_type,
                                               stream));
                else
                    return fail(EIN_ERROR_UNSUPPORTED, "gemm_host_c: split precision is FP64 only");
            } else {
                EIN_TRY(gemm_colmajor<T>(ta, tb, m, nc, k, alpha, A, lda, 0, Bp, ldb, 0, beta, Cd, m, 0, 1, stream));
            }
            EIN_HIP_CHECK(hipEventRecord(ev_comp[b], stream));
            EIN_HIP_CHECK(hipStreamWaitEvent(down, ev_comp[b], 0));
            EIN_HIP_CHECK(hipMemcpy2DAsync(Ch, size_t(ldc) * sizeof(T), Cd, size_t(m) * sizeof(T), size_t(m) * sizeof(T),
                                           size_t(nc), hipMemcpyDeviceToHost, down));
            EIN_HIP_CHECK(hipEventRecord(ev_down[b], down));
        }
        return EIN_SUCCESS;
    };
    status = run();
    // the buffers go back to the pool only after every queued copy and GEMM has finished
    hipError_t const e1 = hipStreamSynchronize(down);
    hipError_t const e2 = hipStreamSynchronize(up);
    hipError_t const e3 = hipStreamSynchronize(stream);
    ein_free(buf[0]);
    ein_free(buf[1]);
    if (status != EIN_SUCCESS) return status;
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
        return fail(EIN_ERROR_HIP, std::string("gemm_host_c: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3));
    return EIN_SUCCESS;
}

} // namespace
} // namespace ein

extern "C" {

int ein_gemm_host_c_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                        int64_t lda, const double *B, int64_t ldb, double beta, double *C, int64_t ldc,
                        size_t panel_bytes, int32_t nslices, int32_t slice_type, ein_stream_t stream) {
    return ein::gemm_host_c<double>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, panel_bytes, nslices,
                                    slice_type, ein::as_stream(stream));
}
int ein_gemm_host_c_f32(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                        int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc, size_t panel_bytes,
                        ein_stream_t stream) {
    return ein::gemm_host_c<float>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, panel_bytes, 0, 0,
                                   ein::as_stream(stream));
}

int ein_gemm_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                 const double *B, int64_t ldb, double beta, double *C, int64_t ldc, ein_stream_t stream) {
    return ein::gemm_colmajor<double>(transa, transb, m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1,
                                      ein::as_stream(stream));
}

int ein_gemm_f32(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda,
                 const float *B, int64_t ldb, float beta, float *C, int64_t ldc, ein_stream_t stream) {
    return ein::gemm_colmajor<float>(transa, transb, m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1,
                                     ein::as_stream(stream));
}

int ein_gemm_strided_batched_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                                 int64_t lda, int64_t strideA, const double *B, int64_t ldb, int64_t strideB, double beta,
                                 double *C, int64_t ldc, int64_t strideC, int64_t batch, ein_stream_t stream) {
    return ein::gemm_colmajor<double>(transa, transb, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                      strideC, batch, ein::as_stream(stream));
}

int ein_gemm_strided_batched_f32(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                                 int64_t lda, int64_t strideA, const float *B, int64_t ldb, int64_t strideB, float beta,
                                 float *C, int64_t ldc, int64_t strideC, int64_t batch, ein_stream_t stream) {
    return ein::gemm_colmajor<float>(transa, transb, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                     strideC, batch, ein::as_stream(stream));
}

int ein_gemm_tuning_mode(int32_t on) {
    std::lock_guard<std::mutex> lock(ein::g_tune_mutex);
    int32_t const prev = ein::g_tuning ? 1 : 0;
    ein::g_tuning      = on != 0;
    return prev;
}

int ein_gemm_tuning_clear(void) {
    std::lock_guard<std::mutex> lock(ein::g_tune_mutex);
    ein::g_tune.clear();
    return EIN_SUCCESS;
}

int64_t ein_gemm_tuning_count(void) {
    std::lock_guard<std::mutex> lock(ein::g_tune_mutex);
    return int64_t(ein::g_tune.size());
}

int ein_gemm_tuning_lookup(int32_t elem_bytes, int64_t m, int64_t n, int64_t k, int64_t batch, int32_t a_kcontig,
                           int32_t b_kcontig, int32_t *config) {
    if (config == nullptr) return -8;
    ein::GemmConfig c;
    if (!ein::tuned_config(ein::TuneKey{elem_bytes, m, n, k, batch, a_kcontig ? 1 : 0, b_kcontig ? 1 : 0}, &c))
        return ein::fail(EIN_ERROR_VALUE, "gemm tuning: no entry for this problem");
    config[0] = c.bm;
    config[1] = c.bn;
    config[2] = c.data_parallel;
    config[3] = c.blocks_per_cu;
    return EIN_SUCCESS;
}

int ein_gemm_tuning_save(const char *path) {
    if (path == nullptr) return -1;
    FILE *f = std::fopen(path, "w");
    if (f == nullptr) return ein::fail(EIN_ERROR_VALUE, std::string("gemm tuning: cannot write ") + path);
    std::lock_guard<std::mutex> lock(ein::g_tune_mutex);
    std::fprintf(f, "# einsums_amd GEMM tuning table: elem m n k batch a_kcontig b_kcontig : bm bn data_parallel blocks_per_cu\n");
    for (auto const &kv : ein::g_tune)
        std::fprintf(f, "%d %lld %lld %lld %lld %d %d : %d %d %d %d\n", kv.first.eb, (long long)kv.first.m,
                     (long long)kv.first.n, (long long)kv.first.k, (long long)kv.first.batch, kv.first.ak, kv.first.bk,
                     kv.second.bm, kv.second.bn, kv.second.data_parallel, kv.second.blocks_per_cu);
    std::fclose(f);
    return EIN_SUCCESS;
}

int ein_gemm_tuning_load(const char *path) {
    if (path == nullptr) return -1;
    FILE *f = std::fopen(path, "r");
    if (f == nullptr) return ein::fail(EIN_ERROR_VALUE, std::string("gemm tuning: cannot read ") + path);
    std::map<ein::TuneKey, ein::GemmConfig> loaded;
    char line[512];
    int lineno = 0;
    while (std::fgets(line, sizeof line, f) != nullptr) {
        ++lineno;
        if (line[0] == '#' || line[0] == '\n') continue;
        int eb, ak, bk, bm, bn, dp, bpc;
        long long m, n, k, b;
        if (std::sscanf(line, "%d %lld %lld %lld %lld %d %d : %d %d %d %d", &eb, &m, &n, &k, &b, &ak, &bk, &bm, &bn, &dp,
                        &bpc) != 11 ||
            (eb != 4 && eb != 8) || (bm != 64 && bm != 96 && bm != 128) || (bn != 64 && bn != 96 && bn != 128) || (bm == 96) != (bn == 96) ||
            dp < 0 || dp > 2 || bpc < 0 ||
            m <= 0 || n <= 0 || k <= 0 || b <= 0) {
            std::fclose(f);
            return ein::fail(EIN_ERROR_VALUE, "gemm tuning: malformed line " + std::to_string(lineno) + " in " + path);
        }
        ein::GemmConfig c;
        c.bm            = bm;
        c.bn            = bn;
        c.data_parallel = dp;
        c.blocks_per_cu = bpc;
        loaded[ein::TuneKey{eb, m, n, k, b, ak ? 1 : 0, bk ? 1 : 0}] = c;
    }
    std::fclose(f);
    std::lock_guard<std::mutex> lock(ein::g_tune_mutex);
    for (auto const &kv : loaded) ein::g_tune[kv.first] = kv.second;
    return EIN_SUCCESS;
}

} // extern "C"
