// Generic binary contraction on gfx950 (any label pattern) without atomics.
//
// Reference behaviour restated: einsum_generic_algorithm (TensorAlgebra/Backends/
// BaseAlgebra.hpp:85-139, device launcher GPUTensorAlgebra.hpp:200-342): C is scaled by
// C_prefactor (zeroed when it is 0), then every target point receives
// AB_prefactor * sum over the link labels of A*B. Repeated labels on one tensor address a
// diagonal (their strides add, as get_stride_for does in Detail/Utilities.hpp:631-652).
//
// Instead of one thread per point of the full index space with an atomicAdd per point,
// the labels are split into four groups -- batch (in A, B, C), M (in C and A only, or C
// only), N (in C and B only) and K (not in C) -- and the contraction becomes
//     C[b, i, j] = alpha * sum_p A[b, i, p] * B[b, p, j] + beta * C[b, i, j]
// over multi-indices decoded with fast 32-bit division. Large M and N go to an MFMA
// strided-gather GEMM; otherwise each target point is reduced by one thread or one
// wave (optionally split over K with a deterministic second pass).
#include <algorithm>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "contract.hpp"
#include "einsum.hpp"
#include "tile_engine.hpp"

namespace ein {

// ============================================================ host analysis

int analyse_contraction(ein_tensor_desc const &C, ein_tensor_desc const &A, ein_tensor_desc const &B,
                        ContractionAnalysis &out) {
    ein_tensor_desc const *T[3] = {&C, &A, &B};
    char const *names[3]        = {"C", "A", "B"};
    int labels[3 * EIN_MAX_RANK];
    out.nlab = 0;
    for (int t = 0; t < 3; ++t) {
        if (T[t]->rank < 0 || T[t]->rank > EIN_MAX_RANK)
            return fail(EIN_ERROR_RANK, std::string("rank of ") + names[t] + " out of range");
    }
    for (int i = 0; i < 3 * EIN_MAX_RANK; ++i) {
        out.ext[i] = -1;
        out.sA[i] = out.sB[i] = out.sC[i] = 0;
        out.inA[i] = out.inB[i] = out.inC[i] = false;
    }
    out.c_repeats = false;
    out.c_size    = 1;
    for (int t = 0; t < 3; ++t) {
        for (int d = 0; d < T[t]->rank; ++d) {
            int lab = T[t]->labels[d];
            int64_t e = T[t]->dims[d];
            if (e < 0) return fail(EIN_ERROR_VALUE, std::string("negative extent in ") + names[t]);
            int u = -1;
            for (int q = 0; q < out.nlab; ++q)
                if (labels[q] == lab) u = q;
            if (u < 0) {
                u           = out.nlab++;
                labels[u]   = lab;
                out.ext[u]  = e;
            } else if (out.ext[u] != e) {
                return fail(EIN_ERROR_DIMENSION, "Inconsistent dimensions found!");
            }
            int64_t s = T[t]->strides[d];
            if (t == 0) {
                if (out.inC[u]) out.c_repeats = true;
                out.inC[u] = true;
                out.sC[u] += s;
            } else if (t == 1) {
                out.inA[u] = true;
                out.sA[u] += s;
            } else {
                out.inB[u] = true;
                out.sB[u] += s;
            }
        }
    }
    for (int d = 0; d < C.rank; ++d) out.c_size *= C.dims[d];
    return EIN_SUCCESS;
}

namespace {

struct Digit {
    int64_t e;
    int64_t s[3];
};

// Sort digits outer -> inner by descending key stride, then fold neighbours that are
// contiguous in every operand (outer stride == inner stride * inner extent).
std::vector<Digit> order_and_fold(std::vector<Digit> v, int nops, int key0, int key1) {
    std::stable_sort(v.begin(), v.end(), [&](Digit const &x, Digit const &y) {
        int64_t ax = std::llabs(x.s[key0]), ay = std::llabs(y.s[key0]);
        if (ax != ay) return ax > ay;
        return std::llabs(x.s[key1]) > std::llabs(y.s[key1]);
    });
    std::vector<Digit> out;
    for (auto const &d : v) {
        if (!out.empty()) {
            bool ok = true;
            for (int o = 0; o < nops; ++o) ok = ok && out.back().s[o] == d.s[o] * d.e;
            if (ok) {
                out.back().e *= d.e;
                for (int o = 0; o < nops; ++o) out.back().s[o] = d.s[o];
                continue;
            }
        }
        out.push_back(d);
    }
    return out;
}

int64_t total_of(std::vector<Digit> const &v) {
    int64_t t = 1;
    for (auto const &d : v) t *= d.e;
    return t;
}

int fill_group(IndexGroup &G, std::vector<Digit> const &v, int nops) {
    if (int(v.size()) > GMAX) return fail(EIN_ERROR_UNSUPPORTED, "contraction: too many index digits in one group");
    G.n     = int32_t(v.size());
    int64_t t = total_of(v);
    G.total = uint32_t(t < (int64_t{1} << 31) ? t : 0);
    for (int d = 0; d < GMAX; ++d) {
        G.ext[d] = 1;
        G.fd[d]  = make_fastdiv(1);
        for (int o = 0; o < 3; ++o) G.s[o][d] = 0;
    }
    for (int d = 0; d < G.n; ++d) {
        if (v[d].e >= (int64_t{1} << 31)) return fail(EIN_ERROR_UNSUPPORTED, "contraction: extent above 2^31");
        G.ext[d] = uint32_t(v[d].e);
        G.fd[d]  = make_fastdiv(G.ext[d]);
        for (int o = 0; o < nops; ++o) G.s[o][d] = v[d].s[o];
    }
    return EIN_SUCCESS;
}

} // namespace

// ====================================================== MFMA gather GEMM

// Marks a row/column outside the problem (offsets themselves may be negative).
constexpr int64_t NO_ROW = INT64_MIN;

// One tile (or the slab range [s0, s1) of it) of a gather-GEMM problem. `As`/`Bs` are
// the workgroup's double-buffered LDS slabs (layouts of tile_engine.hpp; AKG / BKG pick
// the row-major (KROW) layout for an operand whose unit stride runs along k); with
// `piece` != nullptr the raw accumulators go there ([BN][BM], i fastest) instead of
// through the epilogue into C.
template <typename T, int BM, int BN, int BK, bool AKG, bool AW, bool BKG, bool BW>
__device__ __forceinline__ void gather_segment(GatherProblem<T> const &P, uint64_t t, uint64_t s0, uint64_t s1, T *As,
                                               T *Bs, T *piece) {
    using Cfg       = TileCfg<T, BM, BN, BK>;
    using LA        = SlabLayout<T, BM, BK, AKG, AW>;
    using LB        = SlabLayout<T, BN, BK, BKG, BW>;
    int const tid   = threadIdx.x;
    int const lane  = tid & 63;
    int const wave  = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const wm    = wave & 1;
    int const wn    = wave >> 1;
    int const r     = lane & 15;
    int const q     = lane >> 4;
    uint32_t const tiles_mn = P.tiles_m * P.tiles_n;

    uint32_t const bz = uint32_t(t / tiles_mn);
    uint32_t tm, tn;
    tile_coords(uint32_t(t % tiles_mn), P.tiles_m, P.tiles_n, tm, tn);
    int64_t const i0   = int64_t(tm) * BM;
    int64_t const j0   = int64_t(tn) * BN;
    int64_t const kbeg = int64_t(s0) * BK;
    int64_t const kend = int64_t(s1) * BK < P.k ? int64_t(s1) * BK : P.k;

    int64_t boff[3];
    group_decode<3>(bz, P.gb, boff);
    T const *__restrict__ Ab = P.A + boff[0];
    T const *__restrict__ Bb = P.B + boff[1];

    // element offsets from the per-call tables; with k_affine one (wave-uniform) table
    // read per slab gives all 16 contraction offsets
    int64_t kslab = 0, kbA = 0, kbB = 0;
    auto set_slab = [&](int64_t k0) {
        kslab = k0;
        if (P.k_affine && k0 < kend) {
            kbA = P.kA[k0];
            kbB = P.kB[k0];
        }
    };
    // with `check` false (interior tile and slab) the table reads need no guard
    auto offA = [&](int row, int64_t k0, int k, bool check, bool &valid) -> int64_t {
        int64_t const i = i0 + row, p = k0 + k;
        valid           = i < P.m && p < kend;
        if (check && !valid) return 0;
        return P.rowA[i] + (P.k_affine ? kbA + (p - kslab) * P.kA_in : P.kA[p]);
    };
    auto offB = [&](int row, int64_t k0, int k, bool check, bool &valid) -> int64_t {
        int64_t const j = j0 + row, p = k0 + k;
        valid           = j < P.n && p < kend;
        if (check && !valid) return 0;
        return P.colB[j] + (P.k_affine ? kbB + (p - kslab) * P.kB_in : P.kB[p]);
    };
    constexpr bool dword_krow = (AKG && !AW) || (BKG && !BW);
    bool const full_mn        = !dword_krow && i0 + BM <= P.m && j0 + BN <= P.n;
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
    __syncthreads(); // LDS of the previous segment no longer read
    set_slab(kbeg);
    dma_both(As, Bs, kbeg);
    for (int64_t sl = 0; sl < nslabs; ++sl) {
        int const buf = int(sl & 1);
        dma_wait_all();
        auto next_dma = [&](T *An, T *Bn) {
            if (sl + 1 < nslabs) {
                set_slab(kbeg + (sl + 1) * BK);
                dma_both(An, Bn, kbeg + (sl + 1) * BK);
            }
        };
        mfma_slab_lds<T, BM, BN, BK, AKG, AW, BKG, BW>(As + buf * LA::SIZE, Bs + buf * LB::SIZE, As + (buf ^ 1) * LA::SIZE,
                                                       Bs + (buf ^ 1) * LB::SIZE, acc, wm, wn, lane, next_dma);
    }

    // lane coordinates made opaque so the store addresses are not hoisted out of the
    // caller's tile loop
    int ri = wm * Cfg::WM + r, rj = wn * Cfg::WN;
    asm volatile("" : "+v"(ri), "+v"(rj));
    if (piece != nullptr) {
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
            for (int b = 0; b < Cfg::FN; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    int ii = ri + a * 16;
                    int jj = rj + b * 16 + Mfma<T>::row(q, v);
                    piece[jj * BM + ii] = acc.c[a][b][v];
                }
        return;
    }
    T *__restrict__ Cb = P.C + boff[2];
    int64_t rowC[Cfg::FM];
#pragma unroll
    for (int a = 0; a < Cfg::FM; ++a) {
        int64_t i = i0 + ri + a * 16;
        rowC[a]   = i < P.m ? P.rowC[i] : NO_ROW;
    }
    // C reads of one column block are issued together (one latency per block)
#pragma unroll
    for (int b = 0; b < Cfg::FN; ++b) {
        int64_t colj[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int64_t j = j0 + rj + b * 16 + Mfma<T>::row(q, v);
            colj[v]   = j < P.n ? P.colC[j] : NO_ROW;
        }
        T cv[Cfg::FM][4];
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                cv[a][v] = (P.beta != T(0) && rowC[a] != NO_ROW && colj[v] != NO_ROW) ? Cb[rowC[a] + colj[v]] : T(0);
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (rowC[a] != NO_ROW && colj[v] != NO_ROW)
                    Cb[rowC[a] + colj[v]] = P.alpha * acc.c[a][b][v] + P.beta * cv[a][v];
    }
}

template <typename T, int BM, int BN, int BK, bool AKG, bool AW, bool BKG, bool BW>
__global__ __launch_bounds__(GEMM_THREADS, 2) void gemm_gather_kernel(GatherProblem<T> P) {
    __shared__ __attribute__((aligned(16))) T As[2 * SlabLayout<T, BM, BK, AKG, AW>::SIZE];
    __shared__ __attribute__((aligned(16))) T Bs[2 * SlabLayout<T, BN, BK, BKG, BW>::SIZE];
    StreamK const &sk = P.sk;
    uint64_t const S  = sk.slabs;
    uint64_t u0, u1, wlog;
    sk_range(sk, u0, u1, wlog);
    for (uint64_t u = u0; u < u1;) {
        uint64_t const t  = u / S;
        uint64_t const s0 = u - t * S;
        uint64_t const s1 = (S - s0 < u1 - u) ? S : s0 + (u1 - u);
        u += s1 - s0;
        T *piece = (s0 == 0 && s1 == S) ? nullptr : P.partial + sk_slot(sk, wlog, t) * (BM * BN);
        gather_segment<T, BM, BN, BK, AKG, AW, BKG, BW>(P, t, s0, s1, As, Bs, piece);
    }
}

// Stream-K fix-up for the gather GEMM (launch shape as the dense fix-up): sums the pieces
// of every cut tile in workgroup order and writes C through the row/column offset tables.
template <typename T, int BM, int BN>
__global__ __launch_bounds__(256) void gather_streamk_fixup(GatherProblem<T> P) {
    StreamK const &sk = P.sk;
    uint64_t const t  = uint64_t(sk.dp_tiles) + blockIdx.y;
    uint64_t g0, g1;
    sk_owners(sk, t, g0, g1);
    if (g0 == g1) return;
    int const e0     = int(blockIdx.x) * 256 + int(threadIdx.x);
    int const stride = int(gridDim.x) * 256;
    T sum[FIXUP_EPT];
#pragma unroll
    for (int x = 0; x < FIXUP_EPT; ++x) sum[x] = T(0);
    for (uint64_t g = g0; g <= g1; ++g) {
        T const *__restrict__ pc = P.partial + sk_slot(sk, g, t) * (BM * BN);
#pragma unroll
        for (int x = 0; x < FIXUP_EPT; ++x) sum[x] += pc[e0 + x * stride];
    }
    uint32_t const tiles_mn = P.tiles_m * P.tiles_n;
    uint32_t const bz       = uint32_t(t / tiles_mn);
    uint32_t tm, tn;
    tile_coords(uint32_t(t % tiles_mn), P.tiles_m, P.tiles_n, tm, tn);
    int64_t const i0 = int64_t(tm) * BM, j0 = int64_t(tn) * BN;
    int64_t ob[3];
    group_decode<3>(bz, P.gb, ob);
#pragma unroll
    for (int x = 0; x < FIXUP_EPT; ++x) {
        int const e = e0 + x * stride;
        int64_t i = i0 + e % BM, j = j0 + e / BM;
        if (i >= P.m || j >= P.n) continue;
        T *cp = P.C + ob[2] + P.rowC[i] + P.colC[j];
        T val = P.alpha * sum[x];
        if (P.beta != T(0)) val += P.beta * *cp;
        *cp = val;
    }
}

// ================================================ reduction regimes

// Target group digits carry (A, B, C) strides; K digits carry (A, B).
template <typename T>
struct ReduceProblem {
    IndexGroup gt, gk;
    int64_t nt, nk;
    T const *A;
    T const *B;
    T *C;
    T alpha, beta;
    int32_t splits; // K pieces per target: 1, or 4 * parts
    int32_t parts;  // partials per target written for the finish pass
    int64_t k_chunk;
    T *partial;
};

// One thread per target point; the K loop walks the digits as an odometer.
template <typename T>
__global__ __launch_bounds__(256) void reduce_thread_kernel(ReduceProblem<T> P) {
    int64_t const stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < P.nt; t += stride) {
        int64_t o[3];
        group_decode<3>(uint32_t(t), P.gt, o);
        T acc = T(0);
        if (P.nk > 0) {
            uint32_t dig[GMAX];
#pragma unroll
            for (int d = 0; d < GMAX; ++d) dig[d] = 0;
            int64_t oa = o[0], ob = o[1];
            int const nd = P.gk.n;
            for (int64_t p = 0; p < P.nk; ++p) {
                acc += P.A[oa] * P.B[ob];
                // advance the odometer (innermost digit last)
                for (int d = nd - 1; d >= 0; --d) {
                    oa += P.gk.s[0][d];
                    ob += P.gk.s[1][d];
                    if (++dig[d] < P.gk.ext[d]) break;
                    oa -= int64_t(P.gk.ext[d]) * P.gk.s[0][d];
                    ob -= int64_t(P.gk.ext[d]) * P.gk.s[1][d];
                    dig[d] = 0;
                }
            }
        }
        T *cp = P.C + o[2];
        T val = P.alpha * acc;
        if (P.beta != T(0)) val += P.beta * *cp;
        *cp = val;
    }
}

// Waves over (target, K piece). splits == 1: one wave per target, lanes stride its K range.
// splits = 4*parts: the four waves of a block take consecutive contiguous pieces of one
// target's K range and add their sums in LDS in fixed order; with parts > 1 each block
// writes a partial that the finish pass sums in fixed order. Cutting K this way keeps many
// short waves in flight, so the grid has no long low-occupancy tail. VEC: the K index
// counts 16-byte vectors of the innermost link digit, unit-stride in both operands, and
// every offset is a multiple of the vector width.
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void reduce_wave_kernel(ReduceProblem<T> P) {
    __shared__ T red[4];
    int const lane      = threadIdx.x & 63;
    int const wave      = threadIdx.x >> 6;
    int64_t const w     = int64_t(blockIdx.x) * 4 + wave;
    int64_t const nwork = P.nt * P.splits;
    // with splits a multiple of 4 the grid holds exactly nwork waves, so only splits == 1
    // (which uses no barrier) returns here
    if (w >= nwork) return;
    int64_t const t  = w / P.splits;
    int64_t const sp = w - t * P.splits;
    int64_t o[3];
    group_decode<3>(uint32_t(t), P.gt, o);
    int64_t const kb = sp * P.k_chunk;
    int64_t const ke = kb + P.k_chunk < P.nk ? kb + P.k_chunk : P.nk;
    T acc[4]         = {T(0), T(0), T(0), T(0)};
    constexpr int VW = VEC ? 16 / int(sizeof(T)) : 1;
    struct alignas(VEC ? 16 : sizeof(T)) Vec {
        T v[VW];
    };
    auto ld   = [](T const *ptr) { return *reinterpret_cast<Vec const *>(ptr); };
    auto dotv = [](Vec const &a, Vec const &b) {
        T r = a.v[0] * b.v[0];
#pragma unroll
        for (int i = 1; i < VW; ++i) r += a.v[i] * b.v[i];
        return r;
    };
    int64_t p = kb + lane;
    for (; p + 192 < ke; p += 256) {
        int64_t k[4][2];
#pragma unroll
        for (int u = 0; u < 4; ++u) group_decode<2>(uint32_t(p + 64 * u), P.gk, k[u]);
        Vec a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a[u] = ld(P.A + o[0] + k[u][0]);
            b[u] = ld(P.B + o[1] + k[u][1]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] += dotv(a[u], b[u]);
    }
    for (; p < ke; p += 64) {
        int64_t k0[2];
        group_decode<2>(uint32_t(p), P.gk, k0);
        acc[0] += dotv(ld(P.A + o[0] + k0[0]), ld(P.B + o[1] + k0[1]));
    }
    T s = wave_reduce_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
    if (P.splits > 1) {
        if (lane == 0) red[wave] = s;
        __syncthreads();
        if (threadIdx.x != 0) return;
        s = (red[0] + red[1]) + (red[2] + red[3]);
        if (P.parts > 1) {
            P.partial[t * P.parts + sp / 4] = s;
            return;
        }
    } else if (lane != 0) {
        return;
    }
    T *cp = P.C + o[2];
    T val = P.alpha * s;
    if (P.beta != T(0)) val += P.beta * *cp;
    *cp = val;
}

template <typename T>
__global__ __launch_bounds__(256) void reduce_finish_targets_kernel(ReduceProblem<T> P) {
    int64_t const stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < P.nt; t += stride) {
        T s = T(0);
        for (int sp = 0; sp < P.parts; ++sp) s += P.partial[t * P.parts + sp];
        int64_t o[3];
        group_decode<3>(uint32_t(t), P.gt, o);
        T *cp = P.C + o[2];
        T val = P.alpha * s;
        if (P.beta != T(0)) val += P.beta * *cp;
        *cp = val;
    }
}

// ================================================================ driver

namespace {

template <typename T, int BM, int BN, bool AKG, bool AW, bool BKG, bool BW>
void const *gather_kernel_ptr() {
    constexpr int BK = 16;
    if constexpr (SlabLayout<T, BM, BK, AKG, AW>::VALID && SlabLayout<T, BN, BK, BKG, BW>::VALID)
        return reinterpret_cast<void const *>(&gemm_gather_kernel<T, BM, BN, BK, AKG, AW, BKG, BW>);
    else
        return nullptr;
}

template <typename T, int BM, int BN>
void const *gather_kernel_for(OperandStaging sa, OperandStaging sb) {
    switch ((sa.krow ? 8 : 0) | (sa.wide ? 4 : 0) | (sb.krow ? 2 : 0) | (sb.wide ? 1 : 0)) {
    case 0: return gather_kernel_ptr<T, BM, BN, false, false, false, false>();
    case 1: return gather_kernel_ptr<T, BM, BN, false, false, false, true>();
    case 2: return gather_kernel_ptr<T, BM, BN, false, false, true, false>();
    case 3: return gather_kernel_ptr<T, BM, BN, false, false, true, true>();
    case 4: return gather_kernel_ptr<T, BM, BN, false, true, false, false>();
    case 5: return gather_kernel_ptr<T, BM, BN, false, true, false, true>();
    case 6: return gather_kernel_ptr<T, BM, BN, false, true, true, false>();
    case 7: return gather_kernel_ptr<T, BM, BN, false, true, true, true>();
    case 8: return gather_kernel_ptr<T, BM, BN, true, false, false, false>();
    case 9: return gather_kernel_ptr<T, BM, BN, true, false, false, true>();
    case 10: return gather_kernel_ptr<T, BM, BN, true, false, true, false>();
    case 11: return gather_kernel_ptr<T, BM, BN, true, false, true, true>();
    case 12: return gather_kernel_ptr<T, BM, BN, true, true, false, false>();
    case 13: return gather_kernel_ptr<T, BM, BN, true, true, false, true>();
    case 14: return gather_kernel_ptr<T, BM, BN, true, true, true, false>();
    default: return gather_kernel_ptr<T, BM, BN, true, true, true, true>();
    }
}

// 16-byte DMA for one gathered operand: the chunk runs along the innermost digit of `fast`
// (operand index fop), which must have stride +1 and an extent that is a multiple of the
// chunk; every other stride of the operand must keep chunk starts 16-byte aligned.
template <typename T>
bool gather_wide_ok(std::vector<Digit> const &fast, int fop, std::vector<Digit> const &other, int oop,
                    std::vector<Digit> const &batch, int bop, void const *ptr) {
    constexpr int64_t KG = 16 / int64_t(sizeof(T));
    if (fast.empty() || fast.back().s[fop] != 1 || fast.back().e % KG != 0) return false;
    for (size_t d = 0; d + 1 < fast.size(); ++d)
        if (fast[d].s[fop] % KG != 0) return false;
    for (auto const &d : other)
        if (d.s[oop] % KG != 0) return false;
    for (auto const &d : batch)
        if (d.s[bop] % KG != 0) return false;
    return reinterpret_cast<uintptr_t>(ptr) % 16 == 0;
}

__device__ __forceinline__ void fill_offsets(IndexGroup const &G, int64_t total, int64_t *__restrict__ o0,
                                             int64_t *__restrict__ o1) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < total; i += int64_t(gridDim.x) * 256) {
        int64_t o[2];
        group_decode<2>(uint32_t(i), G, o);
        o0[i] = o[0];
        o1[i] = o[1];
    }
}

// Element offsets of every index of the M, N and K groups (blockIdx.y selects the group),
// laid out as gather_table_entries describes.
__global__ __launch_bounds__(256) void group_tables_kernel(IndexGroup gm, IndexGroup gn, IndexGroup gk, int64_t m,
                                                           int64_t n, int64_t k, int64_t *__restrict__ tab) {
    if (blockIdx.y == 0)
        fill_offsets(gm, m, tab, tab + m);
    else if (blockIdx.y == 1)
        fill_offsets(gn, n, tab + 2 * m, tab + 2 * m + n);
    else
        fill_offsets(gk, k, tab + 2 * (m + n), tab + 2 * (m + n) + k);
}

// Host version of the same tables (grouped plans build them once at creation).
void host_offsets(IndexGroup const &G, int64_t total, int64_t *o0, int64_t *o1) {
    for (int64_t i = 0; i < total; ++i) {
        uint64_t idx = uint64_t(i);
        int64_t a = 0, b = 0;
        for (int d = G.n - 1; d >= 0; --d) {
            uint64_t const r = idx % G.ext[d];
            idx /= G.ext[d];
            a += int64_t(r) * G.s[0][d];
            b += int64_t(r) * G.s[1][d];
        }
        o0[i] = a;
        o1[i] = b;
    }
}

// Within a 16-step slab the contraction offsets are affine when the innermost K digit's
// extent is a multiple of 16 (slabs start at multiples of 16 and never cross that digit).
template <typename T>
void set_k_affine(GatherProblem<T> &P, std::vector<Digit> const &gk) {
    P.k_affine = (!gk.empty() && gk.back().e % 16 == 0) ? 1 : 0;
    P.kA_in    = gk.empty() ? 0 : gk.back().s[0];
    P.kB_in    = gk.empty() ? 0 : gk.back().s[1];
}

template <typename T>
void set_tables(GatherProblem<T> &P, int64_t const *tab) {
    P.rowA = tab;
    P.rowC = tab + P.m;
    P.colB = tab + 2 * P.m;
    P.colC = tab + 2 * P.m + P.n;
    P.kA   = tab + 2 * (P.m + P.n);
    P.kB   = tab + 2 * (P.m + P.n) + P.k;
}

} // namespace

namespace {

// ====================================================== transpose-transpose-GEMM route

thread_local int32_t t_route = EIN_ROUTE_AUTO;

// Outer -> inner order of the digits by descending |stride| in operand slot o.
std::vector<Digit> sorted_by(std::vector<Digit> v, int o) {
    std::stable_sort(v.begin(), v.end(),
                     [&](Digit const &x, Digit const &y) { return std::llabs(x.s[o]) > std::llabs(y.s[o]); });
    return v;
}

// The digits (outer -> inner) address one axis of operand slot o with a single positive
// stride: each stride is the next one's times the next extent. No digits: extent 1.
bool folds_to_axis(std::vector<Digit> const &v, int o, int64_t &stride) {
    stride = 0;
    if (v.empty()) return true;
    for (size_t d = 0; d + 1 < v.size(); ++d)
        if (v[d].s[o] != v[d + 1].s[o] * v[d + 1].e) return false;
    stride = v.back().s[o];
    return stride > 0;
}

// dst (dense, the listed digits outermost first) = src, src addressed through slot o of
// each group's digits.
template <typename T>
int pack_digits(T const *src, std::vector<std::pair<std::vector<Digit> const *, int>> const &groups, T *dst,
                hipStream_t stream) {
    std::vector<int64_t> dims, ss, sd;
    for (auto const &g : groups)
        for (auto const &d : *g.first) {
            dims.push_back(d.e);
            ss.push_back(d.s[g.second]);
        }
    if (dims.empty()) {
        dims.push_back(1);
        ss.push_back(1);
    }
    sd.resize(dims.size());
    int64_t pv = 1;
    for (size_t i = dims.size(); i-- > 0;) {
        sd[i] = pv;
        pv *= dims[i];
    }
    std::vector<int32_t> perm(dims.size());
    std::iota(perm.begin(), perm.end(), 0);
    return permute_tensor<T>(int32_t(dims.size()), dims.data(), perm.data(), T(1), src, ss.data(), T(0), dst, sd.data(),
                             stream);
}

// Transpose-transpose-GEMM: an operand whose digits do not fold into one strided matrix is
// packed by the LDS-tiled permute into [batch][rows][k] (k fastest), the product runs on
// the dense MFMA GEMM (or the split-precision GEMM when that mode is on), and an output
// that does not fold is formed in a dense buffer X ([batch][N][M], M fastest) and added
// into C by the permute. Digit orders are chosen to fold as much as possible. The route
// is taken (`taken` = true) when its extra traffic and launches cost less than the
// measured gap between the gather GEMM and the dense GEMM (the gather kernel takes about
// 1.45x the dense kernel's time at 2304^3 FP64 on MI355X).
template <typename T>
int contract_ttgt(std::vector<Digit> const &vb, std::vector<Digit> const &vm, std::vector<Digit> const &vn,
                  std::vector<Digit> const &vk, T alpha, T const *a, T const *b, T beta, T *c, hipStream_t stream,
                  bool &taken) {
    taken = false;
    if (t_route == EIN_ROUTE_GATHER) return EIN_SUCCESS;
    bool const forced = t_route == EIN_ROUTE_TTGT;
    if (vb.size() + vm.size() + vk.size() > size_t(EIN_MAX_RANK) || vb.size() + vn.size() + vk.size() > size_t(EIN_MAX_RANK) ||
        vb.size() + vm.size() + vn.size() > size_t(EIN_MAX_RANK))
        return EIN_SUCCESS;
    int64_t const M = total_of(vm), N = total_of(vn), K = total_of(vk), NB = total_of(vb);
    struct Choice {
        std::vector<Digit> m, n, k, bt;
        bool a_ok, b_ok, c_ok;
        int64_t as[3], bs[3], cs[3];
        double moved; // elements written and read again by packing / staging
    };
    Choice best{};
    best.moved = -1.0;
    std::vector<Digit> const bt = sorted_by(vb, 2);
    for (int om = 0; om < 2; ++om)         // M digits ordered by A (0) or C (1) strides
        for (int on = 0; on < 2; ++on)     // N digits by B (0) or C (1)
            for (int ok = 0; ok < 2; ++ok) { // K digits by A (0) or B (1)
                Choice ch;
                ch.bt   = bt;
                ch.m    = sorted_by(vm, om);
                ch.n    = sorted_by(vn, on);
                ch.k    = sorted_by(vk, ok);
                ch.a_ok = folds_to_axis(ch.bt, 0, ch.as[0]) && folds_to_axis(ch.m, 0, ch.as[1]) &&
                          folds_to_axis(ch.k, 0, ch.as[2]);
                ch.b_ok = folds_to_axis(ch.bt, 1, ch.bs[0]) && folds_to_axis(ch.n, 0, ch.bs[1]) &&
                          folds_to_axis(ch.k, 1, ch.bs[2]);
                ch.c_ok = folds_to_axis(ch.bt, 2, ch.cs[0]) && folds_to_axis(ch.m, 1, ch.cs[1]) &&
                          folds_to_axis(ch.n, 1, ch.cs[2]);
                ch.moved = (ch.a_ok ? 0.0 : double(NB) * M * K) + (ch.b_ok ? 0.0 : double(NB) * N * K) +
                           (ch.c_ok ? 0.0 : double(NB) * M * N);
                if (best.moved < 0.0 || ch.moved < best.moved) best = ch;
            }
    int const launches   = int(!best.a_ok) + int(!best.b_ok) + int(!best.c_ok);
    double const t_gemm  = 2.0 * double(NB) * M * N * K / (sizeof(T) == 8 ? 55e12 : 100e12);
    double const t_extra = 2.0 * sizeof(T) * best.moved / 4e12 + 5e-6 * launches;
    if (!forced && t_extra >= 0.45 * t_gemm) return EIN_SUCCESS;

    auto align = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t const a_bytes = best.a_ok ? 0 : align(size_t(NB * M * K) * sizeof(T));
    size_t const b_bytes = best.b_ok ? 0 : align(size_t(NB * N * K) * sizeof(T));
    size_t const x_bytes = best.c_ok ? 0 : align(size_t(NB * M * N) * sizeof(T));
    size_t const total   = a_bytes + b_bytes + x_bytes;
    char *ws             = nullptr;
    if (total > 0) {
        // under stream capture the buffer must already exist; otherwise the gather runs
        if (!forced && stream_capturing(stream) && !workspace_fits(stream, 1, total)) return EIN_SUCCESS;
        void *p = nullptr;
        EIN_TRY(workspace_slot(stream, 1, total, &p));
        ws = static_cast<char *>(p);
    }
    taken = true;
    GemmProblem<T> P{};
    P.m     = M;
    P.n     = N;
    P.k     = K;
    P.batch = NB;
    P.alpha = alpha;
    if (best.a_ok) {
        P.A    = a;
        P.a_sb = best.as[0];
        P.a_sm = best.as[1];
        P.a_sk = best.as[2];
    } else {
        T *pa = reinterpret_cast<T *>(ws);
        EIN_TRY(pack_digits<T>(a, {{&best.bt, 0}, {&best.m, 0}, {&best.k, 0}}, pa, stream));
        P.A    = pa;
        P.a_sb = M * K;
        P.a_sm = K;
        P.a_sk = 1;
    }
    if (best.b_ok) {
        P.B    = b;
        P.b_sb = best.bs[0];
        P.b_sn = best.bs[1];
        P.b_sk = best.bs[2];
    } else {
        T *pb = reinterpret_cast<T *>(ws + a_bytes);
        EIN_TRY(pack_digits<T>(b, {{&best.bt, 1}, {&best.n, 0}, {&best.k, 1}}, pb, stream));
        P.B    = pb;
        P.b_sb = N * K;
        P.b_sn = K;
        P.b_sk = 1;
    }
    T *X = best.c_ok ? nullptr : reinterpret_cast<T *>(ws + a_bytes + b_bytes);
    if (best.c_ok) {
        P.C    = c;
        P.c_sb = best.cs[0];
        P.c_sm = best.cs[1];
        P.c_sn = best.cs[2];
        P.beta = beta;
    } else {
        P.C    = X;
        P.c_sb = M * N;
        P.c_sm = 1;
        P.c_sn = M;
        P.beta = T(0);
    }
    EIN_TRY(gemm_dispatch<T>(P, stream));
    if (!best.c_ok) {
        // C = beta*C + X over the (batch, N, M) digits
        std::vector<int64_t> dims, sx, sc;
        int64_t pv = 1;
        std::vector<std::pair<std::vector<Digit> const *, int>> const order = {{&best.bt, 2}, {&best.n, 1}, {&best.m, 1}};
        for (auto const &g : order)
            for (auto const &d : *g.first) {
                dims.push_back(d.e);
                sc.push_back(d.s[g.second]);
            }
        sx.resize(dims.size());
        for (size_t i = dims.size(); i-- > 0;) {
            sx[i] = pv;
            pv *= dims[i];
        }
        std::vector<int32_t> perm(dims.size());
        std::iota(perm.begin(), perm.end(), 0);
        EIN_TRY(permute_tensor<T>(int32_t(dims.size()), dims.data(), perm.data(), T(1), X, sx.data(), beta, c, sc.data(),
                                  stream));
    }
    return EIN_SUCCESS;
}

} // namespace

template <typename T>
int contract_generic(T c_prefactor, ein_tensor_desc const &C, T *c, T ab_prefactor, ein_tensor_desc const &A, T const *a,
                     ein_tensor_desc const &B, T const *b, hipStream_t stream) {
    ContractionAnalysis an;
    EIN_TRY(analyse_contraction(C, A, B, an));
    if (an.c_size == 0) return EIN_SUCCESS;

    T beta = c_prefactor;
    if (an.c_repeats) {
        // C addresses a diagonal: the reference scales/zeroes all of C first, then
        // accumulates into the diagonal with prefactor 1.
        EIN_TRY(tensor_scale<T>(C, c, c_prefactor, stream));
        beta = T(1);
    }

    std::vector<Digit> vb, vm, vn, vk;
    bool empty_k = false;
    for (int u = 0; u < an.nlab; ++u) {
        int64_t e = an.ext[u];
        if (e == 0) {
            if (!an.inC[u]) empty_k = true;
            continue;
        }
        if (e == 1) continue;
        if (an.inC[u] && an.inA[u] && an.inB[u])
            vb.push_back({e, {an.sA[u], an.sB[u], an.sC[u]}});
        else if (an.inC[u] && !an.inB[u])
            vm.push_back({e, {an.sA[u], an.sC[u], 0}});
        else if (an.inC[u] && an.inB[u])
            vn.push_back({e, {an.sB[u], an.sC[u], 0}});
        else
            vk.push_back({e, {an.sA[u], an.sB[u], 0}});
    }
    if (empty_k || ab_prefactor == T(0)) {
        // nothing to accumulate: C = beta*C over the target points
        if (!an.c_repeats) EIN_TRY(tensor_scale<T>(C, c, c_prefactor, stream));
        return EIN_SUCCESS;
    }

    int64_t const M = total_of(vm), N = total_of(vn), K = total_of(vk), NB = total_of(vb);
    bool const use_mfma = M >= 16 && N >= 16 && K >= 16 && M < (int64_t{1} << 31) && N < (int64_t{1} << 31) &&
                          K < (int64_t{1} << 31) && NB <= 65535;

    if (use_mfma) {
        bool taken = false;
        EIN_TRY(contract_ttgt<T>(vb, vm, vn, vk, ab_prefactor, a, b, beta, c, stream, taken));
        if (taken) return EIN_SUCCESS;
        auto gm = order_and_fold(vm, 2, 0, 1);
        auto gn = order_and_fold(vn, 2, 0, 1);
        bool const a_i_fast = !gm.empty() && std::llabs(gm.back().s[0]) == 1;
        // K order serves the operand that cannot use its row/column direction
        auto gk = a_i_fast ? order_and_fold(vk, 2, 1, 0) : order_and_fold(vk, 2, 0, 1);
        bool const b_j_fast_n = !gn.empty() && std::llabs(gn.back().s[0]) == 1;
        bool const b_p_unit   = !gk.empty() && std::llabs(gk.back().s[1]) == 1;
        bool const b_j_fast   = b_j_fast_n || !b_p_unit;
        bool const a_fast     = a_i_fast || !(!gk.empty() && std::llabs(gk.back().s[0]) == 1);
        auto gbv              = order_and_fold(vb, 3, 2, 0);
        // The epilogue's lanes run along the innermost M digit. When that digit is far from
        // unit stride in C, the tile is written to a dense buffer X (i fastest, then j, then
        // the batch) and added into C by the LDS-tiled permute, whose reads and writes are
        // both coalesced.
        bool const stage = std::llabs(gm.back().s[1]) >= 16 && int(gm.size() + gn.size() + gbv.size()) <= EIN_MAX_RANK;
        std::vector<Digit> const gm_c = gm, gn_c = gn, gb_c = gbv; // C strides
        if (stage) {
            int64_t pv = 1;
            for (size_t d = gm.size(); d-- > 0;) {
                gm[d].s[1] = pv;
                pv *= gm[d].e;
            }
            for (size_t d = gn.size(); d-- > 0;) {
                gn[d].s[1] = pv;
                pv *= gn[d].e;
            }
            for (size_t d = gbv.size(); d-- > 0;) {
                gbv[d].s[2] = pv;
                pv *= gbv[d].e;
            }
        }
        GatherProblem<T> P{};
        EIN_TRY(fill_group(P.gm, gm, 2));
        EIN_TRY(fill_group(P.gn, gn, 2));
        EIN_TRY(fill_group(P.gk, gk, 2));
        EIN_TRY(fill_group(P.gb, gbv, 3));
        set_k_affine(P, gk);
        P.m     = M;
        P.n     = N;
        P.k     = K;
        P.batch = NB;
        P.A     = a;
        P.B     = b;
        P.C     = c;
        P.alpha = ab_prefactor;
        P.beta  = beta;
        GemmConfig cfg = choose_gemm_config(M, N, K, NB, int(sizeof(T)));
        int const bm   = cfg.bm == 128 ? 128 : 64;
        int const bn   = cfg.bn == 128 ? 128 : 64;
        P.tiles_m      = uint32_t(ceil_div(M, bm));
        P.tiles_n      = uint32_t(ceil_div(N, bn));
        uint64_t tiles   = uint64_t(P.tiles_m) * P.tiles_n * uint64_t(NB);
        if (tiles > 0x7fffffffu) return fail(EIN_ERROR_UNSUPPORTED, "contraction tile grid too large");
        OperandStaging sa{!a_fast, false}, sb{!b_j_fast, false};
        constexpr int EW = int(sizeof(T)) / 4;
        sa.wide = a_fast ? (bm * EW >= 256 && gather_wide_ok<T>(gm, 0, gk, 0, gbv, 0, a))
                         : gather_wide_ok<T>(gk, 0, gm, 0, gbv, 0, a);
        sb.wide = b_j_fast ? (bn * EW >= 256 && gather_wide_ok<T>(gn, 0, gk, 1, gbv, 1, b))
                           : gather_wide_ok<T>(gk, 1, gn, 0, gbv, 1, b);
        void const *kern = bm == 128 && bn == 128 ? gather_kernel_for<T, 128, 128>(sa, sb)
                           : bm == 128            ? gather_kernel_for<T, 128, 64>(sa, sb)
                                                  : gather_kernel_for<T, 64, 64>(sa, sb);
        if (kern == nullptr) return fail(EIN_ERROR_UNSUPPORTED, "contraction: operand staging not realisable for this tile");
        P.sk = plan_stream_k(tiles, uint64_t(ceil_div(K, 16)),
                             uint64_t(num_cus()) * uint64_t(resident_blocks(kern, GEMM_THREADS)));
        // workspace: offset tables, then the stream-K pieces
        size_t const tab_bytes   = (gather_table_entries(M, N, K) * sizeof(int64_t) + 255) / 256 * 256;
        size_t const piece_bytes = (size_t(sk_piece_slots(P.sk)) * size_t(bm) * size_t(bn) * sizeof(T) + 255) / 256 * 256;
        size_t const x_bytes     = stage ? size_t(M) * size_t(N) * size_t(NB) * sizeof(T) : 0;
        void *ws                 = nullptr;
        EIN_TRY(workspace(stream, tab_bytes + piece_bytes + x_bytes, &ws));
        int64_t *tab = static_cast<int64_t *>(ws);
        set_tables(P, tab);
        P.partial = P.sk.units != 0 ? reinterpret_cast<T *>(static_cast<char *>(ws) + tab_bytes) : nullptr;
        T *X      = stage ? reinterpret_cast<T *>(static_cast<char *>(ws) + tab_bytes + piece_bytes) : nullptr;
        if (stage) {
            P.C    = X;
            P.beta = T(0);
        }
        {
            int64_t const mx = std::max(M, std::max(N, K));
            dim3 tgrid{uint32_t(std::min<int64_t>(ceil_div(mx, 256), 4096)), 3};
            group_tables_kernel<<<tgrid, 256, 0, stream>>>(P.gm, P.gn, P.gk, M, N, K, tab);
            EIN_LAUNCH_CHECK();
        }
        void *args[] = {&P};
        EIN_HIP_CHECK(hipLaunchKernel(kern, dim3(P.sk.grid), dim3(GEMM_THREADS), args, 0, stream));
        if (P.sk.units != 0) {
            if (bm == 128 && bn == 128)
                gather_streamk_fixup<T, 128, 128><<<sk_fixup_grid(P.sk, 128 * 128), 256, 0, stream>>>(P);
            else if (bm == 128)
                gather_streamk_fixup<T, 128, 64><<<sk_fixup_grid(P.sk, 128 * 64), 256, 0, stream>>>(P);
            else
                gather_streamk_fixup<T, 64, 64><<<sk_fixup_grid(P.sk, 64 * 64), 256, 0, stream>>>(P);
            EIN_LAUNCH_CHECK();
        }
        if (stage) {
            // C = beta*C + X over the (M, N, batch) digits
            std::vector<int64_t> dims, sx, sc;
            std::vector<int32_t> perm;
            auto add = [&](std::vector<Digit> const &dx, std::vector<Digit> const &dc, int o) {
                for (size_t d = 0; d < dx.size(); ++d) {
                    perm.push_back(int32_t(dims.size()));
                    dims.push_back(dx[d].e);
                    sx.push_back(dx[d].s[o]);
                    sc.push_back(dc[d].s[o]);
                }
            };
            add(gbv, gb_c, 2);
            add(gn, gn_c, 1);
            add(gm, gm_c, 1);
            EIN_TRY(permute_tensor<T>(int32_t(dims.size()), dims.data(), perm.data(), T(1), X, sx.data(), beta, c,
                                      sc.data(), stream));
        }
        return EIN_SUCCESS;
    }

    // ---- reduction regimes: targets = batch + M + N digits, links = K digits
    std::vector<Digit> vt;
    for (auto const &d : vb) vt.push_back({d.e, {d.s[0], d.s[1], d.s[2]}});
    for (auto const &d : vm) vt.push_back({d.e, {d.s[0], 0, d.s[1]}});
    for (auto const &d : vn) vt.push_back({d.e, {0, d.s[0], d.s[1]}});
    int64_t const NT = total_of(vt);
    if (NT >= (int64_t{1} << 31) || K >= (int64_t{1} << 31))
        return fail(EIN_ERROR_UNSUPPORTED, "contraction: index range above 2^31 in the reduction regime");
    // K innermost digit: smallest of the A/B strides
    for (auto &d : vk) d.s[2] = std::min(std::llabs(d.s[0]), std::llabs(d.s[1]));
    auto gk = order_and_fold(vk, 2, 2, 0);
    for (auto &d : gk) d.s[2] = 0;
    bool const k_unit = !gk.empty() && (std::llabs(gk.back().s[0]) == 1 || std::llabs(gk.back().s[1]) == 1);
    // target order: innermost = unit stride in the operand that is read per point
    std::vector<Digit> vt2 = vt;
    std::stable_sort(vt2.begin(), vt2.end(), [](Digit const &x, Digit const &y) {
        auto key = [](Digit const &d) {
            int64_t k = std::llabs(d.s[2]);
            if (d.s[0] != 0) k = std::min<int64_t>(k, std::llabs(d.s[0]));
            if (d.s[1] != 0) k = std::min<int64_t>(k, std::llabs(d.s[1]));
            return k;
        };
        return key(x) > key(y);
    });
    std::vector<Digit> gt;
    for (auto const &d : vt2) {
        if (!gt.empty() && gt.back().s[0] == d.s[0] * d.e && gt.back().s[1] == d.s[1] * d.e &&
            gt.back().s[2] == d.s[2] * d.e) {
            gt.back().e *= d.e;
            for (int o = 0; o < 3; ++o) gt.back().s[o] = d.s[o];
        } else {
            gt.push_back(d);
        }
    }
    ReduceProblem<T> R{};
    EIN_TRY(fill_group(R.gt, gt, 3));
    EIN_TRY(fill_group(R.gk, gk, 2));
    R.nt    = NT;
    R.nk    = K;
    R.A     = a;
    R.B     = b;
    R.C     = c;
    R.alpha = ab_prefactor;
    R.beta  = beta;
    R.splits  = 1;
    R.k_chunk = K;
    R.partial = nullptr;
    int const cus = num_cus();
    bool const thread_mode = K <= 32 || (!k_unit && NT >= 4096);
    if (thread_mode) {
        uint32_t grid = uint32_t(std::max<int64_t>(1, std::min<int64_t>(ceil_div(NT, 256), int64_t(cus) * 16)));
        reduce_thread_kernel<T><<<grid, 256, 0, stream>>>(R);
        EIN_LAUNCH_CHECK();
        return EIN_SUCCESS;
    }
    R.parts = 1;
    // 16-byte vectors along the innermost link digit when it is unit-stride in both operands
    // and every operand offset keeps them aligned; the K index then counts vectors
    constexpr int VW = 16 / int(sizeof(T));
    bool vec = !gk.empty() && gk.back().s[0] == 1 && gk.back().s[1] == 1 && gk.back().e % VW == 0 &&
               (reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) % 16 == 0;
    for (size_t d = 0; vec && d + 1 < gk.size(); ++d) vec = gk[d].s[0] % VW == 0 && gk[d].s[1] % VW == 0;
    for (auto const &d : gt) vec = vec && d.s[0] % VW == 0 && d.s[1] % VW == 0;
    if (vec) {
        auto gkv = gk;
        gkv.back().e /= VW;
        gkv.back().s[0] = VW;
        gkv.back().s[1] = VW;
        EIN_TRY(fill_group(R.gk, gkv, 2));
        R.nk      = K / VW;
        R.k_chunk = R.nk;
    }
    // too few targets for about four rounds of resident waves: cut each target's K range
    // into 4 * parts pieces of >= 256 links; partials (parts > 1, one more launch) only when
    // four waves per target do not even fill the chip once
    int64_t const resident = int64_t(cus) * 32;
    if (NT < 4 * resident && K >= 1024) {
        int64_t s = std::min<int64_t>(ceil_div(4 * resident, NT), K / 256) / 4;
        if (4 * NT >= resident) s = 1;
        s         = std::max<int64_t>(1, std::min<int64_t>(s, 1024));
        R.splits  = int(4 * s);
        R.parts   = int(s);
        R.k_chunk = ceil_div(R.nk, int64_t(R.splits));
    }
    if (R.parts > 1) {
        void *ws = nullptr;
        EIN_TRY(workspace(stream, size_t(NT) * size_t(R.parts) * sizeof(T), &ws));
        R.partial = static_cast<T *>(ws);
    }
    int64_t nwaves = NT * R.splits;
    if (vec)
        reduce_wave_kernel<T, true><<<uint32_t(ceil_div(nwaves, 4)), 256, 0, stream>>>(R);
    else
        reduce_wave_kernel<T, false><<<uint32_t(ceil_div(nwaves, 4)), 256, 0, stream>>>(R);
    EIN_LAUNCH_CHECK();
    if (R.parts > 1) {
        uint32_t grid = uint32_t(std::max<int64_t>(1, std::min<int64_t>(ceil_div(NT, 256), int64_t(cus) * 16)));
        reduce_finish_targets_kernel<T><<<grid, 256, 0, stream>>>(R);
        EIN_LAUNCH_CHECK();
    }
    return EIN_SUCCESS;
}

template int contract_generic<double>(double, ein_tensor_desc const &, double *, double, ein_tensor_desc const &,
                                      double const *, ein_tensor_desc const &, double const *, hipStream_t);
template int contract_generic<float>(float, ein_tensor_desc const &, float *, float, ein_tensor_desc const &, float const *,
                                     ein_tensor_desc const &, float const *, hipStream_t);

} // namespace ein

extern "C" {

int32_t ein_set_contraction_route(int32_t route) {
    if (route != EIN_ROUTE_AUTO && route != EIN_ROUTE_GATHER && route != EIN_ROUTE_TTGT) {
        ein::set_error("set_contraction_route: unknown route");
        return -1;
    }
    int32_t const prev = ein::t_route;
    ein::t_route       = route;
    return prev;
}

int ein_contract_f64(double c_prefactor, const ein_tensor_desc *C, double *c, double ab_prefactor, const ein_tensor_desc *A,
                     const double *a, const ein_tensor_desc *B, const double *b, ein_stream_t stream) {
    if (C == nullptr) return -2;
    if (A == nullptr) return -5;
    if (B == nullptr) return -7;
    return ein::contract_generic<double>(c_prefactor, *C, c, ab_prefactor, *A, a, *B, b, ein::as_stream(stream));
}

int ein_contract_f32(float c_prefactor, const ein_tensor_desc *C, float *c, float ab_prefactor, const ein_tensor_desc *A,
                     const float *a, const ein_tensor_desc *B, const float *b, ein_stream_t stream) {
    if (C == nullptr) return -2;
    if (A == nullptr) return -5;
    if (B == nullptr) return -7;
    return ein::contract_generic<float>(c_prefactor, *C, c, ab_prefactor, *A, a, *B, b, ein::as_stream(stream));
}

} // extern "C"

// ===================================================== grouped launches
//
// A grouped plan holds many independent contractions (the blocks of a block-sparse
// einsum, a batch of uneven GEMMs). Each becomes a gather-GEMM problem; all their tiles
// form one list sorted by contraction length (longest first) and a persistent launch of
// about two workgroups per CU walks the list with a fixed stride -- one launch per wave
// of items whose outputs do not overlap. Descriptors live in device memory, so executing
// a plan is a pure kernel launch and can be captured into a hipGraph.

namespace ein {

template <typename T>
struct GroupedItem {
    GatherProblem<T> P;
    int32_t flags; // bit 0: A walks rows (i-fast), bit 1: B walks columns (j-fast)
    int32_t pad;
};

struct GroupedTile {
    uint32_t item;
    uint32_t tile;
};

template <typename T, int BM>
__global__ __launch_bounds__(GEMM_THREADS, 2) void gemm_grouped_kernel(GroupedItem<T> const *__restrict__ items,
                                                                       GroupedTile const *__restrict__ tiles,
                                                                       uint32_t ntiles) {
    constexpr int BN = 64, BK = 16;
    // the k-major layouts (padded) are the larger ones
    __shared__ __attribute__((aligned(16))) T As[2 * SlabLayout<T, BM, BK, false, false>::SIZE];
    __shared__ __attribute__((aligned(16))) T Bs[2 * SlabLayout<T, BN, BK, false, false>::SIZE];
    static_assert(SlabLayout<T, BM, BK, false, false>::SIZE >= SlabLayout<T, BM, BK, true, false>::SIZE &&
                      SlabLayout<T, BN, BK, false, false>::SIZE >= SlabLayout<T, BN, BK, true, false>::SIZE,
                  "LDS sized for the larger layout");
    for (uint32_t tt = blockIdx.x; tt < ntiles; tt += gridDim.x) {
        GroupedTile const e           = tiles[tt];
        GroupedItem<T> const &it      = items[e.item];
        uint64_t const S              = uint64_t((it.P.k + BK - 1) / BK);
        switch (it.flags) {
        case 3: gather_segment<T, BM, BN, BK, false, false, false, false>(it.P, e.tile, 0, S, As, Bs, nullptr); break;
        case 1: gather_segment<T, BM, BN, BK, false, false, true, false>(it.P, e.tile, 0, S, As, Bs, nullptr); break;
        case 2: gather_segment<T, BM, BN, BK, true, false, false, false>(it.P, e.tile, 0, S, As, Bs, nullptr); break;
        default: gather_segment<T, BM, BN, BK, true, false, true, false>(it.P, e.tile, 0, S, As, Bs, nullptr); break;
        }
    }
}

namespace {

struct PreScale {
    ein_tensor_desc C;
    void *c;
    double factor;
};

struct GroupedWave {
    std::vector<PreScale> prescale;        // C scales that must precede this wave's tiles
    std::vector<ein_contraction_item> large; // items that fill the GPU alone: regular dispatch
    void *d_items[2] = {nullptr, nullptr}; // per tile height (64, 128)
    void *d_tiles[2] = {nullptr, nullptr};
    void *d_tables   = nullptr;            // offset tables of the wave's items
    uint32_t ntiles[2] = {0, 0};
};

// Returns a wave's device buffers to the pool. With `stream` known (the only stream the
// plan ran on) the blocks are reused once that stream's queued work is done; otherwise only
// after a device synchronisation.
void free_wave(GroupedWave &w, hipStream_t stream = nullptr, bool known = false) {
    auto release = [&](void *p) {
        if (known) (void)ein_free_async(p, stream);
        else (void)ein_free(p);
    };
    for (int v = 0; v < 2; ++v) {
        release(w.d_items[v]);
        release(w.d_tiles[v]);
        w.d_items[v] = w.d_tiles[v] = nullptr;
    }
    release(w.d_tables);
    w.d_tables = nullptr;
}

struct GroupedPlan {
    int elem = 8;
    int bm_index = 0; // 0: 64-row tiles, 1: 128-row tiles
    bool tuned = false;
    std::vector<GroupedWave> waves;
    std::vector<ein_contraction_item> items; // host copy (for the autotuner)
    std::vector<int> wave_of;                // wave of each kept item
    int64_t tiles_total[2] = {0, 0};
    std::vector<hipStream_t> streams;        // streams the plan was executed on
};

// Elements a descriptor addresses, and how far below its base pointer the negative strides
// reach (both in elements): a scratch buffer of `span` elements serves it at base + below.
int64_t span_of(ein_tensor_desc const &d, int64_t *below = nullptr) {
    int64_t hi = 0, lo = 0;
    for (int i = 0; i < d.rank; ++i) {
        if (d.dims[i] <= 0) continue;
        int64_t const reach = (d.dims[i] - 1) * d.strides[i];
        if (reach < 0) lo += reach;
        else hi += reach;
    }
    if (below != nullptr) *below = -lo;
    return hi - lo + 1;
}

// Byte range [lo, hi) a descriptor addresses from base p (strides of either sign).
struct ByteRange {
    uintptr_t lo, hi;
    bool overlaps(ByteRange const &o) const { return lo < o.hi && o.lo < hi; }
};

ByteRange range_of(ein_tensor_desc const &d, void const *p, size_t elem) {
    int64_t lo = 0, hi = 0;
    for (int i = 0; i < d.rank; ++i) {
        if (d.dims[i] == 0) return {0, 0};
        int64_t reach = (d.dims[i] - 1) * d.strides[i];
        if (reach < 0)
            lo += reach;
        else
            hi += reach;
    }
    auto const base = reinterpret_cast<uintptr_t>(p);
    return {base + uintptr_t(lo * int64_t(elem)), base + uintptr_t((hi + 1) * int64_t(elem))};
}

// Analyse one item into a gather problem (MFMA path for every size). kind: 0 = tiles to
// run, 1 = only a scale of C, 2 = nothing to do.
template <typename T>
int build_item(ein_contraction_item const &in, GroupedItem<T> (&out)[2], int &kind, bool &c_repeats) {
    ContractionAnalysis an;
    EIN_TRY(analyse_contraction(in.C, in.A, in.B, an));
    c_repeats = an.c_repeats;
    kind      = 0;
    if (an.c_size == 0) {
        kind = 2;
        return EIN_SUCCESS;
    }
    std::vector<Digit> vb, vm, vn, vk;
    bool empty_k = false;
    for (int u = 0; u < an.nlab; ++u) {
        int64_t e = an.ext[u];
        if (e == 0) {
            if (!an.inC[u]) empty_k = true;
            continue;
        }
        if (e == 1) continue;
        if (an.inC[u] && an.inA[u] && an.inB[u])
            vb.push_back({e, {an.sA[u], an.sB[u], an.sC[u]}});
        else if (an.inC[u] && !an.inB[u])
            vm.push_back({e, {an.sA[u], an.sC[u], 0}});
        else if (an.inC[u] && an.inB[u])
            vn.push_back({e, {an.sB[u], an.sC[u], 0}});
        else
            vk.push_back({e, {an.sA[u], an.sB[u], 0}});
    }
    if (empty_k || in.ab_prefactor == 0.0) {
        kind = 1;
        return EIN_SUCCESS;
    }
    int64_t const M = total_of(vm), N = total_of(vn), K = total_of(vk), NB = total_of(vb);
    if (M >= (int64_t{1} << 31) || N >= (int64_t{1} << 31) || K >= (int64_t{1} << 31) || NB >= (int64_t{1} << 31))
        return fail(EIN_ERROR_UNSUPPORTED, "grouped contraction: index range above 2^31");
    auto gm = order_and_fold(vm, 2, 0, 1);
    auto gn = order_and_fold(vn, 2, 0, 1);
    bool const a_i_fast = !gm.empty() && std::llabs(gm.back().s[0]) == 1;
    auto gk = a_i_fast ? order_and_fold(vk, 2, 1, 0) : order_and_fold(vk, 2, 0, 1);
    bool const b_j_fast = (!gn.empty() && std::llabs(gn.back().s[0]) == 1) ||
                          !(!gk.empty() && std::llabs(gk.back().s[1]) == 1);
    bool const a_fast = a_i_fast || !(!gk.empty() && std::llabs(gk.back().s[0]) == 1);
    auto gbv = order_and_fold(vb, 3, 2, 0);
    GatherProblem<T> P{};
    EIN_TRY(fill_group(P.gm, gm, 2));
    EIN_TRY(fill_group(P.gn, gn, 2));
    EIN_TRY(fill_group(P.gk, gk, 2));
    EIN_TRY(fill_group(P.gb, gbv, 3));
    set_k_affine(P, gk);
    P.m = M;
    P.n = N;
    P.k = K;
    P.batch = NB;
    P.A = static_cast<T const *>(in.a);
    P.B = static_cast<T const *>(in.b);
    P.C = static_cast<T *>(in.c);
    P.alpha = T(in.ab_prefactor);
    P.beta = c_repeats ? T(1) : T(in.c_prefactor);
    P.partial = nullptr;
    int const bms[2] = {64, 128};
    for (int v = 0; v < 2; ++v) {
        out[v].P = P;
        out[v].P.tiles_m = uint32_t(ceil_div(M, bms[v]));
        out[v].P.tiles_n = uint32_t(ceil_div(N, 64));
        out[v].flags = (a_fast ? 1 : 0) | (b_j_fast ? 2 : 0);
        out[v].pad = 0;
    }
    return EIN_SUCCESS;
}

// The result equals running the items one after another in list order. Item j depends on
// an earlier item i when j's output overlaps i's output or inputs, or j's inputs overlap
// i's output; j then runs in a wave after i's. Each item goes to the wave after the last
// one holding an item it depends on, so accumulations into one output keep list order and
// a wave's items never touch each other's outputs. Scales of C (c_prefactor with repeated
// C labels, or an item that only scales) run at the start of the item's own wave.
template <typename T>
int plan_create(ein_contraction_item const *items, int count, GroupedPlan &plan,
                std::vector<int> const *forced_waves = nullptr) {
    plan.elem = int(sizeof(T));
    std::vector<std::vector<GroupedItem<T>>> wave_items[2];
    struct Placed {
        ByteRange out, in_a, in_b;
        int wave;
    };
    std::vector<Placed> placed;
    size_t const eb = sizeof(T);
    for (int i = 0; i < count; ++i) {
        GroupedItem<T> gi[2]{};
        int kind = 0;
        bool rep = false;
        EIN_TRY(build_item<T>(items[i], gi, kind, rep));
        if (kind == 2) continue;
        Placed me{range_of(items[i].C, items[i].c, eb), range_of(items[i].A, items[i].a, eb),
                  range_of(items[i].B, items[i].b, eb), 0};
        if (forced_waves != nullptr) {
            me.wave = (*forced_waves)[plan.items.size()];
        } else {
            for (auto const &p : placed)
                if (p.wave >= me.wave &&
                    (me.out.overlaps(p.out) || me.out.overlaps(p.in_a) || me.out.overlaps(p.in_b) ||
                     (kind == 0 && (me.in_a.overlaps(p.out) || me.in_b.overlaps(p.out)))))
                    me.wave = p.wave + 1;
        }
        size_t const w = size_t(me.wave);
        while (w >= plan.waves.size()) {
            plan.waves.emplace_back();
            wave_items[0].emplace_back();
            wave_items[1].emplace_back();
        }
        placed.push_back(me);
        plan.items.push_back(items[i]);
        plan.wave_of.push_back(me.wave);
        // Items with 16 or more 128 x 128 tiles keep their own tuned kernels (wide DMA,
        // 128 x 128 tiles, stream-K over all CUs), and so do items with a long contraction
        // (k >= 2048), whose few tiles would each run k/16 slabs on one CU here while
        // stream-K splits them over the GPU. The persistent kernel takes the many small,
        // short items, for which one launch each would cost more than the work.
        bool const big_tiles = kind == 0 && gi[0].P.m >= 128 && gi[0].P.n >= 128 &&
                               ceil_div(gi[0].P.m, 128) * ceil_div(gi[0].P.n, 128) * gi[0].P.batch >= 16;
        if (kind == 0 && (big_tiles || gi[0].P.k >= 2048)) {
            plan.waves[w].large.push_back(items[i]);
            continue;
        }
        if (kind == 1 || rep) plan.waves[w].prescale.push_back({items[i].C, items[i].c, items[i].c_prefactor});
        if (kind == 1) continue;
        wave_items[0][w].push_back(gi[0]);
        wave_items[1][w].push_back(gi[1]);
    }
    for (size_t w = 0; w < plan.waves.size(); ++w) {
        // offset tables of the wave's items (shared by both tile heights)
        {
            size_t total = 0;
            for (auto const &it : wave_items[0][w]) total += gather_table_entries(it.P.m, it.P.n, it.P.k);
            std::vector<int64_t> host(std::max<size_t>(total, 1));
            EIN_TRY(ein_malloc(&plan.waves[w].d_tables, std::max<size_t>(total * sizeof(int64_t), 256)));
            int64_t *dev = static_cast<int64_t *>(plan.waves[w].d_tables);
            size_t at = 0;
            for (size_t q = 0; q < wave_items[0][w].size(); ++q) {
                GatherProblem<T> &P0 = wave_items[0][w][q].P;
                int64_t *h           = host.data() + at;
                host_offsets(P0.gm, P0.m, h, h + P0.m);
                host_offsets(P0.gn, P0.n, h + 2 * P0.m, h + 2 * P0.m + P0.n);
                host_offsets(P0.gk, P0.k, h + 2 * (P0.m + P0.n), h + 2 * (P0.m + P0.n) + P0.k);
                set_tables(P0, dev + at);
                set_tables(wave_items[1][w][q].P, dev + at);
                at += gather_table_entries(P0.m, P0.n, P0.k);
            }
            if (total) EIN_HIP_CHECK(hipMemcpy(dev, host.data(), total * sizeof(int64_t), hipMemcpyHostToDevice));
        }
        for (int v = 0; v < 2; ++v) {
            auto const &its = wave_items[v][w];
            std::vector<GroupedTile> tl;
            std::vector<int64_t> cost;
            for (size_t q = 0; q < its.size(); ++q) {
                uint64_t nt = uint64_t(its[q].P.tiles_m) * its[q].P.tiles_n * uint64_t(its[q].P.batch);
                if (nt > 0xffffffffu) return fail(EIN_ERROR_UNSUPPORTED, "grouped contraction: too many tiles");
                for (uint64_t t = 0; t < nt; ++t) {
                    tl.push_back({uint32_t(q), uint32_t(t)});
                    cost.push_back(its[q].P.k);
                }
            }
            // longest contraction first (LPT over a fixed-stride persistent walk)
            std::vector<size_t> order(tl.size());
            for (size_t z = 0; z < order.size(); ++z) order[z] = z;
            std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
            std::vector<GroupedTile> sorted(tl.size());
            for (size_t z = 0; z < order.size(); ++z) sorted[z] = tl[order[z]];
            size_t ib = its.size() * sizeof(GroupedItem<T>), tb = sorted.size() * sizeof(GroupedTile);
            EIN_TRY(ein_malloc(&plan.waves[w].d_items[v], std::max<size_t>(ib, 256)));
            EIN_TRY(ein_malloc(&plan.waves[w].d_tiles[v], std::max<size_t>(tb, 256)));
            if (ib) EIN_HIP_CHECK(hipMemcpy(plan.waves[w].d_items[v], its.data(), ib, hipMemcpyHostToDevice));
            if (tb) EIN_HIP_CHECK(hipMemcpy(plan.waves[w].d_tiles[v], sorted.data(), tb, hipMemcpyHostToDevice));
            plan.waves[w].ntiles[v] = uint32_t(sorted.size());
            plan.tiles_total[v] += int64_t(sorted.size());
        }
    }
    // default tile height: 128 rows only when it still gives >= 2 tiles per resident slot
    plan.bm_index = plan.tiles_total[1] >= int64_t(num_cus()) * 4 ? 1 : 0;
    return EIN_SUCCESS;
}

template <typename T>
int plan_launch(GroupedPlan const &plan, int bm_index, hipStream_t stream) {
    void const *k128 = reinterpret_cast<void const *>(&gemm_grouped_kernel<T, 128>);
    void const *k64  = reinterpret_cast<void const *>(&gemm_grouped_kernel<T, 64>);
    uint32_t const slots =
        uint32_t(num_cus()) * uint32_t(resident_blocks(bm_index == 1 ? k128 : k64, GEMM_THREADS));
    for (auto const &w : plan.waves) {
        for (auto const &ps : w.prescale)
            EIN_TRY(tensor_scale<T>(ps.C, static_cast<T *>(ps.c), T(ps.factor), stream));
        uint32_t nt = w.ntiles[bm_index];
        if (nt != 0) {
            uint32_t grid = nt < slots ? nt : slots;
            if (bm_index == 1)
                gemm_grouped_kernel<T, 128><<<grid, GEMM_THREADS, 0, stream>>>(
                    static_cast<GroupedItem<T> const *>(w.d_items[1]), static_cast<GroupedTile const *>(w.d_tiles[1]),
                    nt);
            else
                gemm_grouped_kernel<T, 64><<<grid, GEMM_THREADS, 0, stream>>>(
                    static_cast<GroupedItem<T> const *>(w.d_items[0]), static_cast<GroupedTile const *>(w.d_tiles[0]), nt);
            EIN_LAUNCH_CHECK();
        }
        for (auto const &it : w.large)
            EIN_TRY(einsum_execute<T>(EIN_ALG_INDETERMINATE, T(it.c_prefactor), it.C, static_cast<T *>(it.c),
                                      T(it.ab_prefactor), it.A, static_cast<T const *>(it.a), it.B,
                                      static_cast<T const *>(it.b), nullptr, stream));
    }
    return EIN_SUCCESS;
}

// Autotuner: times both tile heights on scratch operands laid out like the plan's own (so
// the user's tensors are untouched) and keeps the faster one.
template <typename T>
int plan_autotune(GroupedPlan &plan, hipStream_t stream) {
    if (plan.items.empty()) return EIN_SUCCESS;
    // it allocates, clears and times scratch operands: not something a graph can hold
    if (stream_capturing(stream)) return fail(EIN_ERROR_VALUE, "grouped plan autotune: not allowed during graph capture");
    int64_t sa = 1, sb = 1, sc = 1;
    for (auto const &it : plan.items) {
        sa = std::max(sa, span_of(it.A));
        sb = std::max(sb, span_of(it.B));
        sc = std::max(sc, span_of(it.C));
    }
    void *pa = nullptr, *pb = nullptr, *pc = nullptr;
    EIN_TRY(ein_malloc(&pa, size_t(sa) * sizeof(T)));
    EIN_TRY(ein_malloc(&pb, size_t(sb) * sizeof(T)));
    EIN_TRY(ein_malloc(&pc, size_t(sc) * sizeof(T)));
    EIN_HIP_CHECK(hipMemsetAsync(pa, 0, size_t(sa) * sizeof(T), stream));
    EIN_HIP_CHECK(hipMemsetAsync(pb, 0, size_t(sb) * sizeof(T), stream));
    EIN_HIP_CHECK(hipMemsetAsync(pc, 0, size_t(sc) * sizeof(T), stream));
    std::vector<ein_contraction_item> scratch = plan.items;
    for (auto &it : scratch) {
        int64_t ba = 0, bb = 0, bc = 0;
        (void)span_of(it.A, &ba);
        (void)span_of(it.B, &bb);
        (void)span_of(it.C, &bc);
        it.a = static_cast<T *>(pa) + ba;
        it.b = static_cast<T *>(pb) + bb;
        it.c = static_cast<T *>(pc) + bc;
    }
    // the scratch items all write one buffer; keep the plan's own wave structure (the
    // aliasing only affects the values, which are discarded)
    GroupedPlan tp;
    int st = plan_create<T>(scratch.data(), int(scratch.size()), tp, &plan.wave_of);
    if (st == EIN_SUCCESS) {
        hipEvent_t e0, e1;
        EIN_HIP_CHECK(hipEventCreate(&e0));
        EIN_HIP_CHECK(hipEventCreate(&e1));
        float best = 0.f;
        int best_v = plan.bm_index;
        for (int v = 0; v < 2 && st == EIN_SUCCESS; ++v) {
            if (tp.tiles_total[v] == 0) continue;
            st = plan_launch<T>(tp, v, stream); // warm-up
            EIN_HIP_CHECK(hipEventRecord(e0, stream));
            for (int r = 0; r < 3 && st == EIN_SUCCESS; ++r) st = plan_launch<T>(tp, v, stream);
            EIN_HIP_CHECK(hipEventRecord(e1, stream));
            EIN_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            EIN_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (v == 0 || ms < best) {
                best   = ms;
                best_v = v;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        plan.bm_index = best_v;
        plan.tuned    = true;
    }
    (void)hipStreamSynchronize(stream);
    for (auto &w : tp.waves) free_wave(w, stream, true);
    (void)ein_free_async(pa, stream);
    (void)ein_free_async(pb, stream);
    (void)ein_free_async(pc, stream);
    return st;
}

} // namespace
} // namespace ein

extern "C" {

int ein_grouped_plan_create(int32_t elem_bytes, const ein_contraction_item *items, int32_t count, void **plan) {
    if (elem_bytes != 8 && elem_bytes != 4) return ein::fail(-1, "grouped plan: elem_bytes must be 8 or 4");
    if (count < 0 || (count > 0 && items == nullptr)) return ein::fail(-2, "grouped plan: bad item list");
    if (plan == nullptr) return -4;
    auto *p = new ein::GroupedPlan();
    int st = elem_bytes == 8 ? ein::plan_create<double>(items, count, *p) : ein::plan_create<float>(items, count, *p);
    if (st != EIN_SUCCESS) {
        for (auto &w : p->waves) ein::free_wave(w);
        delete p;
        return st;
    }
    *plan = p;
    return EIN_SUCCESS;
}

int ein_grouped_plan_execute(void *plan, ein_stream_t stream) {
    if (plan == nullptr) return -1;
    auto *p = static_cast<ein::GroupedPlan *>(plan);
    hipStream_t const s = ein::as_stream(stream);
    if (std::find(p->streams.begin(), p->streams.end(), s) == p->streams.end()) p->streams.push_back(s);
    return p->elem == 8 ? ein::plan_launch<double>(*p, p->bm_index, s) : ein::plan_launch<float>(*p, p->bm_index, s);
}

int ein_grouped_plan_autotune(void *plan, ein_stream_t stream) {
    if (plan == nullptr) return -1;
    auto *p = static_cast<ein::GroupedPlan *>(plan);
    return p->elem == 8 ? ein::plan_autotune<double>(*p, ein::as_stream(stream))
                        : ein::plan_autotune<float>(*p, ein::as_stream(stream));
}

int ein_grouped_plan_info(void *plan, int32_t *tile_rows, int32_t *waves, int64_t *tiles) {
    if (plan == nullptr) return -1;
    auto *p = static_cast<ein::GroupedPlan *>(plan);
    if (tile_rows) *tile_rows = p->bm_index ? 128 : 64;
    if (waves) *waves = int32_t(p->waves.size());
    if (tiles) *tiles = p->tiles_total[p->bm_index];
    return EIN_SUCCESS;
}

int ein_grouped_plan_destroy(void *plan) {
    if (plan == nullptr) return EIN_SUCCESS;
    auto *p = static_cast<ein::GroupedPlan *>(plan);
    // kernels of the plan may still be queued: the buffers are reused only after them
    bool const one = p->streams.size() == 1;
    for (auto &w : p->waves) ein::free_wave(w, one ? p->streams[0] : nullptr, one);
    delete p;
    return EIN_SUCCESS;
}

} // extern "C"
