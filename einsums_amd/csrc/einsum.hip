// Runtime einsum dispatch: pattern analysis on integer labels, folding of the matched
// index groups into 1-D/2-D strided views, and the call into the matching kernel.
//
// The predicates restate compile_plan (libs/EinsumsPy/TensorAlgebra/src/
// PyTensorAlgebra.cpp:169-300), which mirrors the compile-time ones of
// TensorAlgebra/Backends/Dispatch.hpp:210-262,407-439,565-609. The view folding
// follows einsum_do_matrix_product / einsum_do_matrix_vector / einsum_do_outer_product
// (Dispatch.hpp:267-401,444-560,614-837): an index group folds when its axes are
// contiguous in row-major order (stride[p] == stride[p+1] * dim[p+1]); a group that does
// not fold sends the call to the generic contraction, and the reported algorithm is
// GENERIC -- the same outcome as a reference do_* helper returning false.
#include <algorithm>
#include <utility>
#include <vector>

#include "contract.hpp"
#include "einsum.hpp"
#include "ozaki.hpp"

namespace ein {

namespace {

// Split-precision setting for FP64 GEMM-mapped einsums on this thread (0 = plain FP64;
// -1 = the automatic moduli count of EIN_SLICE_CRT, passed on as nslices 0).
thread_local int t_split_slices = 0;
thread_local int t_split_type   = EIN_SLICE_BF16;

} // namespace

template <typename T>
int gemm_dispatch(GemmProblem<T> const &P, hipStream_t stream) {
    if constexpr (std::is_same_v<T, double>) {
        if (t_split_slices != 0 && P.batch == 1)
            return gemm_split(P, t_split_slices < 0 ? 0 : t_split_slices, t_split_type, stream);
    }
    return gemm_strided<T>(P, nullptr, stream);
}
template int gemm_dispatch<double>(GemmProblem<double> const &, hipStream_t);
template int gemm_dispatch<float>(GemmProblem<float> const &, hipStream_t);

namespace {

template <typename T>
int run_gemm(GemmProblem<T> const &P, hipStream_t stream) {
    return gemm_dispatch<T>(P, stream);
}

using Labels = std::vector<int>;
using PosList = std::vector<std::pair<int, int>>; // (label, position)

Labels labels_of(ein_tensor_desc const &T) { return Labels(T.labels, T.labels + T.rank); }

Labels intersect(Labels const &a, Labels const &b) {
    Labels out;
    for (int x : a)
        for (int y : b)
            if (x == y) out.push_back(x);
    return out;
}

Labels difference(Labels const &a, Labels const &b) {
    Labels out;
    for (int x : a)
        if (std::find(b.begin(), b.end(), x) == b.end()) out.push_back(x);
    return out;
}

Labels unique(Labels const &a) {
    Labels out;
    for (int x : a)
        if (std::find(out.begin(), out.end(), x) == out.end()) out.push_back(x);
    return out;
}

PosList find_with_pos(Labels const &a, Labels const &b) {
    PosList out;
    for (int x : a)
        for (int j = 0; j < int(b.size()); ++j)
            if (x == b[j]) out.push_back({x, j});
    return out;
}

bool contiguous(PosList const &p) {
    for (size_t i = 0; i + 1 < p.size(); ++i)
        if (p[i].second != p[i + 1].second - 1) return false;
    return true;
}

bool same_ordering(PosList const &x, PosList const &y) {
    if (x.size() != y.size() || x.empty()) return false;
    for (size_t i = 0; i < x.size(); ++i)
        if (x[i].first != y[i].first) return false;
    return true;
}

struct Pattern {
    int alg = EIN_ALG_GENERIC;
    bool gemv_swap = false; // matrix operand is B
};

Pattern classify(ein_tensor_desc const &C, ein_tensor_desc const &A, ein_tensor_desc const &B) {
    Labels Ci = labels_of(C), Ai = labels_of(A), Bi = labels_of(B);
    Labels linksAB = intersect(Ai, Bi);
    Labels links   = difference(linksAB, Ci);
    Labels CminusA = difference(Ci, Ai), CminusB = difference(Ci, Bi);
    Labels Au = unique(Ai), Bu = unique(Bi), Cu = unique(Ci), link_u = unique(links);
    bool const A_had = Ai.size() != Au.size(), B_had = Bi.size() != Bu.size(), C_had = Ci.size() != Cu.size();
    bool const no_had = !A_had && !B_had && !C_had;

    PosList link_in_A = find_with_pos(link_u, Ai), link_in_B = find_with_pos(link_u, Bi);
    PosList tgt_in_A = find_with_pos(Cu, Ai), tgt_in_B = find_with_pos(Cu, Bi);
    PosList A_tgt_in_C = find_with_pos(Ai, Ci), B_tgt_in_C = find_with_pos(Bi, Ci);

    bool const c_link_A = contiguous(link_in_A), c_link_B = contiguous(link_in_B);
    bool const c_tgt_A = contiguous(tgt_in_A), c_tgt_B = contiguous(tgt_in_B);
    bool const c_A_in_C = contiguous(A_tgt_in_C), c_B_in_C = contiguous(B_tgt_in_C);
    bool const so_link = same_ordering(link_in_A, link_in_B);
    bool const so_CA = same_ordering(tgt_in_A, A_tgt_in_C), so_CB = same_ordering(tgt_in_B, B_tgt_in_C);

    bool const gemm = !CminusA.empty() && !CminusB.empty() && c_link_A && c_link_B && c_tgt_A && c_tgt_B && c_A_in_C &&
                      c_B_in_C && so_link && so_CA && so_CB && no_had;
    bool const gemv_AB = c_link_A && c_link_B && c_tgt_A && so_link && so_CA && !so_CB && B_tgt_in_C.empty() && no_had;
    bool const gemv_BA = c_link_A && c_link_B && c_tgt_B && so_link && so_CB && !so_CA && A_tgt_in_C.empty() && no_had;
    bool const direct  = Ai == Ci && Bi == Ci && no_had;
    bool const dot     = Ci.empty() && Ai == Bi && no_had;
    bool const outer   = linksAB.empty() && c_tgt_A && c_tgt_B && no_had;

    Pattern p;
    if (dot)
        p.alg = EIN_ALG_DOT;
    else if (direct)
        p.alg = EIN_ALG_DIRECT;
    else if (outer)
        p.alg = EIN_ALG_GER;
    else if (gemv_AB || gemv_BA) {
        p.alg       = EIN_ALG_GEMV;
        p.gemv_swap = gemv_BA && !gemv_AB;
    } else if (gemm)
        p.alg = EIN_ALG_GEMM;
    return p;
}

// Folds the axes at the (adjacent) positions `pos` of T into one strided dimension:
// the non-unit axes must be row-major contiguous; the folded stride is that of the
// innermost non-unit axis.
bool fold_group(ein_tensor_desc const &T, std::vector<int> pos, int64_t &extent, int64_t &stride) {
    std::sort(pos.begin(), pos.end());
    extent = 1;
    stride = 1;
    for (size_t i = 0; i + 1 < pos.size(); ++i)
        if (pos[i + 1] != pos[i] + 1) return false;
    int64_t expect = -1;
    bool first     = true;
    for (int i = int(pos.size()) - 1; i >= 0; --i) {
        int d = pos[i];
        extent *= T.dims[d];
        if (T.dims[d] == 1) continue;
        if (first) {
            stride = T.strides[d];
            first  = false;
        } else if (T.strides[d] != expect) {
            return false;
        }
        expect = T.strides[d] * T.dims[d];
    }
    return true;
}

std::vector<int> positions_of(ein_tensor_desc const &T, Labels const &labs) {
    std::vector<int> pos;
    for (int d = 0; d < T.rank; ++d)
        if (std::find(labs.begin(), labs.end(), T.labels[d]) != labs.end()) pos.push_back(d);
    return pos;
}

// Checks that a group's label order in T equals `order` (so that two folded views walk
// the same multi-index sequence).
bool order_matches(ein_tensor_desc const &T, std::vector<int> const &pos, Labels const &order) {
    if (pos.size() != order.size()) return false;
    std::vector<int> sp = pos;
    std::sort(sp.begin(), sp.end());
    for (size_t i = 0; i < sp.size(); ++i)
        if (T.labels[sp[i]] != order[i]) return false;
    return true;
}

Labels labels_at(ein_tensor_desc const &T, std::vector<int> pos) {
    std::sort(pos.begin(), pos.end());
    Labels out;
    for (int p : pos) out.push_back(T.labels[p]);
    return out;
}

template <typename T>
int run_generic(T cp, ein_tensor_desc const &C, T *c, T abp, ein_tensor_desc const &A, T const *a,
                ein_tensor_desc const &B, T const *b, int32_t *used, hipStream_t stream) {
    if (used) *used = EIN_ALG_GENERIC;
    return contract_generic<T>(cp, C, c, abp, A, a, B, b, stream);
}

int check_operands(ein_tensor_desc const &C, ein_tensor_desc const &A, ein_tensor_desc const &B) {
    ContractionAnalysis an;
    return analyse_contraction(C, A, B, an);
}

} // namespace

int einsum_classify(ein_tensor_desc const &C, ein_tensor_desc const &A, ein_tensor_desc const &B) {
    return classify(C, A, B).alg;
}

template <typename T>
int einsum_execute(int alg, T cp, ein_tensor_desc const &C, T *c, T abp, ein_tensor_desc const &A, T const *a,
                   ein_tensor_desc const &B, T const *b, int32_t *used, hipStream_t stream) {
    EIN_TRY(check_operands(C, A, B));
    int64_t csize = 1;
    for (int d = 0; d < C.rank; ++d) csize *= C.dims[d];
    int64_t asize = 1, bsize = 1;
    for (int d = 0; d < A.rank; ++d) asize *= A.dims[d];
    for (int d = 0; d < B.rank; ++d) bsize *= B.dims[d];
    if (csize == 0 || asize == 0 || bsize == 0) return run_generic<T>(cp, C, c, abp, A, a, B, b, used, stream);

    Labels Ci = labels_of(C), Ai = labels_of(A), Bi = labels_of(B);
    Pattern pat = classify(C, A, B);
    if (alg == EIN_ALG_INDETERMINATE || alg < 0) alg = pat.alg;

    switch (alg) {
    case EIN_ALG_DOT: {
        if (!(Ci.empty() && Ai == Bi)) break;
        int64_t ea, sa, eb, sb;
        std::vector<int> all_a(A.rank), all_b(B.rank);
        for (int d = 0; d < A.rank; ++d) all_a[d] = d;
        for (int d = 0; d < B.rank; ++d) all_b[d] = d;
        if (!fold_group(A, all_a, ea, sa) || !fold_group(B, all_b, eb, sb)) break;
        if (used) *used = EIN_ALG_DOT;
        return dot_impl<T>(ea, abp, a, sa, b, sb, cp, c, stream);
    }
    case EIN_ALG_DIRECT: {
        if (!(Ai == Ci && Bi == Ci)) break;
        int64_t ea, sa, eb, sb, ec, sc;
        std::vector<int> all(C.rank);
        for (int d = 0; d < C.rank; ++d) all[d] = d;
        if (!fold_group(A, all, ea, sa) || !fold_group(B, all, eb, sb) || !fold_group(C, all, ec, sc)) break;
        if (used) *used = EIN_ALG_DIRECT;
        return dirprod_impl<T>(ec, abp, a, sa, b, sb, cp, c, sc, stream);
    }
    case EIN_ALG_GER: {
        if (!intersect(Ai, Bi).empty()) break;
        // C(i, j) += x(i) y(j); i = A's labels, j = B's labels (in C order)
        std::vector<int> cA = positions_of(C, Ai), cB = positions_of(C, Bi);
        if (cA.size() != size_t(A.rank) || cB.size() != size_t(B.rank)) break; // A or B has summed-only labels
        // C has labels of its own (C is constant along them): the outer product would fill
        // one slice only, the generic contraction broadcasts over them
        if (cA.size() + cB.size() != size_t(C.rank)) break;
        if (!order_matches(A, std::vector<int>([&] {
                               std::vector<int> v(A.rank);
                               for (int d = 0; d < A.rank; ++d) v[d] = d;
                               return v;
                           }()),
                           labels_at(C, cA)))
            break;
        if (!order_matches(B, std::vector<int>([&] {
                               std::vector<int> v(B.rank);
                               for (int d = 0; d < B.rank; ++d) v[d] = d;
                               return v;
                           }()),
                           labels_at(C, cB)))
            break;
        int64_t ei, sci, ej, scj, ex, sx, ey, sy;
        std::vector<int> all_a(A.rank), all_b(B.rank);
        for (int d = 0; d < A.rank; ++d) all_a[d] = d;
        for (int d = 0; d < B.rank; ++d) all_b[d] = d;
        if (!fold_group(C, cA, ei, sci) || !fold_group(C, cB, ej, scj) || !fold_group(A, all_a, ex, sx) ||
            !fold_group(B, all_b, ey, sy))
            break;
        bool const col = sci == 1 && scj >= std::max<int64_t>(1, ei);
        bool const row = scj == 1 && sci >= std::max<int64_t>(1, ej);
        if (!col && !row) break;
        if (used) *used = EIN_ALG_GER;
        EIN_TRY(tensor_scale<T>(C, c, cp, stream));
        if (col) return ger_impl<T>(ei, ej, abp, a, sx, b, sy, c, std::max<int64_t>(1, scj), stream);
        return ger_impl<T>(ej, ei, abp, b, sy, a, sx, c, std::max<int64_t>(1, sci), stream);
    }
    case EIN_ALG_GEMV: {
        // matrix operand M (labels: targets + links), vector operand V (links only)
        bool const swap = [&] {
            if (difference(Bi, Ci).size() == Bi.size() && difference(Ai, Ci).size() != Ai.size()) return false;
            if (difference(Ai, Ci).size() == Ai.size()) return true;
            return pat.gemv_swap;
        }();
        ein_tensor_desc const &Md = swap ? B : A;
        ein_tensor_desc const &Vd = swap ? A : B;
        T const *mp = swap ? b : a;
        T const *vp = swap ? a : b;
        Labels Mi = labels_of(Md), Vi = labels_of(Vd);
        Labels links = difference(intersect(Mi, Vi), Ci);
        Labels tgts  = intersect(Mi, Ci);
        if (!difference(Vi, links).empty()) break; // vector must hold only links
        if (tgts.size() != Ci.size()) break;
        std::vector<int> m_t = positions_of(Md, tgts), m_l = positions_of(Md, links);
        std::vector<int> v_l = positions_of(Vd, links), c_t = positions_of(C, tgts);
        if (m_t.size() + m_l.size() != size_t(Md.rank)) break;
        Labels lorder = labels_at(Md, m_l), torder = labels_at(Md, m_t);
        if (!order_matches(Vd, v_l, lorder) || !order_matches(C, c_t, torder)) break;
        int64_t em_t, sm_t, em_l, sm_l, ev, sv, ec, sc;
        if (!fold_group(Md, m_t, em_t, sm_t) || !fold_group(Md, m_l, em_l, sm_l) || !fold_group(Vd, v_l, ev, sv) ||
            !fold_group(C, c_t, ec, sc))
            break;
        if (sm_l == 1 && sm_t >= std::max<int64_t>(1, em_l)) {
            if (used) *used = EIN_ALG_GEMV;
            return gemv_impl<T>('T', em_l, em_t, abp, mp, sm_t, vp, sv, cp, c, sc, stream);
        }
        if (sm_t == 1 && sm_l >= std::max<int64_t>(1, em_t)) {
            if (used) *used = EIN_ALG_GEMV;
            return gemv_impl<T>('N', em_t, em_l, abp, mp, sm_l, vp, sv, cp, c, sc, stream);
        }
        break;
    }
    case EIN_ALG_GEMM: {
        Labels links = unique(difference(intersect(Ai, Bi), Ci));
        Labels a_t = intersect(Ai, Ci), b_t = intersect(Bi, Ci);
        if (links.empty() || a_t.empty() || b_t.empty()) break;
        if (!intersect(a_t, Bi).empty() || !intersect(b_t, Ai).empty()) break; // batch labels
        std::vector<int> A_t = positions_of(A, a_t), A_l = positions_of(A, links);
        std::vector<int> B_t = positions_of(B, b_t), B_l = positions_of(B, links);
        std::vector<int> C_a = positions_of(C, a_t), C_b = positions_of(C, b_t);
        if (A_t.size() + A_l.size() != size_t(A.rank) || B_t.size() + B_l.size() != size_t(B.rank)) break;
        if (C_a.size() + C_b.size() != size_t(C.rank)) break;
        Labels lorder = labels_at(A, A_l);
        if (!order_matches(B, B_l, lorder)) break;
        if (!order_matches(C, C_a, labels_at(A, A_t)) || !order_matches(C, C_b, labels_at(B, B_t))) break;
        int64_t m, a_sm, k, a_sk, kb, b_sk, n, b_sn, cm, c_sm, cn, c_sn;
        if (!fold_group(A, A_t, m, a_sm) || !fold_group(A, A_l, k, a_sk) || !fold_group(B, B_l, kb, b_sk) ||
            !fold_group(B, B_t, n, b_sn) || !fold_group(C, C_a, cm, c_sm) || !fold_group(C, C_b, cn, c_sn))
            break;
        if (used) *used = EIN_ALG_GEMM;
        GemmProblem<T> P{};
        P.m     = m;
        P.n     = n;
        P.k     = k;
        P.batch = 1;
        P.A     = a;
        P.a_sm  = a_sm;
        P.a_sk  = a_sk;
        P.B     = b;
        P.b_sk  = b_sk;
        P.b_sn  = b_sn;
        P.C     = c;
        P.c_sm  = c_sm;
        P.c_sn  = c_sn;
        P.alpha = abp;
        P.beta  = cp;
        return run_gemm<T>(P, stream);
    }
    default:
        break;
    }
    return run_generic<T>(cp, C, c, abp, A, a, B, b, used, stream);
}

template <typename T>
int einsum_adjoint(int32_t wrt, T alpha, const ein_tensor_desc *C, const T *dC, const ein_tensor_desc *A, const T *a,
                          const ein_tensor_desc *B, const T *b, T beta, const ein_tensor_desc *G, T *g,
                          int32_t *algorithm_used, ein_stream_t stream) {
    if (wrt != 0 && wrt != 1) return fail(-1, "einsum_adjoint: wrt must be 0 (A) or 1 (B)");
    if (C == nullptr || A == nullptr || B == nullptr || G == nullptr) return fail(-3, "einsum_adjoint: null descriptor");
    ein_tensor_desc const &X = wrt == 0 ? *A : *B; // differentiated operand
    ein_tensor_desc const &Y = wrt == 0 ? *B : *A; // the other operand
    if (G->rank != X.rank) return fail(EIN_ERROR_RANK, "einsum_adjoint: gradient rank differs from the operand");
    ein_tensor_desc Gd = *G;
    for (int d = 0; d < X.rank; ++d) {
        if (G->dims[d] != X.dims[d]) return fail(EIN_ERROR_DIMENSION, "einsum_adjoint: gradient extents differ");
        Gd.labels[d] = X.labels[d];
    }
    T const *y = wrt == 0 ? b : a;
    // dX[x] = beta*dX + alpha * sum over labels not in X of dC[c] * Y[y]; a label of X found
    // in neither C nor Y (summed in the forward pass) is a label of this product's output
    // only, and dX is constant along it
    return einsum_execute<T>(EIN_ALG_INDETERMINATE, beta, Gd, g, alpha, *C, dC, Y, y, algorithm_used,
                                  as_stream(stream));
}

template int einsum_execute<double>(int, double, ein_tensor_desc const &, double *, double, ein_tensor_desc const &,
                                    double const *, ein_tensor_desc const &, double const *, int32_t *, hipStream_t);
template int einsum_execute<float>(int, float, ein_tensor_desc const &, float *, float, ein_tensor_desc const &,
                                   float const *, ein_tensor_desc const &, float const *, int32_t *, hipStream_t);

} // namespace ein

extern "C" {

int ein_einsum_classify(const ein_tensor_desc *C, const ein_tensor_desc *A, const ein_tensor_desc *B, int32_t *algorithm) {
    if (C == nullptr) return -1;
    if (A == nullptr) return -2;
    if (B == nullptr) return -3;
    if (algorithm == nullptr) return -4;
    for (ein_tensor_desc const *T : {C, A, B})
        if (T->rank < 0 || T->rank > EIN_MAX_RANK) return ein::fail(EIN_ERROR_RANK, "rank out of range");
    *algorithm = ein::einsum_classify(*C, *A, *B);
    return EIN_SUCCESS;
}

int ein_einsum_f64(double c_prefactor, const ein_tensor_desc *C, double *c, double ab_prefactor, const ein_tensor_desc *A,
                   const double *a, const ein_tensor_desc *B, const double *b, int32_t *algorithm_used,
                   ein_stream_t stream) {
    return ein_einsum_execute_f64(EIN_ALG_INDETERMINATE, c_prefactor, C, c, ab_prefactor, A, a, B, b, algorithm_used,
                                  stream);
}

int ein_einsum_f32(float c_prefactor, const ein_tensor_desc *C, float *c, float ab_prefactor, const ein_tensor_desc *A,
                   const float *a, const ein_tensor_desc *B, const float *b, int32_t *algorithm_used, ein_stream_t stream) {
    return ein_einsum_execute_f32(EIN_ALG_INDETERMINATE, c_prefactor, C, c, ab_prefactor, A, a, B, b, algorithm_used,
                                  stream);
}

int ein_einsum_execute_f64(int32_t algorithm, double c_prefactor, const ein_tensor_desc *C, double *c, double ab_prefactor,
                           const ein_tensor_desc *A, const double *a, const ein_tensor_desc *B, const double *b,
                           int32_t *algorithm_used, ein_stream_t stream) {
    if (C == nullptr) return -3;
    if (A == nullptr) return -6;
    if (B == nullptr) return -8;
    return ein::einsum_execute<double>(algorithm, c_prefactor, *C, c, ab_prefactor, *A, a, *B, b, algorithm_used,
                                       ein::as_stream(stream));
}

int ein_einsum_split_f64(int32_t nslices, int32_t slice_type, double c_prefactor, const ein_tensor_desc *C, double *c,
                         double ab_prefactor, const ein_tensor_desc *A, const double *a, const ein_tensor_desc *B,
                         const double *b, int32_t *algorithm_used, ein_stream_t stream) {
    int const type = slice_type & ~EIN_SPLIT_REFINE;
    if (type != EIN_SLICE_BF16 && type != EIN_SLICE_FP16 && type != EIN_SLICE_INT8 && type != EIN_SLICE_CRT)
        return ein::fail(-2, "einsum_split: bad slice type");
    if (type == EIN_SLICE_CRT ? (nslices != 0 && (nslices < 8 || nslices > 16)) : (nslices < 3 || nslices > 6))
        return ein::fail(-1, "einsum_split: nslices must be 3..6 (8..16 moduli, or 0 for automatic, with EIN_SLICE_CRT)");
    ein::t_split_slices = nslices == 0 ? -1 : nslices;
    ein::t_split_type   = slice_type;
    int st = ein_einsum_execute_f64(EIN_ALG_INDETERMINATE, c_prefactor, C, c, ab_prefactor, A, a, B, b, algorithm_used,
                                    stream);
    ein::t_split_slices = 0;
    return st;
}

int ein_einsum_adjoint_f64(int32_t wrt, double alpha, const ein_tensor_desc *C, const double *dC, const ein_tensor_desc *A,
                           const double *a, const ein_tensor_desc *B, const double *b, double beta,
                           const ein_tensor_desc *G, double *g, int32_t *algorithm_used, ein_stream_t stream) {
    return ein::einsum_adjoint<double>(wrt, alpha, C, dC, A, a, B, b, beta, G, g, algorithm_used, stream);
}

int ein_einsum_adjoint_f32(int32_t wrt, float alpha, const ein_tensor_desc *C, const float *dC, const ein_tensor_desc *A,
                           const float *a, const ein_tensor_desc *B, const float *b, float beta, const ein_tensor_desc *G,
                           float *g, int32_t *algorithm_used, ein_stream_t stream) {
    return ein::einsum_adjoint<float>(wrt, alpha, C, dC, A, a, B, b, beta, G, g, algorithm_used, stream);
}

int ein_einsum_execute_f32(int32_t algorithm, float c_prefactor, const ein_tensor_desc *C, float *c, float ab_prefactor,
                           const ein_tensor_desc *A, const float *a, const ein_tensor_desc *B, const float *b,
                           int32_t *algorithm_used, ein_stream_t stream) {
    if (C == nullptr) return -3;
    if (A == nullptr) return -6;
    if (B == nullptr) return -8;
    return ein::einsum_execute<float>(algorithm, c_prefactor, *C, c, ab_prefactor, *A, a, *B, b, algorithm_used,
                                      ein::as_stream(stream));
}

} // extern "C"
