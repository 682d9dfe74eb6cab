// CPU oracle for the einsum path -- TEST INFRASTRUCTURE ONLY.
//
// A from-scratch restatement of the reference's CPU einsum semantics, used by tests/
// as the parity checker and by bench.py as the timed CPU baseline ("port"). It is
// never linked into or called by the product library (einsums_amd/libeinsums_amd.so).
//
// What each function follows (paths relative to the Einsums source tree):
//   oracle_classify          - pattern predicates and precedence of
//                              TensorAlgebra/Backends/Dispatch.hpp:194-262,407-439,565-609,
//                              894-955 and EinsumsPy/TensorAlgebra/src/PyTensorAlgebra.cpp:169-300
//   oracle_einsum_generic_*  - einsum_generic_algorithm, TensorAlgebra/Backends/BaseAlgebra.hpp:
//                              85-139 (C zeroed when C_prefactor == 0, else scaled; each target
//                              gets AB_prefactor * sum of A*B, link loop lexicographic), iterating
//                              the full unique index space of C,A,B as the device launcher does
//                              (GPUTensorAlgebra.hpp:219-234) -- labels present only in A or only
//                              in B are summed
//   oracle_einsum_*          - detail::einsum dispatch (Dispatch.hpp:859-957): DOT/DIRECT/GER/
//                              GEMV/GEMM through folded 2-D views (einsum_do_* at :267-837), the
//                              generic loop otherwise
//   oracle_permute_*         - permute (Permute.hpp:330-407): C = beta*C + alpha*perm(A)
//   oracle_gemm/gemv/dot/ger - column-major BLAS semantics of blas::gemm etc. (BLAS.hpp:156)
// Host BLAS for GEMM is resolved at run time (a system OpenBLAS/BLAS, numpy's OpenBLAS ILP64,
// MKL, then a blocked loop), as the reference links whichever vendor BLAS it finds; a library
// is used only after it reproduces the blocked loop on a small case (blas_agrees).
#include <dlfcn.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int MAXR = 16;

struct Desc { // same layout as ein_tensor_desc
    int32_t rank;
    int32_t labels[MAXR];
    int64_t dims[MAXR];
    int64_t strides[MAXR];
};

using Labels = std::vector<int>;
using PosList = std::vector<std::pair<int, int>>;

Labels labels_of(Desc const &T) { return Labels(T.labels, T.labels + T.rank); }

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

enum { GENERIC = 0, DOT, DIRECT, GER, GEMV, GEMM };

// Dispatch.hpp:894-955 for real, basic tensors: Hadamard -> generic; then dot, direct,
// outer, gemv (A matrix), gemv (B matrix), gemm (ranks >= 2).
int classify(Desc const &C, Desc const &A, Desc const &B, bool *gemv_swap) {
    Labels Ci = labels_of(C), Ai = labels_of(A), Bi = labels_of(B);
    if (gemv_swap) *gemv_swap = false;
    if (unique(Ci).size() != Ci.size() || unique(Ai).size() != Ai.size() || unique(Bi).size() != Bi.size())
        return GENERIC;
    if (Ci.empty() && Ai == Bi) return DOT;                 // einsum_is_dot_product :210-215
    if (Ci == Ai && Ci == Bi) return DIRECT;                // einsum_is_direct_product :220-231
    Labels linksAB = intersect(Ai, Bi);
    Labels Cu = unique(Ci);
    PosList tgt_in_A = find_with_pos(Cu, Ai), tgt_in_B = find_with_pos(Cu, Bi);
    if (linksAB.empty() && contiguous(tgt_in_A) && contiguous(tgt_in_B)) return GER; // :236-262
    auto is_mv = [&](Labels const &X, Labels const &Y) {    // einsum_is_matrix_vector :407-439 (X matrix)
        Labels links = unique(difference(intersect(X, Y), Ci));
        PosList lX = find_with_pos(links, X), lY = find_with_pos(links, Y);
        PosList tX = find_with_pos(Cu, X), tY = find_with_pos(Cu, Y);
        PosList X_in_C = find_with_pos(X, Ci), Y_in_C = find_with_pos(Y, Ci);
        return contiguous(lX) && contiguous(lY) && contiguous(tX) && same_ordering(lX, lY) && same_ordering(tX, X_in_C) &&
               !same_ordering(tY, Y_in_C) && Y_in_C.empty();
    };
    if (is_mv(Ai, Bi)) return GEMV;
    if (is_mv(Bi, Ai)) {
        if (gemv_swap) *gemv_swap = true;
        return GEMV;
    }
    if (C.rank >= 2 && A.rank >= 2 && B.rank >= 2) {       // einsum_is_matrix_product :565-609
        Labels links = unique(difference(linksAB, Ci));
        Labels CmA = difference(Ci, Ai), CmB = difference(Ci, Bi);
        PosList lA = find_with_pos(links, Ai), lB = find_with_pos(links, Bi);
        PosList A_in_C = find_with_pos(Ai, Ci), B_in_C = find_with_pos(Bi, Ci);
        bool cond = !CmA.empty() && !CmB.empty() && contiguous(lA) && contiguous(lB) && contiguous(tgt_in_A) &&
                    contiguous(tgt_in_B) && contiguous(A_in_C) && contiguous(B_in_C) && same_ordering(lA, lB) &&
                    same_ordering(tgt_in_A, A_in_C) && same_ordering(tgt_in_B, B_in_C);
        if (cond) return GEMM;
    }
    return GENERIC;
}

// ------------------------------------------------------------- host BLAS
using dgemm_lp64_t  = void (*)(char const *, char const *, int const *, int const *, int const *, double const *,
                              double const *, int const *, double const *, int const *, double const *, double *,
                              int const *);
using dgemm_ilp64_t = void (*)(char const *, char const *, int64_t const *, int64_t const *, int64_t const *,
                               double const *, double const *, int64_t const *, double const *, int64_t const *,
                               double const *, double *, int64_t const *);
dgemm_lp64_t g_dgemm_lp64   = nullptr;
dgemm_ilp64_t g_dgemm_ilp64 = nullptr;
std::string g_blas_name     = "builtin blocked loop";
std::string g_rejected; // libraries that loaded but failed the self-check
bool g_blas_probed          = false;

// The blocked loop: per (i,j) a sequential sum over p, over arbitrary 2-D strides.
template <typename T>
void gemm_loop(int64_t m, int64_t n, int64_t k, T alpha, T const *A, int64_t a_sm, int64_t a_sk, T const *B,
               int64_t b_sk, int64_t b_sn, T beta, T *C, int64_t c_sm, int64_t c_sn) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t j0 = 0; j0 < n; j0 += 64) {
        for (int64_t i0 = 0; i0 < m; i0 += 64) {
            int64_t i1 = std::min(m, i0 + 64), j1 = std::min(n, j0 + 64);
            for (int64_t j = j0; j < j1; ++j)
                for (int64_t i = i0; i < i1; ++i) {
                    T s = T(0);
                    for (int64_t p = 0; p < k; ++p) s += A[i * a_sm + p * a_sk] * B[p * b_sk + j * b_sn];
                    T *cp = C + i * c_sm + j * c_sn;
                    *cp = beta == T(0) ? alpha * s : alpha * s + beta * *cp;
                }
        }
    }
}

// A dlopen'ed dgemm is trusted only after it reproduces the blocked loop on a small case in
// all four transpose combinations (padded leading dimensions, alpha and beta not 1). A
// library that loads but computes something else in this process (seen with MKL's runtime
// dispatch in a process that also holds the HIP runtime) is rejected and the next one tried.
bool blas_agrees(dgemm_lp64_t f32i, dgemm_ilp64_t f64i) {
    int64_t const m = 37, n = 29, k = 41, ld = 53;
    std::vector<double> A(size_t(ld) * 64), B(size_t(ld) * 64), C0(size_t(ld) * n), C(C0.size()), R(C0.size());
    uint64_t x = 0x9E3779B97F4A7C15ull;
    auto next = [&] {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return double(int64_t(x % 2001) - 1000) / 1000.0;
    };
    for (auto &v : A) v = next();
    for (auto &v : B) v = next();
    for (auto &v : C0) v = next();
    for (int t = 0; t < 4; ++t) {
        char ta = (t & 1) ? 'T' : 'N', tb = (t & 2) ? 'T' : 'N';
        double const alpha = 0.75, beta = -1.25;
        C = C0;
        R = C0;
        if (f32i) {
            int im = int(m), in = int(n), ik = int(k), il = int(ld);
            f32i(&ta, &tb, &im, &in, &ik, &alpha, A.data(), &il, B.data(), &il, &beta, C.data(), &il);
        } else {
            int64_t im = m, in = n, ik = k, il = ld;
            f64i(&ta, &tb, &im, &in, &ik, &alpha, A.data(), &il, B.data(), &il, &beta, C.data(), &il);
        }
        // column-major: op(A)(i,p), op(B)(p,j), C(i,j) = C[i + j*ld]
        int64_t a_sm = ta == 'N' ? 1 : ld, a_sk = ta == 'N' ? ld : 1;
        int64_t b_sk = tb == 'N' ? 1 : ld, b_sn = tb == 'N' ? ld : 1;
        gemm_loop<double>(m, n, k, alpha, A.data(), a_sm, a_sk, B.data(), b_sk, b_sn, beta, R.data(), 1, ld);
        double num = 0, den = 0;
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) {
                double d = C[size_t(i + j * ld)] - R[size_t(i + j * ld)];
                num += d * d;
                den += R[size_t(i + j * ld)] * R[size_t(i + j * ld)];
            }
        if (!(num <= 1e-26 * den)) return false;
    }
    return true;
}

void probe_blas() {
    if (g_blas_probed) return;
    g_blas_probed = true;
    if (std::getenv("ORACLE_NO_BLAS")) return;
    std::vector<std::pair<std::string, bool>> cands; // (path, ILP64 scipy-prefixed symbols)
    for (char const *p : {"libopenblas.so.0", "libopenblas.so", "libblas.so.3"}) cands.push_back({p, false});
    if (char const *np = std::getenv("ORACLE_OPENBLAS64")) cands.push_back({np, true});
    std::string conda = std::getenv("CONDA_PREFIX") ? std::getenv("CONDA_PREFIX") : "/opt/conda";
    cands.push_back({conda + "/lib/libmkl_rt.so", false});
    cands.push_back({"/opt/conda/lib/libmkl_rt.so", false});
    for (auto const &c : cands) {
        void *h = dlopen(c.first.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        dgemm_lp64_t f32i  = c.second ? nullptr : reinterpret_cast<dgemm_lp64_t>(dlsym(h, "dgemm_"));
        dgemm_ilp64_t f64i = c.second ? reinterpret_cast<dgemm_ilp64_t>(dlsym(h, "scipy_dgemm_64_")) : nullptr;
        if ((f32i || f64i) && blas_agrees(f32i, f64i)) {
            g_dgemm_lp64  = f32i;
            g_dgemm_ilp64 = f64i;
            g_blas_name   = c.first;
            return;
        }
        g_rejected += (g_rejected.empty() ? "" : ", ") + c.first;
    }
}

// C(i,j) = alpha*sum_p A(i,p)B(p,j) + beta*C(i,j) over arbitrary 2-D strides
// (beta == 0: C is overwritten, as BLAS does).
template <typename T>
void gemm_strided(int64_t m, int64_t n, int64_t k, T alpha, T const *A, int64_t a_sm, int64_t a_sk, T const *B,
                  int64_t b_sk, int64_t b_sn, T beta, T *C, int64_t c_sm, int64_t c_sn) {
    if (m <= 0 || n <= 0) return;
    if constexpr (std::is_same_v<T, double>) {
        probe_blas();
        // column-major BLAS call when every operand has a unit stride
        bool const c_col = c_sm == 1, c_row = c_sn == 1 && !c_col;
        if ((g_dgemm_lp64 || g_dgemm_ilp64) && (c_col || c_row) && k > 0) {
            // normalise to column-major C: if C is row-major, compute C^T = B^T A^T
            int64_t M = m, N = n, as_m = a_sm, as_k = a_sk, bs_k = b_sk, bs_n = b_sn, ldc = c_sn;
            T const *PA = A, *PB = B;
            if (c_row) {
                M = n;
                N = m;
                PA = B;
                as_m = b_sn;
                as_k = b_sk;
                PB = A;
                bs_k = a_sk;
                bs_n = a_sm;
                ldc = c_sm;
            }
            char ta = 0, tb = 0;
            int64_t lda = 0, ldb = 0;
            if (as_m == 1) { ta = 'N'; lda = as_k; } else if (as_k == 1) { ta = 'T'; lda = as_m; }
            if (bs_k == 1) { tb = 'N'; ldb = bs_n; } else if (bs_n == 1) { tb = 'T'; ldb = bs_k; }
            int64_t nra = ta == 'N' ? M : k, nrb = tb == 'N' ? k : N;
            if (ta && tb && lda >= std::max<int64_t>(1, nra) && ldb >= std::max<int64_t>(1, nrb) &&
                ldc >= std::max<int64_t>(1, M)) {
                if (g_dgemm_lp64 && M < INT32_MAX && N < INT32_MAX && k < INT32_MAX && lda < INT32_MAX &&
                    ldb < INT32_MAX && ldc < INT32_MAX) {
                    int im = int(M), in = int(N), ik = int(k), ila = int(lda), ilb = int(ldb), ilc = int(ldc);
                    g_dgemm_lp64(&ta, &tb, &im, &in, &ik, &alpha, PA, &ila, PB, &ilb, &beta, C, &ilc);
                    return;
                }
                if (g_dgemm_ilp64) {
                    g_dgemm_ilp64(&ta, &tb, &M, &N, &k, &alpha, PA, &lda, PB, &ldb, &beta, C, &ldc);
                    return;
                }
            }
        }
    }
    gemm_loop<T>(m, n, k, alpha, A, a_sm, a_sk, B, b_sk, b_sn, beta, C, c_sm, c_sn);
}

// ------------------------------------------------------------ generic loop
struct Space {
    int n = 0;                  // number of unique labels
    int64_t ext[3 * MAXR];
    int64_t sC[3 * MAXR], sA[3 * MAXR], sB[3 * MAXR];
    bool inC[3 * MAXR];
};

int build_space(Desc const &C, Desc const &A, Desc const &B, Space &S) {
    Desc const *T[3] = {&C, &A, &B};
    int labs[3 * MAXR];
    S.n = 0;
    for (int i = 0; i < 3 * MAXR; ++i) {
        S.ext[i] = -1;
        S.sC[i] = S.sA[i] = S.sB[i] = 0;
        S.inC[i] = false;
    }
    for (int t = 0; t < 3; ++t) {
        if (T[t]->rank < 0 || T[t]->rank > MAXR) return 1004;
        for (int d = 0; d < T[t]->rank; ++d) {
            int u = -1;
            for (int q = 0; q < S.n; ++q)
                if (labs[q] == T[t]->labels[d]) u = q;
            if (u < 0) {
                u = S.n++;
                labs[u] = T[t]->labels[d];
                S.ext[u] = T[t]->dims[d];
            } else if (S.ext[u] != T[t]->dims[d]) {
                return 1003; // dimension_error
            }
            int64_t s = T[t]->strides[d];
            if (t == 0) { S.sC[u] += s; S.inC[u] = true; }
            else if (t == 1) S.sA[u] += s;
            else S.sB[u] += s;
        }
    }
    return 0;
}

template <typename T>
void scale_view(Desc const &C, T *c, T beta) {
    int64_t total = 1;
    for (int d = 0; d < C.rank; ++d) total *= C.dims[d];
    for (int64_t e = 0; e < total; ++e) {
        int64_t off = 0, rem = e;
        for (int d = C.rank - 1; d >= 0; --d) {
            off += (rem % C.dims[d]) * C.strides[d];
            rem /= C.dims[d];
        }
        c[off] = beta == T(0) ? T(0) : beta * c[off];
    }
}

template <typename T, typename Acc>
int generic(T cp, Desc const &C, T *c, T abp, Desc const &A, T const *a, Desc const &B, T const *b) {
    Space S;
    int st = build_space(C, A, B, S);
    if (st) return st;
    // C = 0 or C *= C_prefactor over the whole view (BaseAlgebra.hpp:118-133)
    scale_view<T>(C, c, cp);
    std::vector<int> tgt, lnk;
    for (int u = 0; u < S.n; ++u) (S.inC[u] ? tgt : lnk).push_back(u);
    int64_t nt = 1, nl = 1;
    for (int u : tgt) nt *= S.ext[u];
    for (int u : lnk) nl *= S.ext[u];
    if (nt == 0) return 0;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        int64_t oc = 0, oa = 0, ob = 0, rem = t;
        for (int q = int(tgt.size()) - 1; q >= 0; --q) {
            int u = tgt[q];
            int64_t d = rem % S.ext[u];
            rem /= S.ext[u];
            oc += d * S.sC[u];
            oa += d * S.sA[u];
            ob += d * S.sB[u];
        }
        // link loop, lexicographic (last link label fastest), walked as an odometer
        Acc sum = Acc(0);
        int64_t dig[3 * MAXR] = {0};
        int64_t la = oa, lb = ob;
        int const nlk = int(lnk.size());
        for (int64_t l = 0; l < nl; ++l) {
            sum += Acc(a[la]) * Acc(b[lb]);
            for (int q = nlk - 1; q >= 0; --q) {
                int u = lnk[q];
                la += S.sA[u];
                lb += S.sB[u];
                if (++dig[q] < S.ext[u]) break;
                la -= S.ext[u] * S.sA[u];
                lb -= S.ext[u] * S.sB[u];
                dig[q] = 0;
            }
        }
        c[oc] += T(Acc(abp) * sum);
    }
    return 0;
}

// --------------------------------------------------------------- folding
bool fold(Desc const &T, std::vector<int> pos, int64_t &extent, int64_t &stride) {
    std::sort(pos.begin(), pos.end());
    extent = 1;
    stride = 1;
    for (size_t i = 0; i + 1 < pos.size(); ++i)
        if (pos[i + 1] != pos[i] + 1) return false;
    int64_t expect = -1;
    bool first = true;
    for (int i = int(pos.size()) - 1; i >= 0; --i) {
        int d = pos[i];
        extent *= T.dims[d];
        if (T.dims[d] == 1) continue;
        if (first) { stride = T.strides[d]; first = false; }
        else if (T.strides[d] != expect) return false;
        expect = T.strides[d] * T.dims[d];
    }
    return true;
}

std::vector<int> positions_of(Desc const &T, Labels const &labs) {
    std::vector<int> pos;
    for (int d = 0; d < T.rank; ++d)
        if (std::find(labs.begin(), labs.end(), T.labels[d]) != labs.end()) pos.push_back(d);
    return pos;
}

Labels labels_at(Desc const &T, std::vector<int> pos) {
    std::sort(pos.begin(), pos.end());
    Labels out;
    for (int p : pos) out.push_back(T.labels[p]);
    return out;
}

std::vector<int> all_axes(Desc const &T) {
    std::vector<int> v(T.rank);
    for (int d = 0; d < T.rank; ++d) v[d] = d;
    return v;
}

// Runs the BLAS-shaped branch of detail::einsum; returns false when a view does not fold
// (the reference then runs the generic algorithm).
template <typename T>
bool blas_branch(int alg, bool gemv_swap, T cp, Desc const &C, T *c, T abp, Desc const &A, T const *a, Desc const &B,
                 T const *b) {
    Labels Ci = labels_of(C), Ai = labels_of(A), Bi = labels_of(B);
    switch (alg) {
    case DOT: { // Dispatch.hpp:900-920: C *= C_prefactor; C += AB_prefactor * dot(A, B)
        int64_t ea, sa, eb, sb;
        if (!fold(A, all_axes(A), ea, sa) || !fold(B, all_axes(B), eb, sb)) {
            // linear_algebra::dot on a non-contiguous view loops over the multi-index
            ea = 1;
            for (int d = 0; d < A.rank; ++d) ea *= A.dims[d];
            long double s = 0;
            for (int64_t e = 0; e < ea; ++e) {
                int64_t oa = 0, ob = 0, rem = e;
                for (int d = A.rank - 1; d >= 0; --d) {
                    int64_t x = rem % A.dims[d];
                    rem /= A.dims[d];
                    oa += x * A.strides[d];
                    ob += x * B.strides[d];
                }
                s += (long double)a[oa] * b[ob];
            }
            c[0] = c[0] * cp + abp * T(s);
            return true;
        }
        long double s = 0;
        for (int64_t e = 0; e < ea; ++e) s += (long double)a[e * sa] * b[e * sb];
        c[0] = c[0] * cp + abp * T(s);
        return true;
    }
    case DIRECT: { // linear_algebra::direct_product: C = beta*C + alpha*A.*B
        int64_t ea, sa, eb, sb, ec, sc;
        if (!fold(A, all_axes(A), ea, sa) || !fold(B, all_axes(B), eb, sb) || !fold(C, all_axes(C), ec, sc)) return false;
        for (int64_t e = 0; e < ec; ++e) {
            T v = abp * (a[e * sa] * b[e * sb]);
            c[e * sc] = cp == T(0) ? v : v + cp * c[e * sc];
        }
        return true;
    }
    case GER: { // einsum_do_outer_product :267-401: scale C, then C += alpha * x y^T
        std::vector<int> cA = positions_of(C, Ai), cB = positions_of(C, Bi);
        if (labels_at(C, cA) != Ai || labels_at(C, cB) != Bi) return false;
        int64_t ei, si, ej, sj, ex, sx, ey, sy;
        if (!fold(C, cA, ei, si) || !fold(C, cB, ej, sj) || !fold(A, all_axes(A), ex, sx) || !fold(B, all_axes(B), ey, sy))
            return false;
        scale_view<T>(C, c, cp);
        for (int64_t i = 0; i < ei; ++i)
            for (int64_t j = 0; j < ej; ++j) c[i * si + j * sj] += abp * a[i * sx] * b[j * sy];
        return true;
    }
    case GEMV: { // einsum_do_matrix_vector :444-560
        Desc const &M = gemv_swap ? B : A;
        Desc const &V = gemv_swap ? A : B;
        T const *mp = gemv_swap ? b : a;
        T const *vp = gemv_swap ? a : b;
        Labels Mi = labels_of(M), Vi = labels_of(V);
        Labels links = difference(intersect(Mi, Vi), Ci), tg = intersect(Mi, Ci);
        std::vector<int> mt = positions_of(M, tg), ml = positions_of(M, links), vl = positions_of(V, links),
                         ct = positions_of(C, tg);
        if (labels_at(V, vl) != labels_at(M, ml) || labels_at(C, ct) != labels_at(M, mt)) return false;
        int64_t em, sm, ek, sk, ev, sv, ec, sc;
        if (!fold(M, mt, em, sm) || !fold(M, ml, ek, sk) || !fold(V, vl, ev, sv) || !fold(C, ct, ec, sc)) return false;
        gemm_strided<T>(em, 1, ek, abp, mp, sm, sk, vp, sv, 0, cp, c, sc, 0);
        return true;
    }
    case GEMM: { // einsum_do_matrix_product :614-837
        Labels links = unique(difference(intersect(Ai, Bi), Ci));
        Labels at = intersect(Ai, Ci), bt = intersect(Bi, Ci);
        std::vector<int> At = positions_of(A, at), Al = positions_of(A, links), Bt = positions_of(B, bt),
                         Bl = positions_of(B, links), Ca = positions_of(C, at), Cb = positions_of(C, bt);
        if (labels_at(B, Bl) != labels_at(A, Al)) return false;
        if (labels_at(C, Ca) != labels_at(A, At) || labels_at(C, Cb) != labels_at(B, Bt)) return false;
        int64_t m, asm_, k, ask, kb, bsk, n, bsn, cm, csm, cn, csn;
        if (!fold(A, At, m, asm_) || !fold(A, Al, k, ask) || !fold(B, Bl, kb, bsk) || !fold(B, Bt, n, bsn) ||
            !fold(C, Ca, cm, csm) || !fold(C, Cb, cn, csn))
            return false;
        gemm_strided<T>(m, n, k, abp, a, asm_, ask, b, bsk, bsn, cp, c, csm, csn);
        return true;
    }
    default:
        return false;
    }
}

template <typename T>
int einsum(T cp, Desc const &C, T *c, T abp, Desc const &A, T const *a, Desc const &B, T const *b, int *alg_out) {
    Space S;
    int st = build_space(C, A, B, S);
    if (st) return st;
    bool swap = false;
    int alg = classify(C, A, B, &swap);
    if (alg != GENERIC && blas_branch<T>(alg, swap, cp, C, c, abp, A, a, B, b)) {
        if (alg_out) *alg_out = alg;
        return 0;
    }
    if (alg_out) *alg_out = GENERIC;
    return generic<T, T>(cp, C, c, abp, A, a, B, b);
}

template <typename T>
int permute(int rank, int64_t const *dims_A, int const *perm, T alpha, T const *A, int64_t const *strides_A, T beta,
            T *C, int64_t const *strides_C) {
    int64_t total = 1;
    for (int d = 0; d < rank; ++d) total *= dims_A[d];
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < total; ++e) {
        int64_t idx[MAXR];
        int64_t rem = e;
        for (int d = rank - 1; d >= 0; --d) {
            idx[d] = rem % dims_A[d];
            rem /= dims_A[d];
        }
        int64_t oa = 0, oc = 0;
        for (int d = 0; d < rank; ++d) oa += idx[d] * strides_A[d];
        for (int i = 0; i < rank; ++i) oc += idx[perm[i]] * strides_C[i];
        if (alpha == T(1) && beta == T(0))
            C[oc] = A[oa];
        else
            C[oc] = beta == T(0) ? alpha * A[oa] : alpha * A[oa] + beta * C[oc];
    }
    return 0;
}

} // namespace

extern "C" {

const char *oracle_blas_name() {
    probe_blas();
    static std::string full;
    full = g_blas_name + (g_rejected.empty() ? "" : " (rejected by the self-check: " + g_rejected + ")");
    return full.c_str();
}

int oracle_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

int oracle_classify(Desc const *C, Desc const *A, Desc const *B) { return classify(*C, *A, *B, nullptr); }

int oracle_einsum_generic_f64(double cp, Desc const *C, double *c, double abp, Desc const *A, double const *a,
                              Desc const *B, double const *b) {
    return generic<double, double>(cp, *C, c, abp, *A, a, *B, b);
}
int oracle_einsum_generic_f32(float cp, Desc const *C, float *c, float abp, Desc const *A, float const *a, Desc const *B,
                              float const *b) {
    return generic<float, float>(cp, *C, c, abp, *A, a, *B, b);
}
// Extended-precision accumulation (80-bit long double) for measuring the error of
// reduced-precision kernels.
int oracle_einsum_generic_ld(double cp, Desc const *C, double *c, double abp, Desc const *A, double const *a,
                             Desc const *B, double const *b) {
    return generic<double, long double>(cp, *C, c, abp, *A, a, *B, b);
}
int oracle_einsum_f64(double cp, Desc const *C, double *c, double abp, Desc const *A, double const *a, Desc const *B,
                      double const *b, int *alg) {
    return einsum<double>(cp, *C, c, abp, *A, a, *B, b, alg);
}
int oracle_einsum_f32(float cp, Desc const *C, float *c, float abp, Desc const *A, float const *a, Desc const *B,
                      float const *b, int *alg) {
    return einsum<float>(cp, *C, c, abp, *A, a, *B, b, alg);
}
int oracle_permute_f64(int rank, int64_t const *dims_A, int const *perm, double alpha, double const *A,
                       int64_t const *strides_A, double beta, double *C, int64_t const *strides_C) {
    return permute<double>(rank, dims_A, perm, alpha, A, strides_A, beta, C, strides_C);
}
int oracle_permute_f32(int rank, int64_t const *dims_A, int const *perm, float alpha, float const *A,
                       int64_t const *strides_A, float beta, float *C, int64_t const *strides_C) {
    return permute<float>(rank, dims_A, perm, alpha, A, strides_A, beta, C, strides_C);
}
// Column-major BLAS gemm (blas::gemm semantics).
int oracle_gemm_f64(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, double const *A, int64_t lda,
                    double const *B, int64_t ldb, double beta, double *C, int64_t ldc) {
    bool na = ta == 'n' || ta == 'N', nb = tb == 'n' || tb == 'N';
    gemm_strided<double>(m, n, k, alpha, A, na ? 1 : lda, na ? lda : 1, B, nb ? 1 : ldb, nb ? ldb : 1, beta, C, 1, ldc);
    return 0;
}
// Fast 32-bit division restated for host checking of the device decode.
uint32_t oracle_fastdiv(uint32_t n, uint32_t d) {
    uint32_t l = 0;
    while ((uint64_t{1} << l) < d) ++l;
    uint32_t mul = uint32_t(((uint64_t{1} << 32) * ((uint64_t{1} << l) - d)) / d + 1);
    uint32_t t   = uint32_t((uint64_t{n} * mul) >> 32);
    return (t + n) >> l;
}

} // extern "C"
