// Generic binary contraction: label analysis and the device-side index groups.
#pragma once

#include "blas1.hpp"
#include "gemm.hpp"

namespace ein {

constexpr int GMAX = 8; // digits per index group

// A group of labels walked as one linear index; digit d has extent ext[d] and element
// strides s[0][d], s[1][d], s[2][d] in up to three operands (innermost digit last).
struct IndexGroup {
    int32_t n;
    uint32_t total;
    uint32_t ext[GMAX];
    FastDiv fd[GMAX];
    int64_t s[3][GMAX];
};

// Digits are walked innermost (d = n-1) to outermost. The loop is unrolled over GMAX with
// a uniform guard so every array index is a compile-time constant: a runtime index into
// a kernel-argument struct would force the struct into private memory.
template <int NOPS>
__device__ __forceinline__ void group_decode(uint32_t idx, IndexGroup const &G, int64_t (&off)[NOPS]) {
#pragma unroll
    for (int o = 0; o < NOPS; ++o) off[o] = 0;
#pragma unroll
    for (int d = GMAX - 1; d >= 0; --d) {
        if (d < G.n) {
            uint32_t q = fast_div(idx, G.fd[d]);
            uint32_t r = idx - q * G.ext[d];
#pragma unroll
            for (int o = 0; o < NOPS; ++o) off[o] += int64_t(r) * G.s[o][d];
            idx = q;
        }
    }
}

// Operand order inside the groups: M = (A, C), N = (B, C), K = (A, B), batch = (A, B, C).
// The tile kernels read element offsets from tables (built once per call from the
// groups): rowA/rowC[i] for i < m, colB/colC[j] for j < n, kA/kB[p] for p < k.
template <typename T>
struct GatherProblem {
    IndexGroup gm, gn, gk, gb;
    int64_t m, n, k, batch;
    T const *A;
    T const *B;
    T *C;
    T alpha, beta;
    uint32_t tiles_m, tiles_n;
    StreamK sk;
    T *partial;
    int64_t const *rowA;
    int64_t const *rowC;
    int64_t const *colB;
    int64_t const *colC;
    int64_t const *kA;
    int64_t const *kB;
    // k_affine: the innermost K digit's extent is a multiple of the 16-step slab, so within a
    // slab kA[p] = kA[p0] + (p - p0) * kA_in (same for B)
    int64_t kA_in, kB_in;
    int32_t k_affine;
    int32_t pad_;
};

// Offset tables of one problem: [rowA | rowC | colB | colC | kA | kB] (2(m+n+k) entries).
inline size_t gather_table_entries(int64_t m, int64_t n, int64_t k) { return size_t(2 * (m + n + k)); }

// Result of the host-side label analysis.
struct ContractionAnalysis {
    int nlab;
    int64_t ext[3 * EIN_MAX_RANK];
    int64_t sA[3 * EIN_MAX_RANK], sB[3 * EIN_MAX_RANK], sC[3 * EIN_MAX_RANK];
    bool inA[3 * EIN_MAX_RANK], inB[3 * EIN_MAX_RANK], inC[3 * EIN_MAX_RANK];
    bool c_repeats;
    int64_t c_size;
};

int analyse_contraction(ein_tensor_desc const &C, ein_tensor_desc const &A, ein_tensor_desc const &B,
                        ContractionAnalysis &out);

template <typename T>
int contract_generic(T c_prefactor, ein_tensor_desc const &C, T *c, T ab_prefactor, ein_tensor_desc const &A, T const *a,
                     ein_tensor_desc const &B, T const *b, hipStream_t stream);

} // namespace ein
