// Host-side interface of the MFMA GEMM family.
#pragma once

#include "gemm_core.hpp"

namespace ein {

template <typename T>
struct GemmProblem {
    int64_t m, n, k, batch;
    T const *A;
    int64_t a_sm, a_sk, a_sb;
    T const *B;
    int64_t b_sk, b_sn, b_sb;
    T *C;
    int64_t c_sm, c_sn, c_sb;
    T alpha, beta;
    // filled by the launcher
    uint32_t tiles_m, tiles_n;
    StreamK sk;
    int32_t force_dp;
    T *partial;
};

struct GemmConfig {
    int bm            = 128;
    int bn            = 128;
    int data_parallel = 0; // 1: one tile per workgroup even where stream-K would balance better;
                           // 2: split-K over XCD-local contraction chunks (plan_split_k)
    int blocks_per_cu = 0; // workgroups per CU the schedule plans for; 0 = as many as fit
};

GemmConfig choose_gemm_config(int64_t m, int64_t n, int64_t k, int64_t batch, int elem_bytes);

// Strided (optionally batched) GEMM; `force` overrides the tile heuristic (autotuner).
template <typename T>
int gemm_strided(GemmProblem<T> P, GemmConfig const *force, hipStream_t stream);

// Workspace bytes (from the start of the per-stream workspace) that gemm_strided(P) with
// the default configuration takes; callers that keep their own data in the same workspace
// place it after this many bytes.
template <typename T>
int gemm_workspace_bytes(GemmProblem<T> const &P, size_t *bytes);

} // namespace ein
