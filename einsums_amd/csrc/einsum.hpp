// Runtime einsum dispatch (host side).
#pragma once

#include "ein_common.hpp"

namespace ein {

int einsum_classify(ein_tensor_desc const &C, ein_tensor_desc const &A, ein_tensor_desc const &B);

// Runs `alg` (EIN_ALG_*; EIN_ALG_INDETERMINATE = classify first) and falls back to the
// generic contraction when the operand views do not fold.
template <typename T>
int einsum_execute(int alg, T cp, ein_tensor_desc const &C, T *c, T abp, ein_tensor_desc const &A, T const *a,
                   ein_tensor_desc const &B, T const *b, int32_t *used, hipStream_t stream);

// GEMM of a GEMM-mapped einsum: the split-precision GEMM when this thread's split mode is
// on (FP64, unbatched), else the MFMA GEMM.
template <typename T>
struct GemmProblem;
template <typename T>
int gemm_dispatch(GemmProblem<T> const &P, hipStream_t stream);

} // namespace ein
