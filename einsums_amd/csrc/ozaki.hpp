// Split-precision (Ozaki) FP64 GEMM on BF16/FP16 MFMA: host interface and shared types.
#pragma once

#include "gemm.hpp"

namespace ein {

// Rows of a strided 2-D operand to be sliced: element (r, p) = X[r*s_row + p*s_k].
struct SliceSrc {
    int64_t rows, k, kp;
    double const *X;
    int64_t s_row, s_k;
};

struct OzProblem {
    int64_t m, n, kp;
    void const *As_raw;
    void const *Bs_raw;
    int32_t const *ea;
    int32_t const *eb;
    double *C;
    int64_t c_sm, c_sn;
    double alpha, beta;
    uint32_t tiles_m, tiles_n;
    // split-K (int8 path): split s covers contraction steps [s*kchunk, min(kp, (s+1)*kchunk));
    // with splits > 1 the partial tiles go to `partial` and a reduction applies the epilogue
    int64_t kchunk;
    int32_t splits;
    int32_t pad_;
    double *partial;
};

int gemm_split(GemmProblem<double> P, int nslices, int slice_type, hipStream_t stream);

} // namespace ein
