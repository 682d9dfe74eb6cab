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

// Per-row exponents of a strided operand: e[r] with max_p |X(r,p)| < 2^e[r] (0 for a zero
// row); with `guard` raised by one where the scaled maximum exceeds 127/128. `scratch` holds
// row_exp_scratch(rows) doubles.
int row_exponents(SliceSrc const &S, int32_t *e, bool guard, double *scratch, hipStream_t stream);
size_t row_exp_scratch(int64_t rows);

// CRT split GEMM (crt.hip): nmod moduli, CRT_MIN_MODULI..16.
constexpr int CRT_MIN_MODULI = 8;
int gemm_split_crt(GemmProblem<double> P, int nmod, hipStream_t stream);
// Bits kept per operand element for nmod moduli and contraction length k (-1 if invalid).
int crt_bits(int nmod, int64_t k);
// The fewest moduli (8..16) that keep at least CRT_AUTO_BITS bits at contraction length k.
constexpr int CRT_AUTO_BITS = 40;
int crt_auto_moduli(int64_t k);

} // namespace ein
