/*
 * einsums_amd.h -- C ABI of the MI355X (gfx950) einsum backend.
 *
 * Every entry point is plain C: device pointers, 64-bit sizes, an explicit stream
 * (a hipStream_t passed as void*, NULL = the null stream) and an int status
 * (0 = success, negative = -(index of the bad argument) as in BLAS `info`,
 * positive = an EIN_ERROR_* code; the message is in ein_last_error_string()).
 * All leaves are asynchronous on `stream`; nothing synchronises unless stated.
 *
 * Layout conventions:
 *   - the BLAS-shaped leaves (gemm/gemv/ger/dot/dirprod/scale/axpby) follow
 *     column-major BLAS semantics exactly like the reference's blas::gpu:: layer;
 *   - tensor descriptors (ein_tensor_desc) carry per-axis extents and element
 *     strides, so any row-major / column-major / strided view can be described.
 *
 * Reference interfaces replaced (paths relative to the Einsums source tree):
 *   libs/Einsums/hipBLAS/include/Einsums/hipBLAS.hpp            (blas::gpu::*)
 *   libs/Einsums/TensorAlgebra/include/Einsums/TensorAlgebra/Backends/GPUTensorAlgebra.hpp
 *   libs/Einsums/TensorAlgebra/include/Einsums/TensorAlgebra/Backends/DevicePermute.hpp
 *   libs/EinsumsPy/TensorAlgebra/src/PyTensorAlgebra.cpp        (compile_plan)
 */
#ifndef EINSUMS_AMD_H
#define EINSUMS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *ein_stream_t;

#define EIN_SUCCESS 0
#define EIN_ERROR_HIP 1001          /* a HIP runtime call failed */
#define EIN_ERROR_ALLOC 1002        /* device allocation failed */
#define EIN_ERROR_DIMENSION 1003    /* extents of a shared label disagree (reference: dimension_error) */
#define EIN_ERROR_RANK 1004         /* rank / label count mismatch (reference: rank_error) */
#define EIN_ERROR_UNSUPPORTED 1005  /* a case this backend does not implement */
#define EIN_ERROR_WORKSPACE 1006    /* workspace too small while a stream is being captured */
#define EIN_ERROR_VALUE 1007        /* bad value (reference: std::invalid_argument / std::domain_error) */

#define EIN_MAX_RANK 16

/* Algorithm choice, numbered like detail::AlgorithmChoice
 * (libs/Einsums/TensorAlgebra/include/Einsums/TensorAlgebra/Detail/Utilities.hpp:26-34). */
#define EIN_ALG_GENERIC 0
#define EIN_ALG_DOT 1
#define EIN_ALG_DIRECT 2
#define EIN_ALG_GER 3
#define EIN_ALG_GEMV 4
#define EIN_ALG_GEMM 5
#define EIN_ALG_INDETERMINATE 6

/* A strided tensor view. labels[i] is any small integer (e.g. a character code)
 * naming the index on axis i; equal labels on one tensor form a diagonal. */
typedef struct ein_tensor_desc {
    int32_t rank;
    int32_t labels[EIN_MAX_RANK];
    int64_t dims[EIN_MAX_RANK];
    int64_t strides[EIN_MAX_RANK]; /* in elements */
} ein_tensor_desc;

/* ---------------------------------------------------------------- runtime */
const char *ein_version(void);
const char *ein_last_error_string(void);
int ein_init(void);
/* Pooled device memory (256-byte aligned), replacing the per-tensor hipMalloc/hipFree of
 * GPUAllocator / BlockManager (libs/Einsums/GPUMemory/include/Einsums/GPUMemory/
 * GPUAllocator.hpp:53-180, libs/Einsums/BlockManager/src/BlockManager.cpp:66).
 * ein_free returns a block to the pool without waiting; like hipFree, it is safe with work
 * still queued: the block is handed out again only after the device has been synchronised
 * once after the free (ein_malloc does that when it needs the block).
 * ein_free_async is the stream-ordered form: the block is reused once the work queued on
 * `stream` at the time of the call has completed. */
int ein_malloc(void **ptr, size_t bytes);
int ein_free(void *ptr);
int ein_free_async(void *ptr, ein_stream_t stream);
/* Pre-size the per-stream scratch used by split-K partials, reductions and
 * split-precision slices; required before hipGraph capture of large calls (and before
 * capturing the first call on a stream). A scratch buffer used inside a capture is never
 * freed when a later, larger call grows it: the captured graph keeps pointing at it. */
int ein_workspace_reserve(ein_stream_t stream, size_t bytes);
size_t ein_workspace_size(ein_stream_t stream);
int ein_stream_synchronize(ein_stream_t stream);
/* hipGraph capture of a sequence of calls on one stream. */
int ein_graph_begin(ein_stream_t stream);
int ein_graph_end(ein_stream_t stream, void **graph_exec);
int ein_graph_launch(void *graph_exec, ein_stream_t stream);
int ein_graph_destroy(void *graph_exec);

/* ------------------------------------------------- BLAS-shaped leaves
 * Replace blas::gpu::gemm/gemv/dot/ger/scal/axpby/dirprod
 * (libs/Einsums/hipBLAS/include/Einsums/hipBLAS.hpp:64,133,546,622,668,720,1541). */
int ein_gemm_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                 const double *B, int64_t ldb, double beta, double *C, int64_t ldc, ein_stream_t stream);
int ein_gemm_f32(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda,
                 const float *B, int64_t ldb, float beta, float *C, int64_t ldc, ein_stream_t stream);
/* Strided-batched GEMM (batch index shared by A, B and C). Used by batched einsum
 * patterns; the reference loops over blocks (Dispatch.hpp:1264-1286). */
int ein_gemm_strided_batched_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                                 int64_t lda, int64_t strideA, const double *B, int64_t ldb, int64_t strideB, double beta,
                                 double *C, int64_t ldc, int64_t strideC, int64_t batch, ein_stream_t stream);
int ein_gemm_strided_batched_f32(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                                 int64_t lda, int64_t strideA, const float *B, int64_t ldb, int64_t strideB, float beta,
                                 float *C, int64_t ldc, int64_t strideC, int64_t batch, ein_stream_t stream);
int ein_gemv_f64(char trans, int64_t m, int64_t n, double alpha, const double *A, int64_t lda, const double *x, int64_t incx,
                 double beta, double *y, int64_t incy, ein_stream_t stream);
int ein_gemv_f32(char trans, int64_t m, int64_t n, float alpha, const float *A, int64_t lda, const float *x, int64_t incx,
                 float beta, float *y, int64_t incy, ein_stream_t stream);
/* result (device pointer) = beta*result + alpha*dot(x,y). Partial sums per workgroup, then a
 * second launch adds them in a fixed order, so repeated calls give identical bits.
 * Replaces blas::gpu::dot (hipBLAS.hpp:546). */
int ein_dot_f64(int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy, double beta,
                double *result, ein_stream_t stream);
int ein_dot_f32(int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy, float beta, float *result,
                ein_stream_t stream);
/* ein_dot_f64 for callers that cannot allocate device memory (code under hipGraph capture,
 * serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. Same
 * arguments, kernels and bits as ein_dot_f64; the partial sums live in a static device
 * array of EIN_DOT_NOALLOC_MAX_PARTS doubles (512 KiB) that each GPU the library runs on
 * holds until the process ends, so the call allocates nothing. Limit: unit strides up to
 * EIN_DOT_NOALLOC_MAX_N elements (longer: EIN_ERROR_UNSUPPORTED, nothing launched); other
 * strides any n. All calls on one GPU share that array: calls that overlap in time on one
 * GPU race (serialize them, e.g. on one stream). */
#define EIN_DOT_NOALLOC_MAX_PARTS 65536
#define EIN_DOT_NOALLOC_MAX_N (65536LL * 8192LL)
int ein_dot_f64_noalloc(int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy,
                        double beta, double *result, ein_stream_t stream);
int ein_ger_f64(int64_t m, int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy, double *A,
                int64_t lda, ein_stream_t stream);
int ein_ger_f32(int64_t m, int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy, float *A,
                int64_t lda, ein_stream_t stream);
/* z = alpha*x.*y + beta*z (the reference's dirprod is the beta=1 case). */
int ein_dirprod_f64(int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy, double beta,
                    double *z, int64_t incz, ein_stream_t stream);
int ein_dirprod_f32(int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy, float beta, float *z,
                    int64_t incz, ein_stream_t stream);
int ein_scale_f64(int64_t n, double alpha, double *x, int64_t incx, ein_stream_t stream);
int ein_scale_f32(int64_t n, float alpha, float *x, int64_t incx, ein_stream_t stream);
int ein_axpby_f64(int64_t n, double alpha, const double *x, int64_t incx, double beta, double *y, int64_t incy,
                  ein_stream_t stream);
int ein_axpby_f32(int64_t n, float alpha, const float *x, int64_t incx, float beta, float *y, int64_t incy,
                  ein_stream_t stream);

/* Measured GEMM configurations (the autotuner). The MFMA GEMM behind ein_gemm_* and every
 * GEMM-mapped einsum looks its problem up in a process-wide table keyed by the normalised
 * problem (element size, m, n, k, batch, whether A and B are contiguous along k); an entry
 * overrides the built-in tile heuristic. With tuning mode on, a GEMM whose key is absent
 * first times every candidate (tile 128/64 x 128/64; hybrid stream-K, data-parallel or
 * split-K over XCD-local contraction chunks; planned workgroups per CU) on its own
 * operands, writing a scratch C, and stores the
 * fastest (not while its stream is captured). No counterpart in the reference (its GEMMs
 * are vendor BLAS calls). The table is saved and loaded as text.
 * ein_gemm_tuning_mode returns the previous mode; ein_gemm_tuning_lookup writes
 * {bm, bn, schedule, blocks_per_cu} to config[0..3] (schedule 0 stream-K, 1 data-parallel,
 * 2 split-K). */
int ein_gemm_tuning_mode(int32_t on);
int ein_gemm_tuning_clear(void);
int64_t ein_gemm_tuning_count(void);
int ein_gemm_tuning_lookup(int32_t elem_bytes, int64_t m, int64_t n, int64_t k, int64_t batch, int32_t a_kcontig,
                           int32_t b_kcontig, int32_t *config);
int ein_gemm_tuning_save(const char *path);
int ein_gemm_tuning_load(const char *path);

/* Split-precision FP64 GEMM (Ozaki scheme): op(A), op(B) are cut into `nslices` integer
 * slices with per-row / per-column power-of-two scales, multiplied on MFMA with exact
 * accumulation and recombined in FP64.
 *   EIN_SLICE_BF16 / EIN_SLICE_FP16: 8-bit slices (nslices 3..6) on BF16/FP16 MFMA with FP32
 *     accumulation; relative representation error 2^-(8*nslices).
 *   EIN_SLICE_INT8: 7-bit slices (nslices 3..6) on int8 MFMA with int32 accumulation;
 *     relative representation error 2^-(7*nslices).
 * Same argument order and checks as ein_gemm_f64; no counterpart in the reference (its GEMM
 * leaf is hipblasDgemm). */
#define EIN_SLICE_BF16 0
#define EIN_SLICE_FP16 1
#define EIN_SLICE_INT8 2
/*   EIN_SLICE_CRT: nslices = number of moduli, 8..16 (Chinese remainder theorem, the
 *     integer modular scheme "Ozaki scheme II"): op(A) and op(B) are rounded to kb-bit
 *     integers per row / column, one exact int8 GEMM runs per modulus, and the product is
 *     reconstructed exactly from the residues; kb = ein_split_crt_bits(nslices, k) is the
 *     largest value with 4 k 2^(2 kb) <= prod(moduli), relative representation error 2^-kb
 *     (13 moduli: kb = 44 at k = 2304, 42 at k = 65536). */
#define EIN_SLICE_CRT 3
/* OR'd into EIN_SLICE_CRT: one iterative-refinement correction. The residuals
 * dA = A - rep(A) and dB = B - rep(B) of the kb-bit integer representation are formed in
 * FP64 and C += alpha*(dA*B + A*dB) runs as two more CRT GEMMs with 8 moduli, which takes the
 * result to FP64 rounding (about 1e-15) for about 1.5x the time of the plain CRT GEMM.
 * Rejected (EIN_ERROR_VALUE) with the slice types, whose error is dominated by the dropped
 * slice cross products rather than by the representation of A and B. */
#define EIN_SPLIT_REFINE 0x100
int ein_gemm_split_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                       const double *B, int64_t ldb, double beta, double *C, int64_t ldc, int32_t nslices,
                       int32_t slice_type, ein_stream_t stream);
/* Bits kept per operand element by EIN_SLICE_CRT with nmoduli moduli at contraction length k
 * (-1 for an invalid count or k < 0). nslices = 0 with EIN_SLICE_CRT picks the fewest moduli
 * that keep at least 40 bits (relative representation error <= 2^-40 per element), which
 * ein_split_crt_moduli(k) reports (12 at k = 2304, 13 at k = 12288 and 65536); nmoduli = 0
 * here gives that choice's bits. No device work. */
int ein_split_crt_bits(int32_t nmoduli, int64_t k);
int ein_split_crt_moduli(int64_t k);
/* C = alpha*op(A)*op(B) + beta*C with A, B on the device and C (column-major, ldc) in host
 * memory, for outputs larger than HBM (the out-of-core case of the v^4 CCD intermediates,
 * which the reference keeps as a DiskTensor, libs/Einsums/Tensor/include/Einsums/
 * DiskTensor.hpp:200-306). Column panels of about panel_bytes (0 = 256 MiB) stream through
 * two device buffers; uploads, GEMMs and downloads of consecutive panels overlap. nslices > 0
 * runs each panel's GEMM in split precision (slice_type EIN_SLICE_*). Returns when C is back
 * in host memory; host C should be pinned (hipHostMalloc). Calls are serialised internally. */
int ein_gemm_host_c_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                        int64_t lda, const double *B, int64_t ldb, double beta, double *C, int64_t ldc,
                        size_t panel_bytes, int32_t nslices, int32_t slice_type, ein_stream_t stream);
int ein_gemm_host_c_f32(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                        int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc, size_t panel_bytes,
                        ein_stream_t stream);
/* Tensors on disk: one chunked HDF5 dataset per tensor, its dims in row-major order, as the
 * reference's DiskTensor keeps them (libs/Einsums/Tensor/include/Einsums/Tensor/
 * DiskTensor.hpp:200-306: created on construction when absent, opened otherwise; get()/put()
 * move the data). libhdf5 is loaded at first use ($EINSUMS_AMD_HDF5, libhdf5.so, the conda
 * path); without it these return EIN_ERROR_UNSUPPORTED. ein_disk_create_* creates the file
 * and the dataset when absent (chunks of one leading slice, cut to about 16 MiB) and checks
 * rank, extents and element type when present; ein_disk_write_* does the same, then writes
 * all of data; ein_disk_read_* reads an existing dataset. All HDF5 calls are serialised. */
int ein_disk_create_f64(const char *path, const char *dataset, int32_t rank, const int64_t *dims);
int ein_disk_create_f32(const char *path, const char *dataset, int32_t rank, const int64_t *dims);
int ein_disk_write_f64(const char *path, const char *dataset, int32_t rank, const int64_t *dims, const double *data);
int ein_disk_write_f32(const char *path, const char *dataset, int32_t rank, const int64_t *dims, const float *data);
int ein_disk_read_f64(const char *path, const char *dataset, int32_t rank, const int64_t *dims, double *data);
int ein_disk_read_f32(const char *path, const char *dataset, int32_t rank, const int64_t *dims, float *data);
/* ein_gemm_host_c_* with C in a disk tensor: C (column-major m x n) is the row-major view
 * [n][m] of dataset `dataset` in `path`, whose leading dims multiply to n and the remaining
 * ones to m (the Wabef intermediate as a (v, v, v, v) DiskTensor: n = m = v^2). Panels of
 * whole leading slices (about panel_bytes, 0 = 256 MiB) go disk -> pinned host -> device ->
 * GEMM -> host -> disk; the disk reads and writes of neighbouring panels overlap the device
 * work. beta == 0 skips the reads. nslices > 0 runs each panel's GEMM in split precision
 * (slice_type EIN_SLICE_*). Returns when C is on disk; calls are serialised. */
int ein_gemm_disk_c_f64(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                        int64_t lda, const double *B, int64_t ldb, double beta, const char *path, const char *dataset,
                        size_t panel_bytes, int32_t nslices, int32_t slice_type, ein_stream_t stream);
int ein_gemm_disk_c_f32(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                        int64_t lda, const float *B, int64_t ldb, float beta, const char *path, const char *dataset,
                        size_t panel_bytes, ein_stream_t stream);
/* ein_einsum_f64 with the GEMM-mapped case run by the split-precision GEMM; other
 * patterns run exactly as in ein_einsum_f64. */
int ein_einsum_split_f64(int32_t nslices, int32_t slice_type, double c_prefactor, const ein_tensor_desc *C, double *c,
                         double ab_prefactor, const ein_tensor_desc *A, const double *a, const ein_tensor_desc *B,
                         const double *b, int32_t *algorithm_used, ein_stream_t stream);

/* Fock build: F = beta*F + alpha_j*J + alpha_k*K with J(a,b) = sum_cd g(a,b,c,d) D(c,d) and
 * K(a,c) = sum_bd g(a,b,c,d) D(b,d), g row-major n^4, D and F row-major n x n (ld >= n),
 * n <= 128. Both matrices come from one pass over g, where the reference makes two einsum
 * calls that each read it (QChemTests.cpp:299-304, profile_einsums.cpp:23-30); no single
 * reference function corresponds. Deterministic (fixed summation order). */
int ein_fock_jk_f64(int64_t n, const double *g, const double *D, int64_t ldd, double alpha_j, double alpha_k,
                    double beta, double *F, int64_t ldf, ein_stream_t stream);

/* ------------------------------------------------------------- tensors */
/* Whole-view ops on a strided tensor: C = alpha*C (alpha==0 writes zeros), C = value. */
int ein_tensor_scale_f64(const ein_tensor_desc *C, double *c, double alpha, ein_stream_t stream);
int ein_tensor_scale_f32(const ein_tensor_desc *C, float *c, float alpha, ein_stream_t stream);

/* C = beta*C + alpha*permute(A): C axis i is A axis perm[i].  Replaces
 * detail::gpu_permute / permute_kernel (DevicePermute.hpp:25-73) and the LibreTT path
 * (TensorAlgebra/src/DevicePermute.hip:20-59). alpha==1, beta==0 is a bit-exact copy. */
int ein_permute_f64(int32_t rank, const int64_t *dims_A, const int32_t *perm, double alpha, const double *A,
                    const int64_t *strides_A, double beta, double *C, const int64_t *strides_C, ein_stream_t stream);
int ein_permute_f32(int32_t rank, const int64_t *dims_A, const int32_t *perm, float alpha, const float *A,
                    const int64_t *strides_A, float beta, float *C, const int64_t *strides_C, ein_stream_t stream);

/* Generic contraction C = c_prefactor*C + ab_prefactor * sum_links A*B over any label
 * pattern (Hadamard diagonals, traces, broadcasts, rank-0 C). Replaces the device
 * einsum_generic_algorithm launcher (GPUTensorAlgebra.hpp:200-342) with atomic-free
 * kernels: MFMA strided-gather GEMM, gather GEMV, or a deterministic reduction. */
int ein_contract_f64(double c_prefactor, const ein_tensor_desc *C, double *c, double ab_prefactor, const ein_tensor_desc *A,
                     const double *a, const ein_tensor_desc *B, const double *b, ein_stream_t stream);
int ein_contract_f32(float c_prefactor, const ein_tensor_desc *C, float *c, float ab_prefactor, const ein_tensor_desc *A,
                     const float *a, const ein_tensor_desc *B, const float *b, ein_stream_t stream);

/* Route of the generic contraction's GEMM-sized cases on the calling thread (it holds
 * until changed): EIN_ROUTE_AUTO picks by a cost model, EIN_ROUTE_GATHER takes the MFMA
 * strided-gather kernel, EIN_ROUTE_TTGT packs the operands that do not fold into a matrix
 * with the permute and runs the dense MFMA GEMM (transpose-transpose-GEMM). The reference
 * has one generic route only (GPUTensorAlgebra.hpp:200-342); this is for autotuners and
 * tests. Returns the previous route, or -1 for an unknown value (route unchanged). */
#define EIN_ROUTE_AUTO 0
#define EIN_ROUTE_GATHER 1
#define EIN_ROUTE_TTGT 2
int32_t ein_set_contraction_route(int32_t route);

/* Adjoint of C = alpha * contract(A, B): the gradient G of the operand `wrt` (0 = A,
 * 1 = B) given the cotangent dC, G = beta*G + alpha * contract(dC, other operand), with G
 * labelled like the operand. Labels repeated in the operand make G a diagonal (written
 * once per element, no atomics, the rest of G scaled by beta); labels summed in the
 * operand alone broadcast. No counterpart in the reference (it has no autograd). */
int ein_einsum_adjoint_f64(int32_t wrt, double alpha, const ein_tensor_desc *C, const double *dC, const ein_tensor_desc *A,
                           const double *a, const ein_tensor_desc *B, const double *b, double beta,
                           const ein_tensor_desc *G, double *g, int32_t *algorithm_used, ein_stream_t stream);
int ein_einsum_adjoint_f32(int32_t wrt, float alpha, const ein_tensor_desc *C, const float *dC, const ein_tensor_desc *A,
                           const float *a, const ein_tensor_desc *B, const float *b, float beta, const ein_tensor_desc *G,
                           float *g, int32_t *algorithm_used, ein_stream_t stream);

/* Grouped contractions: many independent generic contractions (blocks of a block-sparse
 * einsum, batches of uneven GEMMs) planned once and run as persistent launches -- one per
 * wave of items whose outputs differ (items writing the same C pointer go to later waves,
 * in list order). Plan creation copies descriptors to device memory (synchronous);
 * execution is launches only, so it can be captured into a hipGraph. The autotuner times
 * the candidate tile heights on scratch operands shaped like the plan and keeps the
 * faster. Replaces the per-block loop of BlockAlgebra.hpp:23-52 / TileAlgebra.hpp:18-93. */
typedef struct ein_contraction_item {
    ein_tensor_desc C, A, B;
    void *c;
    const void *a;
    const void *b;
    double c_prefactor, ab_prefactor;
} ein_contraction_item;
int ein_grouped_plan_create(int32_t elem_bytes, const ein_contraction_item *items, int32_t count, void **plan);
int ein_grouped_plan_execute(void *plan, ein_stream_t stream);
int ein_grouped_plan_autotune(void *plan, ein_stream_t stream);
int ein_grouped_plan_info(void *plan, int32_t *tile_rows, int32_t *waves, int64_t *tiles);
int ein_grouped_plan_destroy(void *plan);

/* Runtime pattern analysis with the precedence of compile_plan
 * (libs/EinsumsPy/TensorAlgebra/src/PyTensorAlgebra.cpp:169-300); writes EIN_ALG_*. */
int ein_einsum_classify(const ein_tensor_desc *C, const ein_tensor_desc *A, const ein_tensor_desc *B, int32_t *algorithm);
/* Full einsum: classify, fold to 2-D views, run the matching kernel, fall back to the
 * generic contraction when a view does not fold (Dispatch.hpp:859-957 semantics).
 * *algorithm_used (may be NULL) receives the EIN_ALG_* that actually ran. */
int ein_einsum_f64(double c_prefactor, const ein_tensor_desc *C, double *c, double ab_prefactor, const ein_tensor_desc *A,
                   const double *a, const ein_tensor_desc *B, const double *b, int32_t *algorithm_used, ein_stream_t stream);
int ein_einsum_f32(float c_prefactor, const ein_tensor_desc *C, float *c, float ab_prefactor, const ein_tensor_desc *A,
                   const float *a, const ein_tensor_desc *B, const float *b, int32_t *algorithm_used, ein_stream_t stream);
/* Execute with a pattern already chosen by the caller (the C++ headers decide it at
 * compile time, Dispatch.hpp:894-955); views that do not fold run the generic kernel
 * and *algorithm_used reports EIN_ALG_GENERIC. EIN_ALG_INDETERMINATE classifies first. */
int ein_einsum_execute_f64(int32_t algorithm, double c_prefactor, const ein_tensor_desc *C, double *c, double ab_prefactor,
                           const ein_tensor_desc *A, const double *a, const ein_tensor_desc *B, const double *b,
                           int32_t *algorithm_used, ein_stream_t stream);
int ein_einsum_execute_f32(int32_t algorithm, float c_prefactor, const ein_tensor_desc *C, float *c, float ab_prefactor,
                           const ein_tensor_desc *A, const float *a, const ein_tensor_desc *B, const float *b,
                           int32_t *algorithm_used, ein_stream_t stream);

/* Dense solvers for the small matrices of the SCF loop around the path, on one workgroup,
 * asynchronous: LAPACK's `info` goes to the device int *info (0 = success) instead of the
 * return value, which reports argument errors only.
 * syev replaces blas::syev(job, uplo, n, a, lda, w, work, lwork)
 * (libs/Einsums/BLAS/include/Einsums/BLAS.hpp:306; linear_algebra::syev, LinearAlgebra.hpp:362).
 * Column-major; uplo names the triangle read; job 'V' overwrites A with orthonormal
 * eigenvectors (columns), 'N' leaves A; W gets the eigenvalues ascending. Cyclic Jacobi,
 * n <= 1024; *info = 1 if 60 sweeps do not converge. Workspace 2*n*n elements from the pool. */
int ein_syev_f64(char job, char uplo, int64_t n, double *A, int64_t lda, double *W, int32_t *info, ein_stream_t stream);
int ein_syev_f32(char job, char uplo, int64_t n, float *A, int64_t lda, float *W, int32_t *info, ein_stream_t stream);
/* gesv replaces blas::gesv(n, nrhs, a, lda, ipiv, b, ldb) (BLAS.hpp:529; linear_algebra::gesv,
 * LinearAlgebra.hpp:495): A = P L U with partial pivoting (ipiv 1-based, device int64),
 * B overwritten with the solution; *info = i > 0 if U(i,i) is exactly zero (the factorization
 * is completed and B is left as it was). Column-major, n <= 4096. */
int ein_gesv_f64(int64_t n, int64_t nrhs, double *A, int64_t lda, int64_t *ipiv, double *B, int64_t ldb, int32_t *info,
                 ein_stream_t stream);
int ein_gesv_f32(int64_t n, int64_t nrhs, float *A, int64_t lda, int64_t *ipiv, float *B, int64_t ldb, int32_t *info,
                 ein_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* EINSUMS_AMD_H */
