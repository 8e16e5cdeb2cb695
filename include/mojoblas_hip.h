/*
 * mojoblas_hip: C-ABI of the MI355X (gfx950) level-3 BLAS hot path (and the
 * level-2 gemv / trsv next to it).
 *
 * Drop-in boundary for mojoBLAS's src/level3 Mojo functions
 * (docs/reference.md:264-327).  A Mojo shim keeps the reference signature
 * and calls the matching entry here through `external_call`; the Python
 * layer (mojoblas_amd) binds the same symbols with ctypes.
 *
 * Conventions (all entries):
 *  - Column-major; element (i,j) of X is X[i + j*ldx]; dimensions int64.
 *  - Flags are single chars with the reference's accepted spellings
 *    ('N'/'n', 'T'/'t', 'C'/'c' [= 'T' for real data], 'L'/'l', 'R'/'r',
 *    'U'/'u').  The reference compares whole strings (gemm.mojo:180-197);
 *    callers pass s[0] only when len(s) == 1 and an invalid char otherwise,
 *    which reproduces the same info code.
 *  - Return value: 0 = success; > 0 = Netlib info (the 1-based index of the
 *    first invalid argument, numbered as in the reference's validation
 *    blocks), in which case no output is touched and no GPU work is issued;
 *    < 0 = runtime error (MBH_E_*), message in mbh_last_error().
 *  - `_host` entries take host pointers (what the reference's BLASPtr holds,
 *    src/type_aliases.mojo:15-17): they copy to HBM, run, copy back and
 *    synchronise; the returned buffers are final.
 *  - `_dev` entries take device pointers and a hipStream_t passed as void*
 *    (NULL = the default stream); they are stream-ordered and asynchronous.
 *    Scratch they need comes from a grow-only workspace kept per (device,
 *    stream); a call holds its stream's workspace from validation until its
 *    last kernel is queued, so calls from several host threads on one stream
 *    queue one after another.  A call captured into a hipGraph must not need
 *    more workspace than the stream already holds (run it once eagerly
 *    first); otherwise it returns MBH_E_ARG (a gemm before queueing
 *    anything; a level-3 routine possibly after part of its work, so such a
 *    capture is to be discarded).
 *    Workspace handed out during capture stays allocated until
 *    mbh_workspace_release or mbh_stream_destroy on that stream.
 *  - `_host` entries are synchronous and thread-safe: each calling thread
 *    stages through its own stream and device buffers.
 *  - Quick returns follow BLAS: m==0 or n==0 returns at once; alpha==0 or
 *    k==0 only scales C by beta (beta==0 writes zeros without reading C).
 */
#ifndef MOJOBLAS_HIP_H
#define MOJOBLAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic used for an fp32 gemm (and the gemms inside the level-3
 * recasts). */
enum mbh_precision {
    MBH_PREC_FP32 = 0,     /* fp32 MFMA (v_mfma_f32_32x32x2_f32), fp32 accumulate */
    MBH_PREC_BF16X3 = 1,   /* split x = hi + lo (bf16 each); hi*hi + hi*lo + lo*hi
                              on bf16 MFMA, fp32 accumulate */
    MBH_PREC_FP8 = 2,      /* opt-in low accuracy: e4m3 with one fp32 scale per
                              128x128 tile of A and of B, fp32 accumulate */
    MBH_PREC_BF16 = 3      /* plain bf16 operands, fp32 accumulate (batched
                              "bf16 input" mode) */
};

#define MBH_E_HIP (-1)        /* a HIP runtime call failed */
#define MBH_E_ALLOC (-2)      /* device or host allocation failed */
#define MBH_E_PRECISION (-3)  /* precision value not valid for this entry */
#define MBH_E_ARG (-4)        /* argument outside what the entry supports
                                 (not a Netlib info case) */
#define MBH_E_SINGULAR (-5)   /* reserved */

/* Library */
const char* mbh_version(void);
const char* mbh_last_error(void);                 /* thread-local message */
int mbh_set_default_precision(int precision);     /* for _host entries; returns the old value */
int mbh_get_default_precision(void);
/* Precision of the _host entries on the calling thread only (no counterpart
 * in the reference, whose signatures carry no precision): -1 clears it, so
 * the thread follows the library default again.  Returns the thread's old
 * setting (-1 if none) or MBH_E_PRECISION. */
int mbh_set_thread_precision(int precision);
int mbh_device_sync(void);
/* Test hook for the error paths (no counterpart in the reference): the
 * (n+1)-th internal device allocation after this call -- workspace growth,
 * host staging, batched plans -- fails once as if the device were out of
 * memory (n = 0: the next one); n < 0 turns the hook off.  The failing call
 * returns MBH_E_HIP or MBH_E_ALLOC before queueing anything and leaves its
 * outputs untouched.  Returns 0. */
int mbh_debug_fail_allocation(long n);
/* Opt-in timing for measurement harnesses (no counterpart in the reference,
 * which has no GPU kernels to time): while enabled, every
 * split-precision (BF16x3 / bf16 / FP8) gemm records HIP events on its launch
 * stream before its pack kernels, before its main kernel and after it.
 * timing_read waits for the recorded events and returns out[4] =
 * {main-kernel ms summed, launches, pack ms summed, launches}; reset != 0
 * starts a new tally.  enable returns the previous state. */
int mbh_timing_enable(int on);
int mbh_timing_read(double* out, int reset);

/* ---- gemm: C := alpha*op(A)*op(B) + beta*C ------------------------------
 * Replaces gemm[dtype] (src/level3/gemm.mojo:56-124) for every transpose
 * variant; info codes as gemm.mojo:179-213 (1,2,3,4,5,8,10,13). */
int mbh_sgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                   const float* a, int64_t lda, const float* b, int64_t ldb, float beta,
                   float* c, int64_t ldc);
int mbh_dgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha,
                   const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                   double* c, int64_t ldc);
int mbh_sgemm_dev(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                  const float* a, int64_t lda, const float* b, int64_t ldb, float beta,
                  float* c, int64_t ldc, int precision, void* stream);
int mbh_dgemm_dev(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha,
                  const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                  double* c, int64_t ldc, void* stream);

/* ---- gemm without allocation (no counterpart in the reference) ----------
 * For callers that cannot allocate device memory: code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments, checks and result (bit for bit) as mbh_sgemm_dev, but the
 * BF16x3 / FP8 pack planes live in a static device array of
 * MBH_NOALLOC_SCRATCH_BYTES per GPU (held from the first call on that GPU
 * until the process ends) instead of the stream's workspace; the entry calls
 * no allocation routine and never synchronises.  A product whose planes need
 * more (BF16x3 takes 4 bytes per element of op(A) and op(B)^T padded to
 * whole tiles: square sizes up to about 2800 fit) returns MBH_E_ARG before
 * queueing anything.  All calls on one GPU share the array: calls that
 * overlap on one GPU (several streams or threads) race, so the caller
 * orders them.  The stream must belong to the current device. */
#define MBH_NOALLOC_SCRATCH_BYTES (64ull * 1024ull * 1024ull)
int mbh_sgemm_dev_noalloc(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                          const float* a, int64_t lda, const float* b, int64_t ldb, float beta,
                          float* c, int64_t ldc, int precision, void* stream);

/* ---- 16-bit gemm: C := alpha*op(A)*op(B) + beta*C on half (h) or bfloat16
 * (b) data.  Replaces gemm[DType.float16] / gemm[DType.bfloat16]
 * (src/level3/gemm.mojo:56-124, whose non-f32 instantiations run the
 * generic v3 / v7 / naive kernels); info codes as for sgemm.  Elements are
 * passed as their 16-bit patterns (the Mojo shim bitcasts its
 * UnsafePointer[Float16] / [BFloat16]); alpha and beta as float (the shim
 * passes Float32(alpha), exact for every 16-bit value).  Products run on the
 * 16-bit MFMA (v_mfma_f32_32x32x16_f16 / _bf16) with fp32 accumulation;
 * alpha*acc + beta*C is formed in fp32 and rounded once, to nearest even,
 * into C's format. */
int mbh_hgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                   const uint16_t* a, int64_t lda, const uint16_t* b, int64_t ldb, float beta,
                   uint16_t* c, int64_t ldc);
int mbh_bgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                   const uint16_t* a, int64_t lda, const uint16_t* b, int64_t ldb, float beta,
                   uint16_t* c, int64_t ldc);
int mbh_hgemm_dev(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                  const uint16_t* a, int64_t lda, const uint16_t* b, int64_t ldb, float beta,
                  uint16_t* c, int64_t ldc, void* stream);
int mbh_bgemm_dev(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                  const uint16_t* a, int64_t lda, const uint16_t* b, int64_t ldb, float beta,
                  uint16_t* c, int64_t ldc, void* stream);

/* ---- trsm: B := alpha*inv(op(A))*B (side L) or alpha*B*inv(op(A)) (side R)
 * Replaces trsm[dtype] (src/level3/trsm.mojo:20-275; info codes :62-89:
 * 1,2,3,4,5,6,9,11).  BLAS semantics, including side R with trans T (the
 * reference solves the untransposed system there, trsm.mojo:223-273).
 * fp32: the 128x128 diagonal blocks are inverted and applied in fp64.
 * refine = 0: B is solved in place in fp32, `precision` selecting the gemm of
 * the off-diagonal updates.  refine = 1 (fp32 data to fp32 accuracy): the
 * solve runs on fp64 copies of alpha*B and of A's triangle with fp64 updates
 * and X is rounded to fp32 once (`precision` has no effect there).  The _host
 * entry uses refine = 1. */
int mbh_strsm_host(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float alpha,
                   const float* a, int64_t lda, float* b, int64_t ldb);
int mbh_dtrsm_host(char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha,
                   const double* a, int64_t lda, double* b, int64_t ldb);
int mbh_strsm_dev(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float alpha,
                  const float* a, int64_t lda, float* b, int64_t ldb, int precision, int refine,
                  void* stream);
int mbh_dtrsm_dev(char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha,
                  const double* a, int64_t lda, double* b, int64_t ldb, void* stream);

/* ---- trmm: B := alpha*op(A)*B (side L) or alpha*B*op(A) (side R) --------
 * Replaces trmm[dtype] (src/level3/trmm.mojo:20-305; info codes :62-89);
 * alpha == 0 zeroes B without reading A (trmm.mojo:94-98). */
int mbh_strmm_host(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float alpha,
                   const float* a, int64_t lda, float* b, int64_t ldb);
int mbh_dtrmm_host(char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha,
                   const double* a, int64_t lda, double* b, int64_t ldb);
int mbh_strmm_dev(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float alpha,
                  const float* a, int64_t lda, float* b, int64_t ldb, int precision, void* stream);
int mbh_dtrmm_dev(char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha,
                  const double* a, int64_t lda, double* b, int64_t ldb, void* stream);

/* ---- syrk: C := alpha*A*A^T + beta*C (trans N) or alpha*A^T*A + beta*C ---
 * Replaces syrk[dtype] (src/level3/syrk.mojo:20-200; info codes :60-83:
 * 1,2,3,4,7,10); only the `uplo` triangle of C is read or written. */
int mbh_ssyrk_host(char uplo, char trans, int64_t n, int64_t k, float alpha, const float* a,
                   int64_t lda, float beta, float* c, int64_t ldc);
int mbh_dsyrk_host(char uplo, char trans, int64_t n, int64_t k, double alpha, const double* a,
                   int64_t lda, double beta, double* c, int64_t ldc);
int mbh_ssyrk_dev(char uplo, char trans, int64_t n, int64_t k, float alpha, const float* a,
                  int64_t lda, float beta, float* c, int64_t ldc, int precision, void* stream);
int mbh_dsyrk_dev(char uplo, char trans, int64_t n, int64_t k, double alpha, const double* a,
                  int64_t lda, double beta, double* c, int64_t ldc, void* stream);

/* ---- syr2k: C := alpha*(A*B^T + B*A^T) + beta*C (trans N) or
 * alpha*(A^T*B + B^T*A) + beta*C.  Replaces syr2k[dtype]
 * (src/level3/syr2k.mojo:20-190; info codes 1,2,3,4,7,9,12). */
int mbh_ssyr2k_host(char uplo, char trans, int64_t n, int64_t k, float alpha, const float* a,
                    int64_t lda, const float* b, int64_t ldb, float beta, float* c, int64_t ldc);
int mbh_dsyr2k_host(char uplo, char trans, int64_t n, int64_t k, double alpha, const double* a,
                    int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc);
int mbh_ssyr2k_dev(char uplo, char trans, int64_t n, int64_t k, float alpha, const float* a,
                   int64_t lda, const float* b, int64_t ldb, float beta, float* c, int64_t ldc,
                   int precision, void* stream);
int mbh_dsyr2k_dev(char uplo, char trans, int64_t n, int64_t k, double alpha, const double* a,
                   int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc,
                   void* stream);

/* ---- symm: C := alpha*A*B + beta*C (side L) or alpha*B*A + beta*C (side R),
 * A symmetric, read from its `uplo` triangle.  Replaces symm[dtype]
 * (src/level3/symm.mojo:20-241; info codes :68-86: 1,2,3,4,7,9,12); the
 * full alpha*temp2 term is applied even where B(l,j) == 0 (the reference
 * skips it, symm.mojo:114,133). */
int mbh_ssymm_host(char side, char uplo, int64_t m, int64_t n, float alpha, const float* a,
                   int64_t lda, const float* b, int64_t ldb, float beta, float* c, int64_t ldc);
int mbh_dsymm_host(char side, char uplo, int64_t m, int64_t n, double alpha, const double* a,
                   int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc);
int mbh_ssymm_dev(char side, char uplo, int64_t m, int64_t n, float alpha, const float* a,
                  int64_t lda, const float* b, int64_t ldb, float beta, float* c, int64_t ldc,
                  int precision, void* stream);
int mbh_dsymm_dev(char side, char uplo, int64_t m, int64_t n, double alpha, const double* a,
                  int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc,
                  void* stream);

/* ---- level 2: gemv  y := alpha*op(A)*x + beta*y -------------------------
 * Replaces gemv[dtype](trans, m, n, alpha, a, lda, x, incx, beta, y, incy),
 * src/level2/gemv.mojo:23-212.  trans is 'N', 'T' or 'C' (uppercase only:
 * the reference compares against "N"/"T"/"C", gemv.mojo:77); info 1,2,3,6,8,11
 * (gemv.mojo:76-91).  x and y are strided (incx, incy != 0; negative
 * increments walk the vector from its far end).  Deterministic fixed-order
 * sums; the arithmetic is that of the data type.  beta == 0 does not read y. */
int mbh_sgemv_host(char trans, int64_t m, int64_t n, float alpha, const float* a, int64_t lda,
                   const float* x, int64_t incx, float beta, float* y, int64_t incy);
int mbh_dgemv_host(char trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
                   const double* x, int64_t incx, double beta, double* y, int64_t incy);
int mbh_sgemv_dev(char trans, int64_t m, int64_t n, float alpha, const float* a, int64_t lda,
                  const float* x, int64_t incx, float beta, float* y, int64_t incy, void* stream);
int mbh_dgemv_dev(char trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
                  const double* x, int64_t incx, double beta, double* y, int64_t incy, void* stream);

/* ---- level 2: trsv  x := inv(op(A))*x, A triangular -----------------------
 * Replaces trsv[dtype](uplo, trans, diag, n, a, lda, x, incx),
 * src/level2/trsv.mojo:20-214; info 1,2,3,4,6,8 (trsv.mojo:57-80).  One
 * kernel: blocked substitution over 128-row blocks with fp64 residual
 * accumulation and fp64 diagonal-block inverses (every arithmetic step in
 * fp64, results rounded to the data type).  The host entry returns
 * MBH_E_HIP if a block's bounded wait for an earlier block ran out. */
int mbh_strsv_host(char uplo, char trans, char diag, int64_t n, const float* a, int64_t lda,
                   float* x, int64_t incx);
int mbh_dtrsv_host(char uplo, char trans, char diag, int64_t n, const double* a, int64_t lda,
                   double* x, int64_t incx);
int mbh_strsv_dev(char uplo, char trans, char diag, int64_t n, const float* a, int64_t lda,
                  float* x, int64_t incx, void* stream);
int mbh_dtrsv_dev(char uplo, char trans, char diag, int64_t n, const double* a, int64_t lda,
                  double* x, int64_t incx, void* stream);
/* Status of the last trsv _dev call queued on `stream` (no counterpart in the
 * reference, whose trsv cannot time out): synchronises the stream and sets
 * *failed nonzero if a block's bounded wait for an earlier block ran out, in
 * which case that call's x is not the solution.  Without this query such a
 * give-up in a _dev call is silent. */
int mbh_trsv_status(void* stream, int* failed);

/* ---- variable-size batched gemm (no counterpart in the reference) --------
 * count independent problems C_p := alpha*A_p*B_p + beta*C_p (each with the
 * semantics of gemm("N","N", m_p, n_p, k_p, ...), gemm.mojo:56-124), run as
 * persistent kernels, 128x128 tiles first, then 64x64.  When every problem
 * with k > 0 has at least four elements in each operand, each workgroup
 * walks a fixed list of C tiles (problems spread over the XCDs, then over
 * each XCD's workgroups, longest-processing-time first) with four waves
 * loading the operands two k-tiles ahead and four running the MFMAs;
 * otherwise workgroups pull tiles from work queues (per tile class one queue
 * per XCD holding whole problems, drained by that XCD's workgroups first,
 * then by any).  Both give the same results.
 *  plan_create : host arrays of shapes and leading dimensions; validates every
 *                problem (MBH_E_ARG names the first bad one), assigns tile
 *                classes from the tuning table, builds the workgroup lists or
 *                queues and stores the plan in HBM (allocates; once per shape
 *                list).  MBH_VBATCH_KERNEL=lean in the environment selects
 *                the work-queue kernels for every plan (for comparisons).
 *  exec        : device arrays of count device pointers; allocates nothing and
 *                never synchronises the host, so it can be captured in a
 *                hipGraph and replayed.  precision MBH_PREC_FP32 (fp32 MFMA) or
 *                MBH_PREC_BF16 (bf16 operands packed on the fly, fp32
 *                accumulate).  Calls sharing one plan must not overlap.
 *  plan_info   : out[8] = {count, tiles, workgroups, flops, problems on
 *                64x64 tiles, on 128x128 tiles, kernel (1: warp-specialized
 *                tiles on fixed per-workgroup lists, 0: work queues, 2: the
 *                two-class kernels with 64-bit offsets, taken when an
 *                operand spans 2^30 elements or more), precision}.
 *  set_tuning  : for one precision, the class table (36 ints,
 *                [bucket(m)][bucket(n)], buckets <=16,32,64,128,256,larger;
 *                0 = 64x64 tiles, 1 = 128x128; NULL = compiled default) and
 *                workgroups per CU (0 = default; each tile class is capped at
 *                what its kernel keeps resident) for plans created afterwards. */
typedef struct mbh_vbatch_plan_s* mbh_vbatch_plan;
int mbh_vbatch_plan_create(int64_t count, const int64_t* m, const int64_t* n, const int64_t* k,
                           const int64_t* lda, const int64_t* ldb, const int64_t* ldc, int precision,
                           mbh_vbatch_plan* plan);
int mbh_sgemm_vbatched_exec(mbh_vbatch_plan plan, float alpha, const float* const* a,
                            const float* const* b, float beta, float* const* c, void* stream);
/* exec with bf16 operands (no counterpart in the reference): a[p] / b[p]
 * point at bf16 bit patterns (uint16, column-major with the plan's lda /
 * ldb, any 2-byte alignment), C stays fp32.  Needs a MBH_PREC_BF16 plan on
 * the warp-specialized kernels (plan_info kernel = 1) whose problems with
 * k > 0 have at least 8 elements per operand (MBH_E_ARG otherwise).  The
 * results are bit-identical to mbh_sgemm_vbatched_exec on fp32 operands
 * that round (nearest even) to the same bf16 values; half the operand bytes
 * are read.  Same capture and overlap rules as exec. */
int mbh_bgemm_vbatched_exec(mbh_vbatch_plan plan, float alpha, const void* const* a, const void* const* b,
                            float beta, float* const* c, void* stream);
int mbh_vbatch_plan_info(mbh_vbatch_plan plan, double* out);
int mbh_vbatch_plan_destroy(mbh_vbatch_plan plan);
int mbh_vbatch_set_tuning(int precision, const int* class_table, int wgs_per_cu);
/* One-shot convenience: plan_create + exec + plan_destroy (synchronous). */
int mbh_sgemm_vbatched_dev(int64_t count, const int64_t* m, const int64_t* n, const int64_t* k,
                           float alpha, const float* const* a, const int64_t* lda,
                           const float* const* b, const int64_t* ldb, float beta, float* const* c,
                           const int64_t* ldc, int precision, void* stream);

/* ---- adjoints (no counterpart in the reference) --------------------------
 * Reverse-mode derivatives on the same kernels, for callers that
 * differentiate through gemm / trsm (the Python layer's
 * torch.autograd.Function binds these).
 *  sgemm_backward: for C = alpha*op(A)*op(B) + beta*C0 and upstream dC,
 *    dA = gradient w.r.t. A as stored (alpha*dC*op(B)^T, or its transpose for
 *    transa T), dB likewise (alpha*op(A)^T*dC or its transpose); either output
 *    may be NULL.  Shapes and flags are the forward call's (validated with the
 *    gemm info codes); dC is m x n.
 *  strsm_backward: for X = alpha*inv(op(A))*B (side L) or alpha*B*inv(op(A))
 *    (side R) and upstream dX: dB = alpha*G with G = op(A)^-T dX (L) or
 *    dX op(A)^-T (R); dA (optional) = -G X^T (L) or -X^T G (R) as a gradient
 *    of op(A), returned for A as stored: only its `uplo` triangle is nonzero
 *    and its diagonal is zero when diag is U.  The solve for G runs in fp64
 *    (refine = 1 above). */
int mbh_sgemm_backward(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                       const float* a, int64_t lda, const float* b, int64_t ldb, const float* dc,
                       int64_t lddc, float* da, int64_t ldda, float* db, int64_t lddb, int precision,
                       void* stream);
int mbh_strsm_backward(char side, char uplo, char transa, char diag, int64_t m, int64_t n,
                       float alpha, const float* a, int64_t lda, const float* x, int64_t ldx,
                       const float* dx, int64_t lddx, float* db, int64_t lddb, float* da,
                       int64_t ldda, int precision, void* stream);

/* ---- device-resident buffers (no counterpart in the reference) ----------
 * The reference's functions take host pointers (docs/reference.md:264-327),
 * so every _host call moves its operands over PCIe.  These entries let a
 * caller without its own HIP bindings (a Mojo program) keep operands in HBM
 * across calls and pass them to the _dev entries.
 *  dev_alloc / dev_free   : device memory (hipMalloc / hipFree); bytes == 0
 *                           gives NULL; freeing NULL is a no-op.
 *  host_alloc / host_free : pinned host memory for fast copies.
 *  set_matrix / get_matrix: copy a rows x cols column-major matrix of
 *                           elem_size-byte elements between host (ldh) and
 *                           device (ldd), ordered on `stream` (NULL = the
 *                           default stream); the destination is complete
 *                           after stream_sync(stream).  MBH_E_ARG for
 *                           rows/cols < 0, elem_size outside [1, 16], an ld
 *                           below max(1, rows) or a NULL buffer.
 *  stream_create / _destroy / _sync: a non-blocking HIP stream at the
 *                           default priority (each stream used with the
 *                           _dev entries keeps its own scratch workspace). */
int mbh_dev_alloc(size_t bytes, void** ptr);
int mbh_dev_free(void* ptr);
int mbh_host_alloc(size_t bytes, void** ptr);
int mbh_host_free(void* ptr);
int mbh_set_matrix(int64_t rows, int64_t cols, int64_t elem_size, const void* host, int64_t ldh, void* dev,
                   int64_t ldd, void* stream);
int mbh_get_matrix(int64_t rows, int64_t cols, int64_t elem_size, const void* dev, int64_t ldd, void* host,
                   int64_t ldh, void* stream);
int mbh_stream_create(void** stream);
int mbh_stream_destroy(void* stream);  /* also releases the stream's workspace */
int mbh_stream_sync(void* stream);
/* Workspace of the _dev entries on one stream (no counterpart in the
 * reference): release synchronises the stream and frees every workspace
 * buffer it holds (graphs captured on it must not be replayed afterwards);
 * bytes reports how much it holds.  NULL = the default stream. */
int mbh_workspace_release(void* stream);
int mbh_workspace_bytes(void* stream, size_t* bytes);

#ifdef __cplusplus
}
#endif
#endif /* MOJOBLAS_HIP_H */
