/*
 * evd4sc.h - C ABI of the MI355X two-stage symmetric eigensolver.
 *
 * Every entry point takes plain device pointers (column-major, 64-bit leading
 * dimensions), runs stream-ordered on the context's HIP stream, and returns an
 * int status: 0 = success, -k = argument k is invalid (LAPACK style),
 * positive = runtime / numerical failure (see EVD_E_* below).  No entry point
 * exits the process; the message of the last failure is available from
 * evd_ctx_last_error().  Scratch memory comes from the context's arena, which
 * the library zero-initialises where an algorithm needs zeros; callers never
 * rely on hipMalloc returning zeroed memory.
 *
 * Each function cites the reference interface it replaces (file:line in
 * hansheng1001/EVD4SC2025).  The reference passes cuBLAS/cuSOLVER handles; here
 * an evd_ctx carries the stream, the arena and the tuning.
 */
#ifndef EVD4SC_H
#define EVD4SC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct evd_ctx evd_ctx;

/* status codes (positive) */
#define EVD_OK 0
#define EVD_E_HIP 1        /* a HIP runtime call failed */
#define EVD_E_NOMEM 2      /* workspace / HBM budget exceeded */
#define EVD_E_NONFINITE 3  /* input held NaN or Inf */
#define EVD_E_TIMEOUT 4    /* a persistent kernel's bounded wait expired */
#define EVD_E_UNSUPPORTED 5
#define EVD_E_NOCONVERGE 6 /* secular equation did not converge */
#define EVD_E_RANGE 7      /* a mixed-precision operand left the 16-bit range */

/* precision of the GEMM work in SBR and the back transformation */
#define EVD_PREC_FP32 0 /* plain FP32 MFMA / FP32 VALU */
#define EVD_PREC_BF16 1 /* BF16 MFMA, FP32 accumulate, hi/lo split */
#define EVD_PREC_FP16 2 /* FP16 MFMA, FP32 accumulate, hi/lo split */

/* stage indices of evd_ctx_stage_ms() */
#define EVD_STAGE_SBR 0
#define EVD_STAGE_BC 1
#define EVD_STAGE_DC 2
#define EVD_STAGE_BCBACK 3
#define EVD_STAGE_SBRBACK 4
#define EVD_STAGE_COUNT 5
/* extra timers of evd_ctx_stage_ms(): the BC back-transformation kernel alone
 * (the stage time EVD_STAGE_BCBACK also covers packing its reflectors) and
 * the bulge-chase kernel alone (EVD_STAGE_BC also covers the band packing
 * and the extraction of D and E) */
#define EVD_TIMER_BCBACK_KERNEL 5
#define EVD_TIMER_CHASE_KERNEL 6
#define EVD_TIMER_COUNT 7

/* ---- context ---------------------------------------------------------- */

/* Create a context on `device`.  `stream` is a hipStream_t (NULL = the
 * context creates its own non-blocking stream).  Replaces the handle creation
 * of src/EVD/myEVD.cu:272-276. */
int evd_ctx_create(int device, void *stream, evd_ctx **out);
int evd_ctx_destroy(evd_ctx *ctx);
int evd_ctx_set_stream(evd_ctx *ctx, void *stream);
void *evd_ctx_get_stream(evd_ctx *ctx);
const char *evd_ctx_last_error(const evd_ctx *ctx);
/* Grow the arena to at least `bytes` now (outside any timed region). */
int evd_ctx_reserve(evd_ctx *ctx, size_t bytes);
size_t evd_ctx_arena_bytes(const evd_ctx *ctx);
/* Refuse any arena growth beyond `bytes` (0 = unlimited). */
int evd_ctx_set_budget(evd_ctx *ctx, size_t bytes);
/* Tuning: panel/bandwidth are arguments of the calls; these are the knobs of
 * the kernels (0 = library default).  bc_sweeps = sweeps per BC-back block,
 * bc_waves = pipelined waves per BC-back workgroup, chase_groups = persistent
 * workgroups of the bulge chase, sbr_back_block = WY block width of the SBR
 * back transformation. */
int evd_ctx_set_tuning(evd_ctx *ctx, int bc_sweeps, int bc_waves, int chase_groups, int sbr_back_block);
/* Set one tuning knob by name (0 restores the default): "bc_sweeps",
 * "bc_waves", "chase_groups", "sbr_back_block", "bc_valu" (1: run the BC
 * back transformation on the FP32 VALU kernel instead of the WY blocks on the
 * matrix cores), "skinny_split" (1..4: split-K workgroups per CU of the
 * long-K skinny SBR products, default 4), "chase_depth" (1..6 sweeps per
 * bulge-chase workgroup, at most 3 and what its LDS window holds are used;
 * default 1), "panel_rows" (256 or 512 rows per panel-QR workgroup, default
 * 512), "gemm_tile" (1: 128 x 128, 2: 64 x 64 output tiles of this
 * library's FP32 products; default by shape), "own_gemm" (1: the large
 * plain FP32 products run on this library's MFMA kernels; by default they go
 * to rocBLAS on the context stream, deterministic mode) and "panel_exchange"
 * (1: every panel on the column-exchange kernel; by default panels of up to
 * 32 columns take the blocked Cholesky QR path, and FP32 panels of 33..64
 * columns take it as two 32-column halves, with the exchange kernel as its
 * on-device fallback).  Returns -2 for an unknown name, -3 for a value
 * out of range.
 * No counterpart in the reference (its tile shapes are compile-time). */
int evd_ctx_set_knob(evd_ctx *ctx, const char *name, int value);
int evd_ctx_set_precision(evd_ctx *ctx, int precision);
/* Per-stage device time (ms) of the last evd_syevd_f32 call, measured with
 * hipEvents on the context stream; enable with evd_ctx_enable_timing.  count
 * may be up to EVD_TIMER_COUNT (ms[5] = the BC back-transformation kernel,
 * ms[6] = the bulge-chase kernel). */
int evd_ctx_enable_timing(evd_ctx *ctx, int on);
int evd_ctx_stage_ms(const evd_ctx *ctx, float *ms, int count);
int evd_ctx_synchronize(evd_ctx *ctx);
/* Fill the whole workspace arena with a byte pattern (0xFF: NaN in FP32 and
 * FP64), stream-ordered.  A check that no result depends on stale scratch;
 * no counterpart in the reference. */
int evd_ctx_poison_arena(evd_ctx *ctx, unsigned byte);

/* ---- building blocks with reference counterparts ------------------------ */

/* Tall-skinny Householder panel QR of the m x nb panel A (nb <= 64):
 * A = (I - W Y^T) [R; 0], Y unit lower (LAPACK V), W = V T (dlarft forward
 * columnwise), R upper.  On exit A holds Y (as the reference's panelQR does),
 * W and R are written.  Reflector generation runs in FP64.
 * Replaces panelQR<T> (api/panelQR.cu:20-246, decl api/inc/PanelQR.h:7-19)
 * and hou_tsqr_panel (api/inc/TallShinnyQR.h:9-134). */
int evd_panel_qr_f32(evd_ctx *ctx, long m, long nb, float *A, long lda, float *W, long ldw, float *R, long ldr);

/* Explicit thin Q (m x n) of a panel factorisation: Q = [I_n; 0] - W Y_1^T,
 * where Q holds Y (unit lower, as left by evd_panel_qr_f32) on entry and Y_1
 * is its leading n x n block; work holds n*n floats.  The explicit-Q output of
 * hou_tsqr_panel (api/inc/TallShinnyQR.h:9-134). */
int evd_orgqr_f32(evd_ctx *ctx, long m, long n, float *Q, long ldq, const float *W, long ldw, float *work);

/* Symmetric rank-2k update of the lower triangle, mirrored to the upper:
 * C = alpha (A B^T + B A^T) + beta C, A and B n x k.
 * Replaces tc_ozimmu_syr2k<T> (api/TCblas/tc_ozimmu_syr2k.cu:209-341,
 * decl api/inc/tc_ozimmu_syr2k.h:6-18). */
int evd_syr2k_f32(evd_ctx *ctx, long n, long k, float alpha, const float *A, long lda, const float *B, long ldb,
                  float beta, float *C, long ldc);

/* Full -> band reduction (SBR) with ZY two-sided updates.  A (n x n,
 * symmetric, full storage) is overwritten: its lower band of bandwidth b is
 * B = Q^T A Q, where Q = prod_p (I - W_p Y_p^T); W and Y (n x n) receive all
 * panels (panel p at row (p+1)b, column pb; zeros elsewhere).
 * Replaces my_ZY_ZY_SBR_Vector<T> (api/zy_zy_sy2sb.cu:24-486,
 * decl api/inc/zy_zy_sy2sb.h:5-25); the reference's dOriA/dZ/dR/work scratch
 * comes from the arena. */
int evd_sy2sb_f32(evd_ctx *ctx, long n, long b, long nb, float *A, long lda, float *W, long ldw, float *Y, long ldy);

/* Dense lower band (bandwidth b) of A -> FP64 band storage
 * band[(r-c) + c*ldband], 0 <= r-c < 2b, rows b+1..2b-1 zeroed.
 * Replaces kernel_bugle_chasing_cpydA2dSubA (src/EVD/myEVD.cu:53-78). */
int evd_band_pack_f32(evd_ctx *ctx, long n, long b, const float *A, long lda, double *band, long ldband);

/* Band -> tridiagonal bulge chasing.  band as produced by evd_band_pack_f32
 * is destroyed; D (n) and E (n-1) receive T's diagonal and subdiagonal.  The
 * reflectors H = I - u u^T (||u||^2 = 2) are stored in U ((n+b) x n, ldu >=
 * n+b): sweep s, step j at U[s+1+j*b .. , s] ("noBandU" layout).  All of U is
 * written (zeros outside the reflectors).
 * Replaces chasing_kernel_one_timeV10<32> and its cooperative launch
 * (util/inc/bugle_chasingV10.h:21-392, src/EVD/myEVD.cu:365-451) plus
 * kernel_bugle_chasing_cpydSubA2dA / cpyATr2Vector (src/EVD/myEVD.cu:105-123,
 * util/kernelOther/kernelOther.cu:399-440). */
int evd_sb2st_f64(evd_ctx *ctx, long n, long b, double *band, long ldband, float *U, long ldu, double *D, double *E);

/* evd_sb2st_f64 without any device allocation, for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results (bit for bit) as evd_sb2st_f64; its scratch (one progress word per
 * sweep and the sweep ticket) is a fixed 4 MiB array per device instead of
 * the context arena, so n is limited to 32768 (EVD_E_UNSUPPORTED above).
 * Calls on one device share that array: calls that overlap on one device
 * (any contexts or streams) race and must be serialised by the caller. */
int evd_sb2st_f64_noalloc(evd_ctx *ctx, long n, long b, double *band, long ldband, float *U, long ldu, double *D,
                          double *E);

/* Symmetric tridiagonal divide and conquer on the GPU in FP64.  On exit D
 * holds the eigenvalues in ascending order and Z (n x n) the orthonormal
 * eigenvectors (compz = 'I').  E is destroyed.
 * Replaces LAPACKE_dstedc(LAPACK_COL_MAJOR,'I',...) (src/EVD/myEVD.cu:529). */
int evd_stedc_f64(evd_ctx *ctx, long n, double *D, double *E, double *Z, long ldz);

/* Merge the N/slide slide-wide WY blocks of W, Y (M x N) left to right into
 * one W with prod_k (I - W_k Y_k^T) = I - W Y^T.
 * Replaces computerWYFromSlide (api/computerWYFromSlide.cu:5-25). */
int evd_wy_from_slide_f32(evd_ctx *ctx, long M, long N, long slide, float *W, long ldw, const float *Y, long ldy);

/* Q = I - W Y^T (M x M).  Replaces computerQFromWY (api/computerQFromWY.cu:5-18). */
int evd_q_from_wy_f32(evd_ctx *ctx, long M, long N, float *Q, long ldq, const float *W, long ldw, const float *Y,
                      long ldy);

/* Apply the bulge-chasing reflectors in U to the ncols columns of Q.
 * order = 0: Q <- H_K ... H_1 Q (first-generated reflector first), which is
 *            what my_BC_back_trans_v8_10_noBandU computes
 *            (api/BC_backTransV8_10_noBandU.cu:234-313, decl
 *            api/inc/BC_backTrans.h:3);
 * order = 1: Q <- H_1 ... H_K Q = Q_BC Q (eigenvectors of the band from those
 *            of T).  b must be 32 or 64. */
int evd_bc_back_f32(evd_ctx *ctx, long n, long b, long ncols, float *Q, long ldq, const float *U, long ldu, int order);

/* Apply Q_SBR = prod_p (I - W_p Y_p^T) from the left to the ncols columns of
 * Z: Z <- Q_SBR Z.  W is overwritten by the merged blocks (as the reference's
 * doubling merge does).  Replaces the SBR part of my_SBRBack_BCBack
 * (src/EVD/myEVD.cu:135-245), applied to Z directly instead of forming Q. */
int evd_sbr_back_f32(evd_ctx *ctx, long n, long b, long ncols, float *Z, long ldz, float *W, long ldw, const float *Y,
                     long ldy);

/* Whole pipeline: eigenvalues w (FP64, ascending, n) and, if jobz == 'V',
 * eigenvectors Z (n x n FP32) of the symmetric A (lower triangle read; A is
 * destroyed).  b = bandwidth (32 or 64), nb = SBR block.  Replaces main's
 * SBR -> BC -> D&C -> back-transformation sequence (src/EVD/myEVD.cu:251-571). */
int evd_syevd_f32(evd_ctx *ctx, char jobz, long n, long b, long nb, float *A, long lda, double *w, float *Z, long ldz);
/* (nb = 0: from the tuning cache, else 1024 rounded down to a multiple of b) */
/* Arena bytes evd_syevd_f32 needs for (n, b, nb, jobz) in plain FP32. */
size_t evd_syevd_workspace_f32(long n, long b, long nb, char jobz);
/* The same for a given EVD_PREC_* (the FP16 form's back transformation keeps
 * its WY blocks, 16 KiB per 32 x 32 reflector block, in the arena). */
size_t evd_syevd_workspace_ex(long n, long b, long nb, char jobz, int precision);

/* ---- FP64 forms of the reference entry points ---------------------------
 * The reference's api/ library is instantiated for double only
 * (api/zy_zy_sy2sb.cu:488-507 and the double signatures of
 * api/inc/computerWYFromSlide.h:4-5, api/inc/computerQFromWY.h:4-5,
 * api/inc/BC_backTrans.h:3); these take its double buffers as they are and
 * compute in FP64 (FP64 matrix cores for the GEMMs).  Arguments and results
 * as the _f32 forms above. */
int evd_panel_qr_f64(evd_ctx *ctx, long m, long nb, double *A, long lda, double *W, long ldw, double *R, long ldr);
/* The explicit thin Q of hou_tsqr_panel<double> (api/inc/TallShinnyQR.h:9-134). */
int evd_orgqr_f64(evd_ctx *ctx, long m, long n, double *Q, long ldq, const double *W, long ldw, double *work);
/* Replaces tc_ozimmu_syr2k<double> (api/TCblas/tc_ozimmu_syr2k.cu:209-341). */
int evd_syr2k_f64(evd_ctx *ctx, long n, long k, double alpha, const double *A, long lda, const double *B, long ldb,
                  double beta, double *C, long ldc);
/* Replaces my_ZY_ZY_SBR_Vector<double> (api/zy_zy_sy2sb.cu:24-486). */
int evd_sy2sb_f64(evd_ctx *ctx, long n, long b, long nb, double *A, long lda, double *W, long ldw, double *Y,
                  long ldy);
/* Replaces computerWYFromSlide (api/computerWYFromSlide.cu:5-25). */
int evd_wy_from_slide_f64(evd_ctx *ctx, long M, long N, long slide, double *W, long ldw, const double *Y, long ldy);
/* Replaces computerQFromWY (api/computerQFromWY.cu:5-18). */
int evd_q_from_wy_f64(evd_ctx *ctx, long M, long N, double *Q, long ldq, const double *W, long ldw, const double *Y,
                      long ldy);
/* Replaces my_BC_back_trans_v8_10_noBandU (api/BC_backTransV8_10_noBandU.cu:
 * 234-313; order 0 is its product, order 1 the reverse).  b must be 32, the
 * bandwidth the reference's kernel is written for. */
int evd_bc_back_f64(evd_ctx *ctx, long n, long b, long ncols, double *Q, long ldq, const double *U, long ldu,
                    int order);

/* ---- batched solves (no counterpart in the reference, which solves one
 * matrix per process) ----
 * count same-size problems in strided storage: A + q a_stride (overwritten),
 * w + q w_stride, Z + q z_stride.  The bulge chase of all problems runs as one
 * persistent launch and the BC back transformation as one launch per kernel;
 * SBR, D&C and the SBR back transformation run per problem in order.  Results
 * equal evd_syevd_f32 on each problem bit for bit. */
int evd_syevd_batched_f32(evd_ctx *ctx, char jobz, int count, long n, long b, long nb, float *A, long lda,
                          long a_stride, double *w, long w_stride, float *Z, long ldz, long z_stride);
size_t evd_syevd_batched_workspace_f32(int count, long n, long b, long nb, char jobz, int precision);
/* Bulge chasing of count bands (band + q band_stride, U + q u_stride,
 * D + q d_stride, E + q e_stride) in one persistent launch: the batched form of
 * evd_sb2st_f64. */
int evd_sb2st_batched_f64(evd_ctx *ctx, int count, long n, long b, double *band, long ldband, long band_stride,
                          float *U, long ldu, long u_stride, double *D, long d_stride, double *E, long e_stride);

/* ---- mixed-size batches (no counterpart in the reference, which solves one
 * matrix per process run, src/EVD/myEVD.cu:251-706) ----
 * count problems of any orders: host arrays n[q], A[q] (device, column major,
 * lower triangle read, not modified), lda[q], w[q], Z[q]
 * (NULL allowed for jobz[q] = 'N'), ldz[q], jobz[q].  Problems are embedded in
 * the next bucket order (64, 96, 128, 192, ... up to 4096) with a diagonal pad
 * above ||A||_inf, and each (bucket, jobz) group is one batched solve; larger
 * orders are solved alone.  Results equal an unpadded solve within rounding
 * (not bit for bit).  b = 32 or 64, or 0: each bucket of order N at bandwidth
 * min(32, N / 4).  Synchronises the stream. */
int evd_syevd_mixed_f32(evd_ctx *ctx, int count, const long *n, float *const *A, const long *lda, double *const *w,
                        float *const *Z, const long *ldz, const char *jobz, long b);

/* ---- captured solves (no counterpart in the reference, which launches every
 * kernel from the host on each call) ----
 * evd_graph_create runs one evd_syevd_f32 on the given buffers (so A is
 * consumed and w, Z receive its results), then captures the same solve on
 * those buffers as a hipGraph.  evd_graph_launch replays it on the context's
 * stream (asynchronous; write the next matrix into A first, stream-ordered);
 * evd_graph_wait synchronises the stream and reports the device error word
 * (e.g. EVD_E_NONFINITE for a non-finite input, whose entries were zeroed).
 * While a graph lives, the context's arena cannot grow (a larger solve on the
 * context fails with EVD_E_UNSUPPORTED) and evd_ctx_destroy refuses; other
 * solves on the context may run between replays.  Results of a replay equal
 * evd_syevd_f32 on the same input bit for bit. */
typedef struct evd_graph evd_graph;
int evd_graph_create(evd_ctx *ctx, char jobz, long n, long b, long nb, float *A, long lda, double *w, float *Z,
                     long ldz, evd_graph **out);
/* The same for a same-size batch (evd_syevd_batched_f32 on fixed strided
 * buffers): one graph per size bucket of a batched workload. */
int evd_graph_create_batched(evd_ctx *ctx, char jobz, int count, long n, long b, long nb, float *A, long lda,
                             long a_stride, double *w, long w_stride, float *Z, long ldz, long z_stride,
                             evd_graph **out);
/* The same for a mixed-size workload (evd_syevd_mixed_f32 on fixed buffers):
 * the whole call, every bucket's padding, batched solve and extraction, as one
 * graph.  Replays read the current contents of A[q]; problems of order above
 * 4096 are solved in place (A[q] consumed), so refill those before a replay.
 * Results of a replay equal evd_syevd_mixed_f32 on the same inputs bit for
 * bit.  The graph holds 48 bytes of device descriptors per problem. */
int evd_graph_create_mixed(evd_ctx *ctx, int count, const long *n, float *const *A, const long *lda, double *const *w,
                           float *const *Z, const long *ldz, const char *jobz, long b, evd_graph **out);
int evd_graph_launch(evd_graph *g);
int evd_graph_wait(evd_graph *g);
int evd_graph_destroy(evd_graph *g);

/* ---- autotuning (no counterpart in the reference: its tile shapes are
 * compile-time template arguments and the myEVD nb argument) ----
 * Knob vector (EVD_TUNE_KNOBS entries, 0 = default): {nb, chase_groups,
 * bc_valu, sbr_back_block, chase_depth, panel_rows, gemm_tile, bc_sweeps,
 * bc_waves} - the SBR block and panel width (nb, panel rows per workgroup),
 * the MFMA output tile of the large products, the chase's sweeps per
 * workgroup and workgroup count, and the BC back transformation's kernel and
 * sweep depth (sweeps per block, blocks in flight).
 * evd_autotune solves a synthetic matrix of order n (device buffers of 3 n^2
 * floats allocated for the call) for candidate knob values, one knob at a
 * time, and stores the fastest setting in the process-wide cache under
 * (next power of two >= n, b, the context's precision); knobs_out (nullable)
 * receives it.  evd_syevd_f32 fills every knob the context leaves at 0, and
 * nb = 0, from the cache entry of its (n, b, precision) when there is one.
 * The cache is saved / loaded as text ("n_bucket b precision" and the knob
 * vector per line; lines with only nb .. sbr_back_block load the others as
 * 0); load merges and rejects malformed files as a whole.  evd_tune_cache_get returns 1 when no entry exists. */
#define EVD_TUNE_KNOBS 9
int evd_autotune(evd_ctx *ctx, long n, long b, int reps, int *knobs_out);
int evd_tune_cache_put(long n, long b, int precision, const int *knobs);
int evd_tune_cache_get(long n, long b, int precision, int *knobs);
int evd_tune_cache_clear(void);
int evd_tune_cache_save(const char *path);
int evd_tune_cache_load(const char *path);
/* Whether evd_syevd_f32 on this context consults the cache (default 1). */
int evd_ctx_use_tune_cache(evd_ctx *ctx, int on);

/* ---- metrics (util/tool/checkMetric.cu:5-25, src/EVD/myEVD.cu:598-639) ---- */

/* ||Q^T Q - I||_F / n, computed in FP64 on the GPU. */
int evd_orthogonality_f32(evd_ctx *ctx, long m, long n, const float *Q, long ldq, double *out);
/* The same for an FP64 Q (the reference's checkOrthogonality takes double). */
int evd_orthogonality_f64(evd_ctx *ctx, long m, long n, const double *Q, long ldq, double *out);
/* ||A Z - Z diag(w)||_F / (n ||A||_F), FP64 on the GPU (A full symmetric). */
int evd_backward_error_f32(evd_ctx *ctx, long n, const float *A, long lda, const double *w, const float *Z, long ldz,
                           double *out);

/* ---- dense kernels exposed for tests ------------------------------------ */

/* C = alpha op(A) op(B) + beta C on MFMA (FP32 in/out, FP32 accumulate).
 * ta, tb in {'N','T'}. */
int evd_sgemm(evd_ctx *ctx, char ta, char tb, long m, long n, long k, float alpha, const float *A, long lda,
              const float *B, long ldb, float beta, float *C, long ldc);
/* Same in FP64 on the FP64 MFMA path. */
int evd_dgemm(evd_ctx *ctx, char ta, char tb, long m, long n, long k, double alpha, const double *A, long lda,
              const double *B, long ldb, double beta, double *C, long ldc);

/* ---- synthetic inputs (device-side, seeded) ------------------------------ */

/* A = (X + X^T)/2 with X i.i.d. uniform[-1,1] (kind 0) or N(0,1) (kind 1)
 * from a counter-based Philox4x32-10 stream keyed by seed. */
int evd_gen_symmetric_f32(evd_ctx *ctx, long n, unsigned long long seed, int kind, float *A, long lda);
/* A = Q diag(lam) Q^T with Q the product of n Householder reflectors whose
 * vectors are N(0, I) from the same Philox stream (stream 1); built in FP64
 * on the GPU, rounded to FP32.  lam: device array of n doubles. */
int evd_gen_spectrum_f32(evd_ctx *ctx, long n, unsigned long long seed, const double *lam, float *A, long lda);

/* ---------------------------------------------------------------- serving
 * Request loop for eigenproblems of uneven sizes.  No counterpart in the
 * reference, whose driver solves one matrix per process run
 * (src/EVD/myEVD.cu:251-706).  `slots` contexts, each with its own stream,
 * run admitted requests concurrently.  With hbm_budget > 0 the budget is one
 * device allocation made here; each admitted request (or same-size batch)
 * borrows a contiguous region of it as its workspace and returns it when it
 * completes, so nothing is allocated or freed while requests are in flight
 * (hbm_budget = 0: unlimited, each slot grows its own workspace).  A, w, Z are
 * the caller's device buffers (A column-major n x n FP32, lower triangle read
 * and overwritten; w n FP64; Z n x n FP32 or NULL for eigenvalues only): A must
 * be complete when submitted (the slots run on their own non-blocking streams,
 * unordered with the caller's), and all three must stay valid until the
 * request completes.  Results do not depend on
 * what else runs: a request returns the same bits alone and in a mix. */
typedef struct evd_server evd_server;
int evd_server_create(int device, int slots, size_t hbm_budget, evd_server **out);
int evd_server_destroy(evd_server *srv);
/* Same-size batching of queued requests (default on): a slot that takes a
 * request also takes up to 31 queued requests of the same (n, b, nb, jobz)
 * that fit the HBM budget together and solves them as one
 * evd_syevd_batched_f32 call (bitwise the results of single solves). */
int evd_server_set_batching(evd_server *srv, int on); /* finishes queued requests first */
/* Admission order: 0 (default) = the queued request of the largest order
 * first (FIFO among equal orders), 1 = FIFO.  Largest-first starts the
 * slowest requests as soon as a slot is free, which shortens the latency tail
 * of a mix of orders. */
int evd_server_set_policy(evd_server *srv, int policy);
/* GEMM precision of every slot (EVD_PREC_*, as evd_ctx_set_precision); only
 * while no request is queued or running (else EVD_E_UNSUPPORTED). */
int evd_server_set_precision(evd_server *srv, int precision);
/* Non-blocking; *id receives the request id.  EVD_E_NOMEM if the request's
 * workspace alone exceeds the budget. */
int evd_server_submit(evd_server *srv, long n, long b, long nb, float *A, long lda, double *w, float *Z, long ldz,
                      long *id);
int evd_server_wait(evd_server *srv, long id); /* status of the request */
int evd_server_wait_all(evd_server *srv);      /* first nonzero status, else 0 */
int evd_server_request_ms(evd_server *srv, long id, double *latency_ms);
/* completed requests, solves/s from first submission to last completion,
 * p50 / p99 latency (submission to completion), bytes held by slot arenas */
int evd_server_stats(evd_server *srv, long *completed, double *solves_per_s, double *p50_ms, double *p99_ms,
                     size_t *arena_bytes);

#ifdef __cplusplus
}
#endif

#endif /* EVD4SC_H */
