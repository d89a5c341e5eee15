/*
 * fmm_amd.h -- C-ABI of the MI355X (gfx950) fast-matrix-multiplication library.
 *
 * Two kinds of shared library export these symbols:
 *   libfmm_amd.so      the runtime: plans built from [U,V,W] tables at run time;
 *   lib<job>.so        one per generated job (control.py N a1 L1 ... abc),
 *                      with that job's composed tables compiled in; it also
 *                      exports the reference's per-job entry points below.
 *
 * Conventions (the reference's, common/include/bl_dgemm.h:52-130):
 * column-major storage, C += A*B, 32-bit int dimensions and leading
 * dimensions.  Unlike the reference, C needs no padding: nothing outside the
 * m x n block of C is read or written.
 */
#ifndef FMM_AMD_H
#define FMM_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by the int-valued entry points. */
#define FMM_OK            0
#define FMM_EINVAL       (-1)  /* bad argument (shape, leading dimension, table) */
#define FMM_EHIP         (-2)  /* a HIP runtime call failed                      */
#define FMM_ENOMEM       (-3)  /* device or host allocation failed               */
#define FMM_EUNSUPPORTED (-4)  /* dtype / variant not built                      */
#define FMM_ESPIN        (-5)  /* a product-order wait exceeded its time bound   */
#define FMM_ETOOBIG      (-6)  /* larger than fmm_abc_noalloc's fixed scratch    */

/* Element types (dtype argument).  A and B hold the named type; C is FP64
 * for FMM_F64 and FP32 for every other type.  Operand sums are formed in the
 * input precision's FP32 (FP64 for FMM_F64) and rounded once to the MFMA
 * operand type; products accumulate in FP32 (FP64 for FMM_F64).
 *   FMM_F64     FP64 MFMA (v_mfma_f64_16x16x4f64)          -- the drop-in path
 *   FMM_F32     FP32 MFMA (v_mfma_f32_16x16x4f32)
 *   FMM_BF16    BF16 MFMA (v_mfma_f32_16x16x32_bf16)
 *   FMM_F16     FP16 MFMA (v_mfma_f32_16x16x32_f16)
 *   FMM_BF16X3  FP32 A, B; each operand sum split into BF16 hi + lo and every
 *               product formed as hi*hi + hi*lo + lo*hi on BF16 MFMA
 * The reference has FP64 only (common/include/bl_dgemm.h:52-130); the other
 * types have no counterpart there.  Every pack type takes every dtype; the
 * AB workspace holds M in C's type, the NAIVE one S, T in the input type. */
#define FMM_F64     0
#define FMM_F32     1
#define FMM_BF16    2
#define FMM_F16     3
#define FMM_BF16X3  4

/* Pack types (the reference's abc / ab / naive generators,
 * meta/abc_gen_mix.py, meta/ab_gen_mix.py, meta/naive_gen_mix.py):
 *   ABC    operand sums formed while packing, each product updates its C
 *          blocks directly;
 *   AB     operand sums formed while packing, product r into a workspace
 *          block M_r, then C_j += w_j M_r;
 *   NAIVE  S_r, T_r summed into workspaces, M_r = S_r T_r, C_j += w_j M_r
 *          (here applied from the product's accumulators, in product order,
 *          instead of through a stored M_r).
 * The three apply the same C updates in the same order and agree bit for bit. */
#define FMM_PACK_ABC    0
#define FMM_PACK_AB     1
#define FMM_PACK_NAIVE  2

typedef struct fmm_plan_s *fmm_plan_t;

/* ---- plans ---------------------------------------------------------------
 * A plan is one (possibly multi-level) algorithm: `nlevels` partitions
 * dims[3*l .. 3*l+2] = (mx, kx, nx) of level l (level 0 first), and the
 * composed tables U [Mx*Kx x rank], V [Kx*Nx x rank], W [Mx*Nx x rank],
 * row-major, rows in the reference's composed order (level 0 = most
 * significant digit; meta/common_mix.py:65-92, 284-310, 324-341).
 * Replaces: meta/abc_gen_mix.py:407-431 (read_coeffs + generateCoeffs at
 * generation time), done here at run time. */
int fmm_plan_create(fmm_plan_t *plan, int nlevels, const int *dims, int rank,
                    const double *U, const double *V, const double *W);
int fmm_plan_destroy(fmm_plan_t plan);
int fmm_plan_info(fmm_plan_t plan, int *Mx, int *Kx, int *Nx, int *rank);

/* Device workspace (bytes) fmm_abc needs for this shape: a few ints of
 * product-order counters; 0 when no ordering is needed.  The count is that of
 * FMM_F64 and is enough for every dtype (the others use larger C tiles). */
size_t fmm_workspace_bytes(fmm_plan_t plan, int m, int n, int k);
/* The same for any pack type: AB needs M (R [+ Mx*Nx] leaf blocks ms x ns),
 * NAIVE S (R blocks ms x ks), T (R blocks ks x ns) and product-order
 * counters.  Sized for FMM_F64 elements, which bounds every other dtype. */
size_t fmm_workspace_bytes_pack(fmm_plan_t plan, int pack, int m, int n, int k);

/* ---- device-pointer entries (stream ordered, reentrant) -------------------
 * A, B, C are device pointers; `stream` is a hipStream_t (NULL = default).
 * ws may be NULL, in which case the workspace is taken and returned with
 * stream-ordered allocation on `stream`.  Concurrent calls must not share
 * one ws.
 * fmm_abc replaces: bl_dgemm_strassen_abc (meta/abc_gen_mix.py:448-472) with
 * the peeling of common/dgemm/bl_dgemm_util.c:316-353 in the same launch. */
int fmm_abc(fmm_plan_t plan, int dtype, int m, int n, int k,
            const void *A, int lda, const void *B, int ldb, void *C, int ldc,
            void *ws, size_t ws_bytes, void *stream);

/* Any pack type; fmm_abc, fmm_ab and fmm_naive are this with `pack` fixed.
 * fmm_ab replaces bl_dgemm_strassen_ab (meta/ab_gen_mix.py:341-372),
 * fmm_naive bl_dgemm_strassen_naive (meta/naive_gen_mix.py:244-361).  ws of
 * at least fmm_workspace_bytes_pack bytes, or NULL (allocated on `stream`). */
int fmm_run(fmm_plan_t plan, int pack, int dtype, int m, int n, int k,
            const void *A, int lda, const void *B, int ldb, void *C, int ldc,
            void *ws, size_t ws_bytes, void *stream);
int fmm_ab(fmm_plan_t plan, int dtype, int m, int n, int k,
           const void *A, int lda, const void *B, int ldb, void *C, int ldc,
           void *ws, size_t ws_bytes, void *stream);
int fmm_naive(fmm_plan_t plan, int dtype, int m, int n, int k,
              const void *A, int lda, const void *B, int ldb, void *C, int ldc,
              void *ws, size_t ws_bytes, void *stream);

/* fmm_abc without any device allocation.  No counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments as fmm_abc minus the workspace, same kernels, results equal to
 * fmm_abc's bit for bit.  The workspace (product-order counters, one per C
 * tile of the leaf block) comes from a static device array of the library
 * (1.25 MiB per GPU, held until the process ends): at most 65,272 C tiles of
 * 128 x 128 per leaf block (FP64; leaf blocks up to ~32,000 x 32,000), and
 * -- when the plan has not run on this device through another entry yet --
 * at most 1 MiB of its product tables.  Larger problems return FMM_ETOOBIG.
 * The first call on a device sets the array up (symbol lookup, a fill, a
 * one-off placement probe; it synchronises the device).  All calls on one
 * GPU share the array: calls that overlap on one GPU (other streams or host
 * threads) race and their results are undefined.  Order-wait flags are read
 * by fmm_ws_status(NULL, ...). */
int fmm_abc_noalloc(fmm_plan_t plan, int dtype, int m, int n, int k,
                    const void *A, int lda, const void *B, int ldb, void *C,
                    int ldc, void *stream);

/* Plain GEMM C += A*B on the same kernel (identity plan), any dtype above.
 * Replaces: bl_dgemm (common/dgemm/my_dgemm.c:115-210). */
int fmm_gemm(int dtype, int m, int n, int k, const void *A, int lda,
             const void *B, int ldb, void *C, int ldc, void *stream);

/* Error flags (FMM_ESPIN when nonzero) of the last call that used ws:
 * every workspace (fmm_workspace_bytes / _pack, fmm_refine_workspace_bytes)
 * starts with a status header holding them.  With ws == NULL: the flags of
 * all calls since the last such query that allocated their workspace
 * themselves (ws == NULL), on the current device; they are cleared by the
 * query, together with those of fmm_abc_noalloc calls.  Synchronises the
 * device first.  A flag is set when a
 * product-order wait exceeds its time bound (5 s, or FMM_SPIN_LIMIT_MS
 * milliseconds from the environment); C is then not reliable.  The
 * host-pointer entries return FMM_ESPIN themselves. */
int fmm_ws_status(const void *ws, int *flags);

/* Kernel timing for benchmarks (no counterpart in the reference): while
 * enabled, the product-kernel launch of every call made on this host thread
 * is bracketed by two HIP events on its stream; fmm_timing_last_ms waits for
 * the last pair and returns its elapsed time (the kernel alone, without the
 * workspace fill, pre-sum or M_Add launches of the same call). */
int fmm_timing(int enable);
int fmm_timing_last_ms(float *ms);

/* ---- batches: many independent problems in one persistent launch ----------
 * No counterpart in the reference (one FMM call per entry there).  A batch is
 * built from a problem list (device pointers; each problem C += A*B with its
 * own plan -- an index into plans[], or -1 for the plain GEMM -- and its own
 * C-tile shape: 0 = 256x128, 1 = 128x256 for FMM_F32 / FMM_BF16 / FMM_F16;
 * FMM_F64 and FMM_BF16X3 have one shape and ignore it).  The batch owns its
 * device job table and workspace (grow-only).  fmm_batch_create builds and
 * uploads synchronously (and prepares every kernel variant of the dtype, so
 * no later run needs device set-up); fmm_batch_set rebuilds in place with the upload
 * ordered on `stream` (the batch's previous run must have completed).  Each
 * run is one workspace fill and one persistent kernel over all CUs on
 * `stream`, with no allocation or copy; its launch arguments depend only on
 * the batch's buffers, so a HIP graph captured from fmm_batch_run replays any
 * later contents of the same fmm_batch_variant.  Every problem gets exactly
 * the result of a single fmm_run / fmm_gemm call with the same plan and
 * dtype.  Runs of one batch must not overlap; no problem's C may overlap
 * another's operands. */
typedef struct fmm_batch_s *fmm_batch_t;
typedef struct {
    int m, n, k;
    const void *A; int lda;
    const void *B; int ldb;
    void *C; int ldc;
    int plan;   /* index into plans[] or -1 (plain GEMM) */
    int tile;   /* C-tile shape, 0 or 1 */
} fmm_problem_t;
int fmm_batch_create(fmm_batch_t *batch, int dtype, int nprob,
                     const fmm_problem_t *probs, int nplans,
                     const fmm_plan_t *plans);
int fmm_batch_set(fmm_batch_t batch, int nprob, const fmm_problem_t *probs,
                  int nplans, const fmm_plan_t *plans, void *stream);
int fmm_batch_run(fmm_batch_t batch, void *stream);
/* Kernel variant (0..3) the batch's current contents need. */
int fmm_batch_variant(fmm_batch_t batch);
/* Grow the batch's device buffers to `jobs` jobs and `ws_ints` workspace
 * ints now.  Every reallocation (here or in fmm_batch_set) increments the
 * batch's generation; a graph captured from fmm_batch_run is valid while the
 * generation and the variant are unchanged. */
int fmm_batch_reserve(fmm_batch_t batch, int jobs, int ws_ints);
int fmm_batch_generation(fmm_batch_t batch);
int fmm_batch_info(fmm_batch_t batch, int *jobs, int *units);
int fmm_batch_status(fmm_batch_t batch, int *flags);
/* The flags of the batch's last run that has completed, without
 * synchronising: each run ends by copying them into host-mapped memory, so a
 * caller that has seen the run finish (an event after fmm_batch_run) reads
 * them here.  FMM_ESPIN when nonzero. */
int fmm_batch_last_status(fmm_batch_t batch, int *flags);
int fmm_batch_destroy(fmm_batch_t batch);

/* ---- FP64 product through low-precision products and one correction -------
 * No counterpart in the reference.  C += A*B for FP64 A, B, C (device
 * pointers): with A1 = fl_T(A), B1 = fl_T(B) (T = the input type of
 * main_dtype: FMM_F32, FMM_BF16, FMM_F16 or FMM_BF16X3) and the FP32
 * remainders dA = fl32(A - A1), dB = fl32(B - B1),
 *     C~ = A1*B1              by main_plan (NULL: plain GEMM), FP32 result;
 *     E  = [dA | A1]*[B1; dB] by corr_plan (NULL: plain GEMM) of corr_dtype
 *                             (FMM_F32 or FMM_BF16X3; FMM_BF16 when main_dtype
 *                             is FMM_BF16, with dA, dB rounded to BF16),
 *                             depth 2k, FP32 result;
 *     C += C~ + E             in FP64.
 * E is the residual A*B - C~ up to dA*dB and C~'s own arithmetic error, so
 * the correction removes the input-rounding error of the main product.  ws
 * of fmm_refine_workspace_bytes bytes, or NULL (stream-ordered allocation). */
size_t fmm_refine_workspace_bytes(fmm_plan_t main_plan, int main_dtype,
                                  fmm_plan_t corr_plan, int corr_dtype,
                                  int m, int n, int k);
int fmm_refine(fmm_plan_t main_plan, int main_dtype, fmm_plan_t corr_plan,
               int corr_dtype, int m, int n, int k, const double *A, int lda,
               const double *B, int ldb, double *C, int ldc, void *ws,
               size_t ws_bytes, void *stream);

/* ---- host-pointer drop-ins -------------------------------------------------
 * Copy in, run on the current GPU, copy C back; return when done.
 * fmm_host_abc replaces bl_dgemm_strassen_abc for a run-time plan. */
int fmm_host_abc(fmm_plan_t plan, int m, int n, int k, const double *A, int lda,
                 const double *B, int ldb, double *C, int ldc);
/* ... and of any pack type (bl_dgemm_strassen_{abc,ab,naive}). */
int fmm_host_run(fmm_plan_t plan, int pack, int m, int n, int k,
                 const double *A, int lda, const double *B, int ldb,
                 double *C, int ldc);

/* The reference's plain entries, same signatures and C += A*B semantics:
 * common/include/bl_dgemm.h:60-70 (bl_dgemm), :160-170 (bl_dgemm_ref),
 * :207 (bl_finalize).  bl_dgemm runs on the GPU; bl_dgemm_ref is the
 * textbook triple loop on the CPU (common/dgemm/bl_dgemm_ref.c:15-50). */
void bl_dgemm(int m, int n, int k, double *XA, int lda, double *XB, int ldb,
              double *XC, int ldc);
void bl_dgemm_ref(int m, int n, int k, double *XA, int lda, double *XB,
                  int ldb, double *XC, int ldc);
void bl_finalize(void);

/* ---- per-job libraries only (lib<job>.so from control.py) -----------------
 * A job library exports the entries of its own pack type: the reference's
 * generated entry (meta/abc_gen_mix.py:448, meta/ab_gen_mix.py:341,
 * meta/naive_gen_mix.py:244; declared in common/include/bl_dgemm.h:96-130),
 * host pointers, the job's own algorithm: */
void bl_dgemm_strassen_abc(int m, int n, int k, double *XA, int lda,
                           double *XB, int ldb, double *XC, int ldc);
void bl_dgemm_strassen_ab(int m, int n, int k, double *XA, int lda,
                          double *XB, int ldb, double *XC, int ldc);
void bl_dgemm_strassen_naive(int m, int n, int k, double *XA, int lda,
                             double *XB, int ldb, double *XC, int ldc);
/* and the device-pointer form of the same job (the contract of fmm_run): */
int fmm_job_abc(int m, int n, int k, const double *A, int lda, const double *B,
                int ldb, double *C, int ldc, void *ws, size_t ws_bytes,
                void *stream);
int fmm_job_ab(int m, int n, int k, const double *A, int lda, const double *B,
               int ldb, double *C, int ldc, void *ws, size_t ws_bytes,
               void *stream);
int fmm_job_naive(int m, int n, int k, const double *A, int lda,
                  const double *B, int ldb, double *C, int ldc, void *ws,
                  size_t ws_bytes, void *stream);
/* every job library: the job's plan (for the run-time entries above) and its
 * pack type (FMM_PACK_*) */
fmm_plan_t fmm_job_plan(void);
int fmm_job_pack(void);

#ifdef __cplusplus
}
#endif

#endif /* FMM_AMD_H */
