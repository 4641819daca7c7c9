/*
 * mi_llt.h -- C ABI of the MI355X dense Cholesky (LLT) factor + solve library.
 *
 * Plain pointers and sizes only.  Matrices are column-major with a leading dimension, as
 * Eigen's default MatrixXd.  "d" pointers are device (HBM) pointers, and work is queued on
 * the handle's HIP stream.  Host-pointer drop-ins are in mi_lapacke.h.
 *
 * Reference interfaces these entry points replace (Eigen 3.3.90, paths relative to the
 * reference root):
 *   internal::llt_inplace<double,Lower|Upper>::blocked   Eigen/src/Cholesky/LLT.h:338-368, 377-399
 *   LLT::_solve_impl / solveInPlace                      Eigen/src/Cholesky/LLT.h:483-522
 *   LAPACKE_dpotrf (the EIGEN_USE_LAPACKE plug)          Eigen/src/Cholesky/LLT_LAPACKE.h:42-93
 *   dpotrf_ / dpotrs_ (Eigen's LAPACK exports)           lapack/cholesky.cpp:14-72
 *
 * Status codes: 0 on success; -i when argument i (1-based, counting the handle) is
 * illegal, as LAPACK's xerbla convention; MI_LLT_E_* below for runtime failures.
 * Numerical failure is not a status: it is reported through `info` with the LAPACK
 * convention (0 = success, k+1 = the leading minor of order k+1 is not positive definite,
 * i.e. Eigen's failing column k, LLT.h:327-330 and lapack/cholesky.cpp:32-33).
 *
 * Threading: a handle is not internally synchronised; use one handle per host thread
 * (handles are independent and may share a device).
 */
#ifndef MI_LLT_H
#define MI_LLT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mi_llt_context *mi_llt_handle;

#define MI_LLT_SUCCESS 0
#define MI_LLT_E_HIP (-1001)      /* a HIP runtime call failed (see mi_llt_last_error) */
#define MI_LLT_E_ALLOC (-1002)    /* device workspace allocation failed */
#define MI_LLT_E_HANDLE (-1003)   /* null or destroyed handle */
#define MI_LLT_E_UNSUPPORTED (-1004)
#define MI_LLT_E_DEVICE (-1005)   /* the device reported a failure of the queued work (a status bit of
                                     mi_llt_status: a timeout or a ticket outside its launch); the
                                     results of that work are invalid.  Synchronous entry points read
                                     and clear the status word and return this code. */

/* Trailing-update precision for mi_dposv_mixed. */
typedef enum {
    MI_PREC_FP64 = 0,  /* FP64 MFMA trailing update (plain path, no refinement needed) */
    MI_PREC_FP32 = 1,  /* FP32 factor storage, FP32 MFMA trailing update, FP32 accumulation */
    MI_PREC_BF16 = 2,  /* FP32 storage, BF16 MFMA trailing update, FP32 accumulation */
    MI_PREC_BF16X3 = 3, /* FP32 storage, split-BF16 (3 products) trailing update, FP32 acc. */
    MI_PREC_BF16X6 = 4  /* FP32 storage, 3-part split BF16 (6 products, ~FP32 accuracy) trailing
                           update, FP32 MFMA inside the panel step, FP32 acc. */
} mi_prec;

/* ---- handle ------------------------------------------------------------------------ */
int mi_llt_create(mi_llt_handle *h);
int mi_llt_destroy(mi_llt_handle h);
/* Use `stream` (a hipStream_t) for all later work; NULL selects the null stream.  When the
   stream changes, the call first waits for the work queued on the previous one (unless that
   stream is being captured into a graph), since the handle's workspace and counters serve one
   stream's calls in order. */
int mi_llt_set_stream(mi_llt_handle h, void *stream);
/* Text of the last runtime failure on this handle ("" if none). */
const char *mi_llt_last_error(mi_llt_handle h);
/* Device bytes currently held by the handle: its workspace, the solve's scratch, the batched
   workspace and the synchronisation words (all grow-only, freed by mi_llt_destroy). */
int64_t mi_llt_workspace_bytes(mi_llt_handle h);
/* Waits for the handle's stream, then returns (and clears) the device status bits of the
 * asynchronous calls since the last query: 0, or a combination of MI_LLT_STATUS_SOLVE_TIMEOUT
 * (a solve sweep gave up waiting for a block), MI_LLT_STATUS_PANEL_TIMEOUT (a panel solve gave
 * up waiting for its diagonal block) and MI_LLT_STATUS_FACTOR_TICKET (a factorization workgroup
 * drew a ticket outside its launch and stopped); the affected result is then invalid.  While a
 * bit is set, the factorization and solve kernels queued on the handle exit without work, so
 * an asynchronous caller checks the status before trusting later results.  Synchronous entry
 * points (mi_dpotrf, mi_dposv_mixed, the host shims) read and clear it themselves and return
 * MI_LLT_E_DEVICE when it is set. */
#define MI_LLT_STATUS_SOLVE_TIMEOUT 1
#define MI_LLT_STATUS_PANEL_TIMEOUT 2
#define MI_LLT_STATUS_FACTOR_TICKET 4
int mi_llt_status(mi_llt_handle h, int *status);
/* Sets `bits` in the handle's status word on its stream, as a device failure would (no
 * reference counterpart: for tests of the failure path, which must not make a kernel fail). */
int mi_llt_inject_status(mi_llt_handle h, int bits);

/* ---- measurement (no reference counterpart: Eigen times with bench/BenchTimer.h) -------
 * When tracing is on, each launch of the kinds below is bracketed by HIP events on the
 * handle's stream; mi_llt_trace_read waits for the stream and returns, per kind, the summed
 * event time (ms), the summed algorithmic flops (LAPACK counts) and the launch count since
 * the last reset.  MI_TRACE_SOLVE brackets a whole forward+backward solve. */
#define MI_TRACE_POTRF 0   /* factorization step: panel (update, POTRF, TRSM) + trailing SYRK */
#define MI_TRACE_TRSM 1    /* panel A21 <- A21 L11^{-T} */
#define MI_TRACE_SYRK 2    /* trailing A22 -= A21 A21^T */
#define MI_TRACE_SOLVE 3   /* forward + backward substitution */
#define MI_TRACE_TRTRI 4   /* diagonal-block inverses L_bb^{-T} built for the solve */
#define MI_TRACE_GEMM 5    /* FP64 residual R = B - A X of the mixed-precision refinement */
#define MI_TRACE_BATCH 6   /* batched factor + solve (plan + persistent launch) */
#define MI_TRACE_KINDS 8
int mi_llt_trace(mi_llt_handle h, int enable);
int mi_llt_trace_read(mi_llt_handle h, int kind, double *ms, double *flops, int64_t *launches);
int mi_llt_trace_reset(mi_llt_handle h);

/* ---- factor ---------------------------------------------------------------------------
 * In-place Cholesky of the `uplo` triangle ('L' or 'U') of the n x n column-major matrix
 * at dA.  The other triangle is never read or written (LLT.h:61-62, 225-229).
 * Replaces internal::llt_inplace<double,Lower|Upper>::blocked (LLT.h:338-368, 377-399),
 * which LAPACKE_dpotrf replaces under EIGEN_USE_LAPACKE (LLT_LAPACKE.h:59-61).
 * `info` is a host int; the call waits for the stream. */
int mi_dpotrf(mi_llt_handle h, char uplo, int64_t n, double *dA, int64_t lda, int *info);
/* Same, asynchronous: `d_info` is a device int, written with the LAPACK info. */
int mi_dpotrf_async(mi_llt_handle h, char uplo, int64_t n, double *dA, int64_t lda, int *d_info);

/* ---- solve ----------------------------------------------------------------------------
 * B <- A^{-1} B with A = L L^T (uplo 'L') or U^T U (uplo 'U') as left in dA by mi_dpotrf.
 * Replaces LLT::solveInPlace (LLT.h:510-522): forward solve with L then backward with L^T
 * (triangular_solve_matrix, TriangularSolverMatrix.h:36-184), and dpotrs_
 * (lapack/cholesky.cpp:41-72).  Asynchronous on the handle's stream. */
int mi_dpotrs(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, const double *dA, int64_t lda,
              double *dB, int64_t ldb);

/* Factor + solve in one call (LLT(A).solve(B), LLT.h:104-110 + 483-498); reuses the
 * diagonal-block inverses built during the factor.  Asynchronous; `d_info` device int.
 * On a failed factor B is left unchanged. */
int mi_dposv_async(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, double *dA, int64_t lda,
                   double *dB, int64_t ldb, int *d_info);

/* ---- factor + solve without allocation --------------------------------------------------
 * No reference counterpart: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same factor +
 * solve as mi_dposv_async with uplo 'L' (the same kernels, results identical), on `stream` (a
 * hipStream_t, NULL: the null stream), without a handle: every scratch buffer lives in one
 * fixed device array of MI_NOALLOC_SCRATCH_BYTES per GPU (held from the library's first use on
 * that GPU until the process ends), and the call allocates nothing.  The problem must fit it:
 *   512 + 73728 nb + (nrhs > 0 ? 1048576 nb + 24576 nb ceil(nrhs/16) : 0) bytes, nb = ceil(n/128)
 * (e.g. n = 8192 with up to 112 right-hand sides, n = 8064 with 128; a factor alone, nrhs = 0,
 * up to n = 145536); larger problems return MI_LLT_E_TOO_LARGE and queue nothing.  uplo 'U'
 * returns MI_LLT_E_UNSUPPORTED.  d_info (device int): 0, k+1 for a failing pivot at column k
 * (B then unchanged), or MI_LLT_E_DEVICE when the device recorded a failure (a timeout).  All
 * calls on one GPU share its array: calls that overlap on one GPU (from two streams or two
 * threads) race and give wrong results.  Asynchronous. */
#define MI_NOALLOC_SCRATCH_BYTES (80ll << 20)
#define MI_LLT_E_TOO_LARGE (-1006) /* the problem needs more than the fixed scratch array holds */
int mi_dposv_noalloc(char uplo, int64_t n, int64_t nrhs, double *dA, int64_t lda, double *dB, int64_t ldb,
                     int *d_info, void *stream);

/* ---- single precision (LLT<MatrixXf>) -------------------------------------------------
 * The same factorization and solve for float storage, the plug of LLT_LAPACKE.h:90
 * (EIGEN_LAPACKE_LLT(float, float, s) -> LAPACKE_spotrf) and of spotrf_/spotrs_
 * (lapack/cholesky.cpp:14-72, the s-prefixed instance).  FP32 storage; the trailing updates
 * run on FP32 MFMA with FP32 accumulation, the diagonal-block factor and the panel solve in
 * FP64 rounded to FP32, the solve in FP64 on the FP32 factor with the result rounded to FP32.
 * Same argument checks, info convention and uplo semantics as the FP64 entry points. */
int mi_spotrf(mi_llt_handle h, char uplo, int64_t n, float *dA, int64_t lda, int *info);
int mi_spotrf_async(mi_llt_handle h, char uplo, int64_t n, float *dA, int64_t lda, int *d_info);
int mi_spotrs(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, const float *dA, int64_t lda, float *dB,
              int64_t ldb);

/* m_l1_norm of LLT::compute (LLT.h:442-453): max column sum of |.| of the symmetric matrix
 * whose `uplo` triangle is stored at dA.  Writes one double to the device pointer d_norm. */
int mi_dsy_l1norm(mi_llt_handle h, char uplo, int64_t n, const double *dA, int64_t lda,
                  double *d_norm);

/* ---- rest of the LLT surface ----------------------------------------------------------
 * mi_dpocon: reciprocal condition number estimate of A from its factor (as left by mi_dpotrf
 *   in the `uplo` triangle of dL) and anorm = ||A||_1 (mi_dsy_l1norm; LLT keeps it as
 *   m_l1_norm): LLT::rcond (LLT.h:166-171) = rcond_estimate_helper (ConditionEstimator.h:
 *   158-169) over rcond_invmatrix_L1_norm_estimate (:55-141, Hager's method with Higham's
 *   alternating-sign vector, at most 10 solves with the factor, run on the device).  n == 0
 *   gives +inf, anorm == 0 gives 0, n == 1 gives 1.  Host `rcond`; waits for the stream.
 * mi_dllt_rank_update: the factor of L L^T + sigma v v^T in place of L (`uplo` triangle of
 *   dL, n x n; v: n doubles with stride incv >= 1): LLT::rankUpdate (LLT.h:467-480) ->
 *   llt_rank_update_lower (LLT.h:239-307), Givens rotations for sigma > 0 and the update of
 *   LLT.h:271-296 otherwise.  d_info (device int): 0, or j+1 when the updated matrix is not
 *   positive definite at column j (sigma <= 0; columns before j are updated, the rest of the
 *   factor is left as it was, as the reference leaves it).  Asynchronous. */
int mi_dpocon(mi_llt_handle h, char uplo, int64_t n, const double *dL, int64_t ldl, double anorm, double *rcond);
int mi_dllt_rank_update(mi_llt_handle h, char uplo, int64_t n, double *dL, int64_t ldl, const double *dv,
                        int64_t incv, double sigma, int *d_info);

/* ---- level-3 BLAS on the device (the EIGEN_USE_BLAS route of this path) ------------------
 * mi_dtrsm: B <- alpha op(A)^{-1} B (side 'L', A m x m) or alpha B op(A)^{-1} (side 'R', A
 *   n x n); A triangular in its `uplo` triangle (the other is not read), op(A) = A ('N') or A^T
 *   ('T' or 'C'), diag 'U' unit (the diagonal is not read) or 'N'; B m x n.  Replaces
 *   triangular_solve_matrix as specialised for ?trsm_ (TriangularSolverMatrix_BLAS.h:41-161;
 *   LLT::solve's sweeps and the panel solve of LLT.h:364), semantics of blas/level3_impl.h:78-178.
 * mi_dsyrk: C <- alpha op(A) op(A)^T + beta C on the `uplo` triangle of C (n x n; the other
 *   triangle is not touched), op(A) = A (n x k, trans 'N') or A^T (A k x n, trans 'T' or 'C');
 *   beta == 0 does not read C.  Replaces general_matrix_matrix_rankupdate as specialised for
 *   ?syrk_ (GeneralMatrixMatrixTriangular_BLAS.h:47-134; the trailing update of LLT.h:365),
 *   semantics of blas/level3_impl.h:357-420.
 * Both asynchronous on the handle's stream. */
int mi_dtrsm(mi_llt_handle h, char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha,
             const double *dA, int64_t lda, double *dB, int64_t ldb);
int mi_dsyrk(mi_llt_handle h, char uplo, char trans, int64_t n, int64_t k, double alpha, const double *dA,
             int64_t lda, double beta, double *dC, int64_t ldc);

/* ---- mixed precision + FP64 iterative refinement --------------------------------------
 * No reference counterpart: LLT computes in one precision.  Solves A X = B for symmetric
 * positive definite A (FP64, `uplo` triangle at dA, not modified) and writes X to dX
 * (n x nrhs, ld ldx).  The factorization runs on an FP32 working copy of A: trailing
 * updates in the MFMA arithmetic `prec` (mi_prec) with FP32 accumulation, the diagonal-block
 * factor and the panel solves in FP64.  Refinement in FP64: X = solve(B); then repeatedly
 * R = B - A X, stop when max_j ||R_j||_2 / (||A||_F ||X_j||_2) <= tol, else X += solve(R),
 * at most `maxit` corrections.  When the low-precision factor is not positive definite, or
 * refinement does not reach tol, the plain FP64 factorization solves the system instead
 * (*iters = -1, resp. -2; cf. LAPACK dsposv's negative ITER).  prec = MI_PREC_FP64 factors
 * in FP64 at once (*iters = 0).  *iters: corrections applied; *berr: the final backward
 * error; *info: 0, or k+1 when the FP64 factor fails at column k (A not positive definite,
 * X then unchanged).  Host pointers iters/berr/info; the call waits for the stream. */
int mi_dposv_mixed(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, const double *dA, int64_t lda,
                   const double *dB, int64_t ldb, double *dX, int64_t ldx, int prec, double tol, int maxit,
                   int *iters, double *berr, int *info);

/* ---- batched, uneven sizes ------------------------------------------------------------
 * Counterpart: a loop of LLT<MatrixXd>(A_i).solve(B_i) (LLT.h:104-110, 430-460, 483-498) over
 * independent systems.  Factors each A_i (n_i x n_i, lower triangle, in place) and overwrites
 * B_i (n_i x nrhs_i) with A_i^{-1} B_i, for i < batch, in one persistent launch (plus a
 * one-workgroup plan launch); nothing is allocated when the handle's batched workspace is
 * already large enough, so repeated calls can be captured in a HIP graph.  All arrays are
 * device arrays of `batch` entries.  0 <= n_i <= 256; uplo 'L' only (MI_LLT_E_UNSUPPORTED
 * for 'U').  d_info[i]: 0, k+1 for a failing pivot at column k (B_i unchanged), or -1 n_i out
 * of range, -2 lda_i < max(1, n_i), -3 nrhs_i < 0, -4 ldb_i < max(1, n_i) (system skipped). */
int mi_dposv_vbatched(mi_llt_handle h, char uplo, int64_t batch, const int *d_n, double *const *d_A,
                      const int *d_lda, const int *d_nrhs, double *const *d_B, const int *d_ldb, int *d_info);
/* Same in FP32 (FP32 storage and arithmetic, FP32 MFMA). */
int mi_sposv_vbatched(mi_llt_handle h, char uplo, int64_t batch, const int *d_n, float *const *d_A,
                      const int *d_lda, const int *d_nrhs, float *const *d_B, const int *d_ldb, int *d_info);
/* Schedule of the batched launch (no reference counterpart; set by the autotuner in
 * mi_llt.batched): workgroups of 4 worker waves per CU (1..64), tile rows per register block
 * of the panel update (2 or 4; 0, the default: 2 for FP64 and 4 for FP32, the largest each
 * precision holds in registers without spilling), and longest-first ordering (1) or the given
 * order (0). */
int mi_llt_set_vbatched_config(mi_llt_handle h, int wg_per_cu, int group, int longest_first);
int mi_llt_get_vbatched_config(mi_llt_handle h, int *wg_per_cu, int *group, int *longest_first);

/* ---- reverse mode -----------------------------------------------------------------------
 * No reference counterpart (Eigen does not differentiate LLT).  For A = L L^T, uplo 'L'
 * (MI_LLT_E_UNSUPPORTED for 'U'):
 * mi_dpotrf_adjoint: dG holds Lbar = dloss/dL on entry (its lower triangle is read) and Abar =
 *   dloss/dA on return, as a full symmetric matrix (both triangles written; the symmetric
 *   convention of torch.linalg.cholesky): Abar = sym(L^{-T} M L^{-1}), M the halved symmetric
 *   completion of tril(L^T Lbar).  The diagonal-block inverses of L are recomputed, not
 *   stored by the forward pass.  Device workspace: 3 n^2 doubles plus 0.75 MiB per 128 rows
 *   (the diagonal-block inverses and couplings).  Asynchronous.
 * mi_dpotrs_adjoint: for X = A^{-1} B computed with the factor L: dXbar holds Xbar on entry
 *   and Bbar = A^{-1} Xbar on return; dLbar (lower triangle, ld ldlb) is accumulated:
 *   Lbar -= tril(Bbar (L^T X)^T + X (L^T Bbar)^T).  Asynchronous. */
int mi_dpotrf_adjoint(mi_llt_handle h, char uplo, int64_t n, const double *dL, int64_t ldl, double *dG,
                      int64_t ldg);
int mi_dpotrs_adjoint(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, const double *dL, int64_t ldl,
                      const double *dX, int64_t ldx, double *dXbar, int64_t ldxb, double *dLbar, int64_t ldlb);

/* ---- batched, uniform order -------------------------------------------------------------
 * Counterpart: a loop over independent systems of LLT::compute (LLT.h:430-460), LLT::solve
 * (LLT.h:483-498) and the two reverse-mode entry points above.  `batch` (0..65535) systems of
 * the same order n, system z's matrix at dA + z * strideA (likewise every array, strides in
 * elements; systems must not overlap: stride >= ld * columns), uplo 'L' only
 * (MI_LLT_E_UNSUPPORTED for 'U').  Every kernel launch covers all systems, so a batch of small
 * systems (e.g. 256 of order 512) fills the GPU where a loop of single calls would not.
 * Results are those of the single-system entry points, system by system.  Asynchronous.
 * mi_dpotrf_batched_async: d_info is a device array of `batch` ints (LAPACK info per system).
 * mi_dpotrs_batched: solve with each system's factor as left by mi_dpotrf_batched_async.
 * mi_dpotrf_adjoint_batched / mi_dpotrs_adjoint_batched: mi_dpotrf_adjoint / mi_dpotrs_adjoint
 *   per system.  Device workspace per system: 1.07 MiB per 128 rows (panel images and block
 *   inverses) plus 3 n^2 doubles (factor adjoint) or 2 n nrhs doubles and the solve's packed
 *   copies (solve adjoint). */
int mi_dpotrf_batched_async(mi_llt_handle h, char uplo, int64_t n, int64_t batch, double *dA, int64_t lda,
                            int64_t strideA, int *d_info);
int mi_dpotrs_batched(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, int64_t batch, const double *dL,
                      int64_t ldl, int64_t strideL, double *dB, int64_t ldb, int64_t strideB);
int mi_dpotrf_adjoint_batched(mi_llt_handle h, char uplo, int64_t n, int64_t batch, const double *dL, int64_t ldl,
                              int64_t strideL, double *dG, int64_t ldg, int64_t strideG);
int mi_dpotrs_adjoint_batched(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, int64_t batch, const double *dL,
                              int64_t ldl, int64_t strideL, const double *dX, int64_t ldx, int64_t strideX,
                              double *dXbar, int64_t ldxb, int64_t strideXbar, double *dLbar, int64_t ldlb,
                              int64_t strideLbar);
/* mi_dposv_batched_async: factor and solve per system (LLT(A_z).solve(B_z), LLT.h:104-110 +
 *   483-498), the solve reusing the factor's panel images; B_z unchanged when system z's factor
 *   fails (d_info[z] = k+1).
 * mi_dposv_adjoint_batched: the reverse mode of that call.  With L as left in dA and X the
 *   solution: on entry dXbar holds Xbar and dLbar (lower triangle) any gradient already
 *   flowing into L (zero if none); on return dXbar holds Bbar = A^{-1} Xbar and dLbar holds
 *   Abar (full symmetric): the factor adjoint of Lbar - tril(Bbar (L^T X)^T + X (L^T Bbar)^T),
 *   i.e. mi_dpotrs_adjoint followed by mi_dpotrf_adjoint, sharing one set of diagonal-block
 *   inverses.  Workspace per system as above with max(3 n^2, 2 n nrhs) doubles of scratch. */
int mi_dposv_batched_async(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, int64_t batch, double *dA,
                           int64_t lda, int64_t strideA, double *dB, int64_t ldb, int64_t strideB, int *d_info);
int mi_dposv_adjoint_batched(mi_llt_handle h, char uplo, int64_t n, int64_t nrhs, int64_t batch, const double *dL,
                             int64_t ldl, int64_t strideL, const double *dX, int64_t ldx, int64_t strideX,
                             double *dXbar, int64_t ldxb, int64_t strideXbar, double *dLbar, int64_t ldlb,
                             int64_t strideLbar);

#ifdef __cplusplus
}
#endif
#endif /* MI_LLT_H */
