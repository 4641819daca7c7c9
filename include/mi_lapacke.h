/*
 * mi_lapacke.h -- host-pointer drop-ins with the exact LAPACKE / LAPACK signatures, so an
 * unmodified Eigen built with -DEIGEN_USE_LAPACKE routes LLT<MatrixXd>::compute to the
 * MI355X (Eigen/src/Cholesky/LLT_LAPACKE.h:59-61 calls LAPACKE_dpotrf, declared in
 * Eigen/src/misc/lapacke.h:2823), and LAPACK callers of dpotrf_/dpotrs_ (the ABI Eigen
 * itself exports in lapack/cholesky.cpp:14-72) do too.
 *
 * Each call stages the matrix to device memory, runs the kernels and copies back, on a
 * per-host-thread handle (thread safe).  lapack_int is int (Eigen/src/misc/lapacke.h:48-49).
 * matrix_layout is LAPACK_ROW_MAJOR (101) or LAPACK_COL_MAJOR (102), lapacke.h:122-123.
 *
 * Return values follow LAPACKE: 0, -i for an illegal argument i, or the positive info of
 * the factorization.  A runtime failure of the device returns -1010 (LAPACK_WORK_MEMORY_ERROR
 * is -1010, Eigen/src/misc/lapacke.h:125) and prints the cause to stderr.
 */
#ifndef MI_LAPACKE_H
#define MI_LAPACKE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Eigen/src/misc/lapacke.h:2823 -- used by LLT_LAPACKE.h:59-61 */
int LAPACKE_dpotrf(int matrix_layout, char uplo, int n, double *a, int lda);
/* Eigen/src/misc/lapacke.h:2821 -- LLT<MatrixXf> through LLT_LAPACKE.h:90 (mi_spotrf) */
int LAPACKE_spotrf(int matrix_layout, char uplo, int n, float *a, int lda);
/* Eigen/src/misc/lapacke.h:2842 (dpotrs) -- the solve of lapack/cholesky.cpp:41-72 */
int LAPACKE_dpotrs(int matrix_layout, char uplo, int n, int nrhs, const double *a, int lda, double *b, int ldb);
/* Eigen/src/misc/lapacke.h:2839 (spotrs) -- the float solve (mi_spotrs) */
int LAPACKE_spotrs(int matrix_layout, char uplo, int n, int nrhs, const float *a, int lda, float *b, int ldb);

/* Fortran ABI: lapack/cholesky.cpp:14-36 (potrf) and :41-72 (potrs), instantiated for d and s;
 * names via EIGEN_LAPACK_FUNC -> prefix + name + '_' (lapack/lapack_common.h:16-18). */
void dpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info);
void spotrf_(const char *uplo, const int *n, float *a, const int *lda, int *info);
void dpotrs_(const char *uplo, const int *n, const int *nrhs, const double *a, const int *lda, double *b,
             const int *ldb, int *info);
void spotrs_(const char *uplo, const int *n, const int *nrhs, const float *a, const int *lda, float *b,
             const int *ldb, int *info);

#ifdef __cplusplus
}
#endif
#endif /* MI_LAPACKE_H */
