/*
 * mi_blas.h -- host-pointer level-3 BLAS entry points with the exact signatures of Eigen's
 * BLAS binding (Eigen/src/misc/blas.h:375 dtrsm_, :400 dsyrk_), so that an unmodified Eigen
 * built with -DEIGEN_USE_BLAS sends the triangular solves of LLT::solve (LLT.h:492-497, via
 * TriangularSolverMatrix_BLAS.h:41-161) and the symmetric rank updates of the blocked factor
 * (LLT.h:365, via GeneralMatrixMatrixTriangular_BLAS.h:47-134) to the MI355X.
 *
 * Semantics and argument checks follow Eigen's own BLAS (blas/level3_impl.h:78-178 trsm,
 * :357-420 syrk): an illegal argument is reported on stderr with its position (1-based, as
 * there) and the call returns 0 without computing; 'C' means 'T' for real data.  Each call
 * stages its operands to device memory on a per-host-thread handle (thread safe), computes
 * with mi_dtrsm / mi_dsyrk (mi_llt.h) and copies the result back.  A device failure is printed
 * to stderr and leaves the output as it was on entry.
 */
#ifndef MI_BLAS_H
#define MI_BLAS_H

#ifdef __cplusplus
extern "C" {
#endif

/* B <- alpha op(A)^{-1} B (side 'L') or alpha B op(A)^{-1} (side 'R'). */
int dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m, const int *n,
           const double *alpha, const double *a, const int *lda, double *b, const int *ldb);
/* C <- alpha op(A) op(A)^T + beta C on the uplo triangle. */
int dsyrk_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha, const double *a,
           const int *lda, const double *beta, double *c, const int *ldc);

#ifdef __cplusplus
}
#endif
#endif /* MI_BLAS_H */
