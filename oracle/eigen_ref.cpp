// eigen_ref.cpp -- C-ABI driver around the genuine Eigen LLT, compiled from the
// reference's own headers (oracle/Makefile passes -I to the reference root).
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement, to write golden
// fixtures, and as the CPU baseline timed beside the GPU in bench.py.  Nothing in the
// product links or loads it.
//
// What is called, in the reference:
//   Eigen::LLT<MatrixXd>::compute / solve     Eigen/src/Cholesky/LLT.h:430-460, 483-498
//   internal::llt_inplace<double,Lower>::blocked   Eigen/src/Cholesky/LLT.h:338-368
//   LLT::rcond                                 Eigen/src/Cholesky/LLT.h:166-171
#include <Eigen/Cholesky>
#include <chrono>
#include <cstdint>

using Eigen::Index;
using Eigen::MatrixXd;
typedef Eigen::Map<MatrixXd, 0, Eigen::OuterStride<> > MapXd;
typedef Eigen::Map<const MatrixXd, 0, Eigen::OuterStride<> > CMapXd;

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" {

// Version string of the compiled Eigen (Macros.h:18-20).
int eigen_ref_version(void)
{
    return EIGEN_WORLD_VERSION * 10000 + EIGEN_MAJOR_VERSION * 100 + EIGEN_MINOR_VERSION;
}

// In-place blocked factor on the lower triangle; returns Eigen's index (-1 = success,
// else first failing column).  Upper triangle untouched.
int64_t eigen_ref_llt_inplace(int64_t n, double *a, int64_t lda)
{
    MapXd A(a, n, n, Eigen::OuterStride<>(lda));
    return (int64_t)Eigen::internal::llt_inplace<double, Eigen::Lower>::blocked(A);
}

// Same for the Upper variant (LLT.h:377-399).
int64_t eigen_ref_llt_inplace_upper(int64_t n, double *a, int64_t lda)
{
    MapXd A(a, n, n, Eigen::OuterStride<>(lda));
    return (int64_t)Eigen::internal::llt_inplace<double, Eigen::Upper>::blocked(A);
}

// Full LLT<MatrixXd> path: compute(A) then X = solve(B).  Writes matrixLLT() into l
// (the whole n x n storage, so the untouched upper triangle is visible), X into x.
// Returns 0 for Success, 1 for NumericalIssue.  Times are wall-clock seconds.
int eigen_ref_llt_factor_solve(int64_t n, const double *a, int64_t lda, double *l, int64_t ldl,
                               const double *b, int64_t nrhs, int64_t ldb, double *x, int64_t ldx,
                               double *t_factor, double *t_solve, double *l1_norm, double *rcond)
{
    CMapXd A(a, n, n, Eigen::OuterStride<>(lda));
    CMapXd B(b, n, nrhs, Eigen::OuterStride<>(ldb));
    Eigen::LLT<MatrixXd> llt(n);
    double t0 = now_s();
    llt.compute(A);
    double t1 = now_s();
    MatrixXd X = llt.solve(B);
    double t2 = now_s();
    if (t_factor) *t_factor = t1 - t0;
    if (t_solve) *t_solve = t2 - t1;
    if (l) {
        MapXd L(l, n, n, Eigen::OuterStride<>(ldl));
        L = llt.matrixLLT();
    }
    if (x) {
        MapXd Xo(x, n, nrhs, Eigen::OuterStride<>(ldx));
        Xo = X;
    }
    if (rcond) *rcond = (llt.info() == Eigen::Success && n > 0) ? llt.rcond() : 0.0;
    (void)l1_norm;
    return llt.info() == Eigen::Success ? 0 : 1;
}

// Timed factor+solve only (bench CPU leg): copies are outside the timed region.
// Returns elapsed seconds of compute()+solve(), or -1 on NumericalIssue.
double eigen_ref_llt_time(int64_t n, const double *a, int64_t lda, const double *b, int64_t nrhs,
                          int64_t ldb, double *x, int64_t ldx)
{
    CMapXd A(a, n, n, Eigen::OuterStride<>(lda));
    CMapXd B(b, n, nrhs, Eigen::OuterStride<>(ldb));
    Eigen::LLT<MatrixXd> llt(n);
    MatrixXd X(n, nrhs);
    double t0 = now_s();
    llt.compute(A);
    X = llt.solve(B);
    double t1 = now_s();
    if (x) {
        MapXd Xo(x, n, nrhs, Eigen::OuterStride<>(ldx));
        Xo = X;
    }
    return llt.info() == Eigen::Success ? (t1 - t0) : -1.0;
}

}  // extern "C"
