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
//   llt_inplace::rankUpdate                    Eigen/src/Cholesky/LLT.h:239-307, 401-404
//   llt_inplace<float,Lower>::blocked          Eigen/src/Cholesky/LLT.h:338-368 (float)
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

// Threads Eigen's parallel products may use (Eigen::setNbThreads, Parallelizer.h:50-59);
// returns the number Eigen then reports.
int eigen_ref_set_threads(int n)
{
    Eigen::setNbThreads(n);
    return Eigen::nbThreads();
}

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

// In-place blocked factor of a float matrix (LLT<MatrixXf>'s llt_inplace<float,Lower>).
int64_t eigen_ref_llt_inplace_f32(int64_t n, float *a, int64_t lda)
{
    typedef Eigen::Map<Eigen::MatrixXf, 0, Eigen::OuterStride<> > MapXf;
    MapXf A(a, n, n, Eigen::OuterStride<>(lda));
    return (int64_t)Eigen::internal::llt_inplace<float, Eigen::Lower>::blocked(A);
}

// Rank-one update of a lower factor in place: llt_inplace<double,Lower>::rankUpdate
// (LLT.h:239-307, 401-404).  Returns Eigen's index (-1 = success, else the failing column).
int64_t eigen_ref_llt_rank_update(int64_t n, double *l, int64_t ldl, const double *v, double sigma)
{
    MapXd L(l, n, n, Eigen::OuterStride<>(ldl));
    Eigen::Map<const Eigen::VectorXd> V(v, n);
    return (int64_t)Eigen::internal::llt_inplace<double, Eigen::Lower>::rankUpdate(L, V, sigma);
}

// LLT<MatrixXd>(A).rcond() (LLT.h:166-171); -1 when the factorization fails.
double eigen_ref_llt_rcond(int64_t n, const double *a, int64_t lda)
{
    CMapXd A(a, n, n, Eigen::OuterStride<>(lda));
    Eigen::LLT<MatrixXd> llt(A);
    if (llt.info() != Eigen::Success) return -1.0;
    return llt.rcond();
}

// Gradient of the Gaussian-process log marginal likelihood with respect to K, from Eigen's
// LLT on the CPU: G = 1/2 (alpha alpha^T - K^{-1}), alpha = K^{-1} y (K symmetric positive
// definite, lower triangle read).  Returns 0, or 1 when the factorization fails.
int eigen_ref_gp_grad(int64_t n, const double *k, int64_t ldk, const double *y, double *g, int64_t ldg)
{
    CMapXd K(k, n, n, Eigen::OuterStride<>(ldk));
    Eigen::LLT<MatrixXd> llt(K);
    if (llt.info() != Eigen::Success) return 1;
    Eigen::VectorXd alpha = llt.solve(Eigen::Map<const Eigen::VectorXd>(y, n));
    MatrixXd kinv = llt.solve(MatrixXd::Identity(n, n));
    MapXd G(g, n, n, Eigen::OuterStride<>(ldg));
    G = 0.5 * (alpha * alpha.transpose() - kinv);
    return 0;
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
