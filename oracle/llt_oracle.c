/*
 * llt_oracle.c -- CPU restatement of Eigen's dense Cholesky (LLT) factor + solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the MI355X
 * kernels: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product path never calls it.
 *
 * Every function restates one piece of Eigen 3.3.90 and cites the file:line it follows
 * (paths relative to the reference root).  Storage is column-major with leading
 * dimension `ld`, exactly as Eigen's default MatrixXd.  Index arithmetic is 64-bit.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   oracle_llt_unblocked / oracle_llt_blocked
 *       - 3x3 known answer of doc/snippets/LLT_example.cpp:1-11 (closed form);
 *       - n=1 (sqrt), 2x2 closed form, diagonal matrices;
 *       - genuine Eigen LLT compiled from the reference (oracle/_ref), element-wise;
 *       - golden fixtures written by tests/golden/make_golden.py from that build;
 *       - failing-column index on matrices with a planted non-PD leading minor.
 *   oracle_llt_solve        - genuine Eigen solve (oracle/_ref) + residual invariants.
 *   oracle_l1_norm_lower    - brute-force symmetric L1 norm in the test.
 *   oracle_*_berr           - exact rational checks on tiny integer matrices.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef int64_t idx_t;

#define A_(i, j) a[(i) + (j) * ld]

/* Eigen/src/Cholesky/LLT.h:312-336 -- llt_inplace<Scalar,Lower>::unblocked.
 * Left-looking column algorithm.  For column k:
 *   x = a_kk - ||A10||^2 ; if x <= 0 return k ; a_kk = sqrt(x)
 *   A21 -= A20 * A10^T   (GEMV, GeneralMatrixVector.h:107-295) ; A21 /= a_kk
 * Returns -1 on success, else the first failing column.  NaN pivots pass the
 * `x <= 0` test exactly as in Eigen. */
idx_t oracle_llt_unblocked(double *a, idx_t n, idx_t ld)
{
    for (idx_t k = 0; k < n; ++k) {
        idx_t rs = n - k - 1;
        double x = A_(k, k);
        if (k > 0) {
            double s = 0.0;
            for (idx_t j = 0; j < k; ++j) s += A_(k, j) * A_(k, j);
            x -= s;
        }
        if (x <= 0.0) return k;
        x = sqrt(x);
        A_(k, k) = x;
        if (k > 0 && rs > 0) {
            for (idx_t j = 0; j < k; ++j) {
                double akj = A_(k, j);
                for (idx_t i = k + 1; i < n; ++i) A_(i, k) -= A_(i, j) * akj;
            }
        }
        if (rs > 0)
            for (idx_t i = k + 1; i < n; ++i) A_(i, k) /= x;
    }
    return -1;
}

/* Eigen/src/Cholesky/LLT.h:346-348 -- the block size rule of llt_inplace::blocked. */
idx_t oracle_llt_block_size(idx_t n)
{
    idx_t bs = n / 8;
    bs = (bs / 16) * 16;
    if (bs < 8) bs = 8;
    if (bs > 128) bs = 128;
    return bs;
}

/* A21 <- A21 * L11^{-T}: the right-side triangular solve of LLT.h:364, i.e.
 * triangular_solve_matrix<..,OnTheRight,Upper,..> (TriangularSolverMatrix.h:186-329)
 * applied to A11^T.  Each row x of A21 solves L11 * x^T = a^T by forward substitution. */
static void trsm_right_lower_trans(const double *l, idx_t ldl, double *b, idx_t ldb,
                                   idx_t rows, idx_t bs)
{
    for (idx_t j = 0; j < bs; ++j) {
        for (idx_t k = 0; k < j; ++k) {
            double ljk = l[j + k * ldl];
            for (idx_t i = 0; i < rows; ++i) b[i + j * ldb] -= b[i + k * ldb] * ljk;
        }
        double d = l[j + j * ldl];
        for (idx_t i = 0; i < rows; ++i) b[i + j * ldb] /= d;
    }
}

/* A22 (lower triangle only) -= A21 * A21^T: the rank update of LLT.h:365,
 * SelfadjointProduct.h:82-117 -> general_matrix_matrix_triangular_product<..,Lower>
 * (GeneralMatrixMatrixTriangular.h:57-194).  Only the lower triangle is written. */
static void syrk_lower_minus(const double *p, idx_t ldp, idx_t rows, idx_t k,
                             double *c, idx_t ldc)
{
    for (idx_t j = 0; j < rows; ++j)
        for (idx_t kk = 0; kk < k; ++kk) {
            double pjk = p[j + kk * ldp];
            for (idx_t i = j; i < rows; ++i) c[i + j * ldc] -= p[i + kk * ldp] * pjk;
        }
}

/* Eigen/src/Cholesky/LLT.h:338-368 -- llt_inplace<Scalar,Lower>::blocked.
 * n < 32 -> unblocked; otherwise right-looking blocked Cholesky with the block size of
 * oracle_llt_block_size().  Returns -1 on success, else k + (failing column of the
 * diagonal block), i.e. the global index of the first failing column. */
idx_t oracle_llt_blocked(double *a, idx_t n, idx_t ld)
{
    if (n < 32) return oracle_llt_unblocked(a, n, ld);
    idx_t block = oracle_llt_block_size(n);
    for (idx_t k = 0; k < n; k += block) {
        idx_t bs = (block < n - k) ? block : n - k;
        idx_t rs = n - k - bs;
        double *a11 = &A_(k, k);
        double *a21 = &A_(k + bs, k);
        double *a22 = &A_(k + bs, k + bs);
        idx_t ret = oracle_llt_unblocked(a11, bs, ld);
        if (ret >= 0) return k + ret;
        if (rs > 0) trsm_right_lower_trans(a11, ld, a21, ld, rs, bs);
        if (rs > 0) syrk_lower_minus(a21, ld, rs, bs, a22, ld);
    }
    return -1;
}

/* Eigen/src/Cholesky/LLT.h:377-399 -- llt_inplace<Scalar,Upper>: the upper factor U of
 * a column-major matrix is the lower factor of its transpose.  The transposed view is
 * taken by copying the upper triangle into a lower-triangle scratch (`work`, n*n). */
idx_t oracle_llt_blocked_upper(double *a, idx_t n, idx_t ld, double *work)
{
    for (idx_t j = 0; j < n; ++j)
        for (idx_t i = j; i < n; ++i) work[i + j * n] = A_(j, i);
    idx_t ret = oracle_llt_blocked(work, n, n);
    for (idx_t j = 0; j < n; ++j)
        for (idx_t i = j; i < n; ++i) A_(j, i) = work[i + j * n];
    return ret;
}

/* Eigen/src/Cholesky/LLT.h:483-498 (_solve_impl / _solve_impl_transposed<true>) with
 * TriangularSolverMatrix.h:36-184 (left, Lower, ColMajor) and the Upper solve on the
 * adjoint view: dst = rhs; L * Y = dst (forward); L^T * X = Y (backward).  In place on
 * b (n x nrhs, leading dimension ldb).  For nrhs == 1 Eigen dispatches to
 * TriangularSolverVector.h:86-148; the arithmetic result is the same substitution. */
void oracle_llt_solve(const double *l, idx_t n, idx_t ldl, double *b, idx_t nrhs, idx_t ldb)
{
    for (idx_t c = 0; c < nrhs; ++c) {
        double *x = b + c * ldb;
        for (idx_t j = 0; j < n; ++j) {           /* forward: column-oriented */
            x[j] /= l[j + j * ldl];
            double xj = x[j];
            for (idx_t i = j + 1; i < n; ++i) x[i] -= l[i + j * ldl] * xj;
        }
        for (idx_t j = n - 1; j >= 0; --j) {      /* backward with L^T: row j of L^T = column j of L */
            double s = x[j];
            for (idx_t i = j + 1; i < n; ++i) s -= l[i + j * ldl] * x[i];
            x[j] = s / l[j + j * ldl];
        }
    }
}

/* Eigen/src/Cholesky/LLT.h:442-453 -- m_l1_norm of the symmetric matrix whose lower
 * triangle is stored: max over columns of sum|col(col).tail| + sum|row(col).head|. */
double oracle_l1_norm_lower(const double *a, idx_t n, idx_t ld)
{
    double best = 0.0;
    for (idx_t col = 0; col < n; ++col) {
        double s = 0.0;
        for (idx_t i = col; i < n; ++i) s += fabs(A_(i, col));
        for (idx_t j = 0; j < col; ++j) s += fabs(A_(col, j));
        if (s > best) best = s;
    }
    return best;
}

/* ||A - L L^T||_F / ||A||_F over the full symmetric matrix (lower triangle of A
 * mirrored), accumulated in long double so the metric itself does not add FP64
 * rounding at the size it measures. */
double oracle_factor_berr(const double *a, idx_t lda, const double *l, idx_t ldl, idx_t n)
{
    long double num = 0.0L, den = 0.0L;
    for (idx_t j = 0; j < n; ++j)
        for (idx_t i = j; i < n; ++i) {
            long double s = 0.0L;
            for (idx_t k = 0; k <= j; ++k)
                s += (long double)l[i + k * ldl] * (long double)l[j + k * ldl];
            long double aij = a[i + j * lda];
            long double d = aij - s;
            long double w = (i == j) ? 1.0L : 2.0L;
            num += w * d * d;
            den += w * aij * aij;
        }
    return (double)sqrtl(num / den);
}

/* max over columns c of ||A x_c - b_c||_2 / (||A||_F ||x_c||_2), A symmetric with the
 * lower triangle stored.  Long double accumulation. */
double oracle_solve_berr(const double *a, idx_t lda, idx_t n, const double *x, idx_t ldx,
                         const double *b, idx_t ldb, idx_t nrhs)
{
    long double anorm2 = 0.0L;
    for (idx_t j = 0; j < n; ++j)
        for (idx_t i = j; i < n; ++i) {
            long double v = a[i + j * lda];
            anorm2 += (i == j ? 1.0L : 2.0L) * v * v;
        }
    long double anorm = sqrtl(anorm2), worst = 0.0L;
    for (idx_t c = 0; c < nrhs; ++c) {
        const double *xc = x + c * ldx, *bc = b + c * ldb;
        long double r2 = 0.0L, x2 = 0.0L;
        for (idx_t i = 0; i < n; ++i) {
            long double s = 0.0L;
            for (idx_t k = 0; k < n; ++k) {
                double aik = (i >= k) ? a[i + k * lda] : a[k + i * lda];
                s += (long double)aik * (long double)xc[k];
            }
            long double r = (long double)bc[i] - s;
            r2 += r * r;
            x2 += (long double)xc[i] * (long double)xc[i];
        }
        long double e = sqrtl(r2) / (anorm * sqrtl(x2));
        if (e > worst) worst = e;
    }
    return (double)worst;
}
