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

/* ------------------------------------------------------------------------------------
 * Single precision: llt_inplace<float,Lower>::blocked / unblocked (LLT.h:312-368), the same
 * algorithm in FP32 arithmetic (LLT<MatrixXf>, the LAPACKE_spotrf plug of LLT_LAPACKE.h:90). */
idx_t oracle_llt_unblocked_f32(float *a, idx_t n, idx_t ld)
{
    for (idx_t k = 0; k < n; ++k) {
        idx_t rs = n - k - 1;
        float x = A_(k, k);
        if (k > 0) {
            float s = 0.0f;
            for (idx_t j = 0; j < k; ++j) s += A_(k, j) * A_(k, j);
            x -= s;
        }
        if (x <= 0.0f) return k;
        x = sqrtf(x);
        A_(k, k) = x;
        if (k > 0 && rs > 0) {
            for (idx_t j = 0; j < k; ++j) {
                float akj = A_(k, j);
                for (idx_t i = k + 1; i < n; ++i) A_(i, k) -= A_(i, j) * akj;
            }
        }
        if (rs > 0)
            for (idx_t i = k + 1; i < n; ++i) A_(i, k) /= x;
    }
    return -1;
}

idx_t oracle_llt_blocked_f32(float *a, idx_t n, idx_t ld)
{
    if (n < 32) return oracle_llt_unblocked_f32(a, n, ld);
    idx_t block = oracle_llt_block_size(n);
    for (idx_t k = 0; k < n; k += block) {
        idx_t bs = (block < n - k) ? block : n - k;
        idx_t rs = n - k - bs;
        float *a11 = &A_(k, k), *a21 = &A_(k + bs, k), *a22 = &A_(k + bs, k + bs);
        idx_t ret = oracle_llt_unblocked_f32(a11, bs, ld);
        if (ret >= 0) return k + ret;
        for (idx_t j = 0; j < bs && rs > 0; ++j) { /* A21 <- A21 L11^{-T} (LLT.h:364) */
            for (idx_t q = 0; q < j; ++q) {
                float ljq = a11[j + q * ld];
                for (idx_t i = 0; i < rs; ++i) a21[i + j * ld] -= a21[i + q * ld] * ljq;
            }
            for (idx_t i = 0; i < rs; ++i) a21[i + j * ld] /= a11[j + j * ld];
        }
        for (idx_t j = 0; j < rs; ++j) /* A22 -= A21 A21^T, lower (LLT.h:365) */
            for (idx_t q = 0; q < bs; ++q) {
                float pjq = a21[j + q * ld];
                for (idx_t i = j; i < rs; ++i) a22[i + j * ld] -= a21[i + q * ld] * pjq;
            }
    }
    return -1;
}

/* ------------------------------------------------------------------------------------
 * Eigen/src/Cholesky/LLT.h:239-307 -- llt_rank_update_lower: the factor of L L^T + sigma v v^T
 * in place of the lower factor l.  sigma > 0: Givens rotations, makeGivens(L_ii, -temp_i,
 * &L_ii) (Eigen/src/Jacobi/Jacobi.h:231-266) applied to (L(i+1:, i), temp(i+1:)) as
 * x' = c x + s y, y' = -s x + c y (Jacobi.h:331-332); sigma <= 0: the update of LLT.h:271-296.
 * `work`: n doubles.  Returns -1, or the column j at which x <= 0 (sigma <= 0). */
idx_t oracle_llt_rank_update(double *a, idx_t n, idx_t ld, const double *v, double sigma, double *work)
{
    double *temp = work;
    if (sigma > 0.0) {
        double sq = sqrt(sigma);
        for (idx_t i = 0; i < n; ++i) temp[i] = sq * v[i];
        for (idx_t i = 0; i < n; ++i) {
            double p = A_(i, i), q = -temp[i], c, s, r;
            if (q == 0.0) {
                c = p < 0.0 ? -1.0 : 1.0;
                s = 0.0;
                r = fabs(p);
            } else if (p == 0.0) {
                c = 0.0;
                s = q < 0.0 ? 1.0 : -1.0;
                r = fabs(q);
            } else if (fabs(p) > fabs(q)) {
                double t = q / p, u = sqrt(1.0 + t * t);
                if (p < 0.0) u = -u;
                c = 1.0 / u;
                s = -t * c;
                r = p * u;
            } else {
                double t = p / q, u = sqrt(1.0 + t * t);
                if (q < 0.0) u = -u;
                s = -1.0 / u;
                c = -t * s;
                r = q * u;
            }
            A_(i, i) = r;
            for (idx_t k = i + 1; k < n; ++k) {
                double xi = A_(k, i), yi = temp[k];
                A_(k, i) = c * xi + s * yi;
                temp[k] = -s * xi + c * yi;
            }
        }
        return -1;
    }
    for (idx_t i = 0; i < n; ++i) temp[i] = v[i];
    double beta = 1.0;
    for (idx_t j = 0; j < n; ++j) {
        double ljj = A_(j, j), dj = ljj * ljj, wj = temp[j], swj2 = sigma * wj * wj;
        double gamma = dj * beta + swj2;
        double x = dj + swj2 / beta;
        if (x <= 0.0) return j;
        double nljj = sqrt(x);
        A_(j, j) = nljj;
        beta += swj2 / dj;
        for (idx_t k = j + 1; k < n; ++k) temp[k] -= (wj / ljj) * A_(k, j);
        if (gamma != 0.0)
            for (idx_t k = j + 1; k < n; ++k)
                A_(k, j) = (nljj / ljj) * A_(k, j) + (nljj * sigma * wj / gamma) * temp[k];
    }
    return -1;
}

/* ------------------------------------------------------------------------------------
 * Eigen/src/Core/ConditionEstimator.h:55-141 (rcond_invmatrix_L1_norm_estimate, real case)
 * and :158-169 (rcond_estimate_helper) for an LLT decomposition: the solves use
 * oracle_llt_solve with the lower factor l.  `work`: 4 n doubles. */
static double norm1(const double *v, idx_t n)
{
    double s = 0.0;
    for (idx_t i = 0; i < n; ++i) s += fabs(v[i]);
    return s;
}

double oracle_llt_rcond(const double *l, idx_t n, idx_t ld, double anorm, double *work)
{
    if (n == 0) return INFINITY;
    if (anorm == 0.0) return 0.0;
    if (n == 1) return 1.0;
    double *v = work, *sgn = work + n, *old_sgn = work + 2 * n;
    for (idx_t i = 0; i < n; ++i) v[i] = 1.0 / (double)n;
    oracle_llt_solve(l, n, ld, v, 1, n);
    double lower = norm1(v, n), old_lower = lower;
    idx_t old_j = -1;
    for (int k = 0; k < 4; ++k) {
        for (idx_t i = 0; i < n; ++i) sgn[i] = v[i] < 0.0 ? -1.0 : 1.0;
        if (k > 0 && memcmp(sgn, old_sgn, (size_t)n * sizeof(double)) == 0) break;
        memcpy(v, sgn, (size_t)n * sizeof(double));
        oracle_llt_solve(l, n, ld, v, 1, n);
        idx_t j = 0;
        for (idx_t i = 1; i < n; ++i)
            if (fabs(v[i]) > fabs(v[j])) j = i;
        if (j == old_j) break;
        for (idx_t i = 0; i < n; ++i) v[i] = 0.0;
        v[j] = 1.0;
        oracle_llt_solve(l, n, ld, v, 1, n);
        lower = norm1(v, n);
        if (lower <= old_lower) break;
        memcpy(old_sgn, sgn, (size_t)n * sizeof(double));
        old_j = j;
        old_lower = lower;
    }
    double alt = 1.0;
    for (idx_t i = 0; i < n; ++i) {
        v[i] = alt * (1.0 + (double)i / (double)(n - 1));
        alt = -alt;
    }
    oracle_llt_solve(l, n, ld, v, 1, n);
    double alternate = (2.0 * norm1(v, n)) / (3.0 * (double)n);
    double inv = lower > alternate ? lower : alternate;
    return inv == 0.0 ? 0.0 : (1.0 / inv) / anorm;
}

/* ------------------------------------------------------------------------------------
 * The level-3 routines of the EIGEN_USE_BLAS route, with the semantics of Eigen's BLAS
 * (blas/level3_impl.h:78-178 trsm, :357-420 syrk): characters as blas/common.h:47-63 reads
 * them ('C' is 'T' for real data).
 *
 * trsm: B <- alpha op(A)^{-1} B (side 'L', A m x m) or alpha B op(A)^{-1} (side 'R', A n x n),
 * by plain substitution on the elements of the `uplo` triangle (diag 'U': ones). */
static double tri_at(const double *a, idx_t lda, int upper, int unit, int trans, idx_t i, idx_t j)
{
    /* element (i, j) of op(A) restricted to its triangle; zero outside */
    idx_t r = trans ? j : i, c = trans ? i : j;
    if (upper ? (r > c) : (r < c)) return 0.0;
    if (r == c && unit) return 1.0;
    return a[r + c * lda];
}

void oracle_trsm(char side, char uplo, char transa, char diag, idx_t m, idx_t n, double alpha, const double *a,
                 idx_t lda, double *b, idx_t ldb)
{
    int left = side == 'L' || side == 'l';
    int upper = uplo == 'U' || uplo == 'u';
    int trans = !(transa == 'N' || transa == 'n');
    int unit = diag == 'U' || diag == 'u';
    /* op(A) is lower triangular when (lower, no transpose) or (upper, transpose) */
    int oplower = upper ? trans : !trans;
    if (left) {
        for (idx_t c = 0; c < n; ++c) {
            double *x = b + c * ldb;
            if (oplower) {
                for (idx_t i = 0; i < m; ++i) {
                    double s = x[i];
                    for (idx_t k = 0; k < i; ++k) s -= tri_at(a, lda, upper, unit, trans, i, k) * x[k];
                    x[i] = s / tri_at(a, lda, upper, unit, trans, i, i);
                }
            } else {
                for (idx_t i = m - 1; i >= 0; --i) {
                    double s = x[i];
                    for (idx_t k = i + 1; k < m; ++k) s -= tri_at(a, lda, upper, unit, trans, i, k) * x[k];
                    x[i] = s / tri_at(a, lda, upper, unit, trans, i, i);
                }
            }
        }
    } else { /* X op(A) = B, row by row: op(A)^T x^T = b^T */
        for (idx_t r = 0; r < m; ++r) {
            if (oplower) { /* op(A)^T upper: backward over columns */
                for (idx_t j = n - 1; j >= 0; --j) {
                    double s = b[r + j * ldb];
                    for (idx_t k = j + 1; k < n; ++k) s -= b[r + k * ldb] * tri_at(a, lda, upper, unit, trans, k, j);
                    b[r + j * ldb] = s / tri_at(a, lda, upper, unit, trans, j, j);
                }
            } else {
                for (idx_t j = 0; j < n; ++j) {
                    double s = b[r + j * ldb];
                    for (idx_t k = 0; k < j; ++k) s -= b[r + k * ldb] * tri_at(a, lda, upper, unit, trans, k, j);
                    b[r + j * ldb] = s / tri_at(a, lda, upper, unit, trans, j, j);
                }
            }
        }
    }
    for (idx_t c = 0; c < n; ++c)
        for (idx_t i = 0; i < m; ++i) b[i + c * ldb] *= alpha;
}

/* syrk: C <- alpha op(A) op(A)^T + beta C on the `uplo` triangle of C (n x n);
 * op(A) = A (n x k) for 'N', A^T (A k x n) otherwise; beta == 0 does not read C. */
void oracle_syrk(char uplo, char trans, idx_t n, idx_t k, double alpha, const double *a, idx_t lda, double beta,
                 double *c, idx_t ldc)
{
    int upper = uplo == 'U' || uplo == 'u';
    int notr = trans == 'N' || trans == 'n';
    for (idx_t j = 0; j < n; ++j) {
        idx_t lo = upper ? 0 : j, hi = upper ? j + 1 : n;
        for (idx_t i = lo; i < hi; ++i) {
            double s = 0.0;
            for (idx_t q = 0; q < k; ++q) {
                double ai = notr ? a[i + q * lda] : a[q + i * lda];
                double aj = notr ? a[j + q * lda] : a[q + j * lda];
                s += ai * aj;
            }
            double *p = c + i + j * ldc;
            *p = alpha * s + (beta == 0.0 ? 0.0 : beta * *p);
        }
    }
}
