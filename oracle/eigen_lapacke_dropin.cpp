// eigen_lapacke_dropin.cpp -- the drop-in check: an unmodified Eigen::LLT, built from the
// reference headers with -DEIGEN_USE_LAPACKE and linked against libmi_llt.so, whose
// LAPACKE_dpotrf / LAPACKE_spotrf (include/mi_lapacke.h) it then calls
// (Eigen/src/Cholesky/LLT_LAPACKE.h:59-61, :90).
//
// Built a second time with -DEIGEN_USE_BLAS too (MI_DROPIN_BLAS): LLT::solve's two triangular
// solves then go to the library's dtrsm_ (TriangularSolverMatrix_BLAS.h:41-99), and the
// program also runs the panel solve and the rank update of llt_inplace::blocked (LLT.h:364-365)
// as Eigen writes them, which go to dtrsm_ (side 'R', TriangularSolverMatrix_BLAS.h:102-161)
// and dsyrk_ (GeneralMatrixMatrixTriangular_BLAS.h:47-134).  Every other product in this
// program is a plain loop, so that the BLAS build needs no other BLAS symbol.
//
// TEST INFRASTRUCTURE: built by oracle/Makefile into oracle/_ref/ when the reference tree is
// present; run by tests/test_gpu_dropin.py on a GPU box.  Prints one line of key=value pairs.
// Usage: eigen_lapacke_dropin n nrhs [d|f]
#include <Eigen/Cholesky>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

template <typename M>
static M gram(int n, std::mt19937_64 &gen)
{
    std::normal_distribution<double> nd(0.0, 1.0);
    Eigen::MatrixXd B(n, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) B(i, j) = nd(gen);
    M A(n, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            long double s = 0.0L;
            for (int k = 0; k < n; ++k) s += (long double)B(i, k) * B(j, k);
            A(i, j) = (typename M::Scalar)(s / n + (i == j ? 0.1L : 0.0L));
        }
    return A;
}

// ||A - L L^T||_F / ||A||_F and max_c ||A x_c - b_c|| / (||A||_F ||x_c||), long double loops.
template <typename M>
static void berrs(const M &A, const M &L, const M &X, const M &R, double *fb, double *sb)
{
    const int n = (int)A.rows();
    long double num = 0.0L, den = 0.0L;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            long double s = 0.0L;
            for (int k = 0; k <= (i < j ? i : j); ++k) s += (long double)L(i, k) * L(j, k);
            num += (A(i, j) - s) * (A(i, j) - s);
            den += (long double)A(i, j) * A(i, j);
        }
    *fb = (double)std::sqrt(num / den);
    const long double an = std::sqrt(den);
    long double worst = 0.0L;
    for (int c = 0; c < X.cols(); ++c) {
        long double r2 = 0.0L, x2 = 0.0L;
        for (int i = 0; i < n; ++i) {
            long double s = 0.0L;
            for (int k = 0; k < n; ++k) s += (long double)A(i, k) * X(k, c);
            r2 += (R(i, c) - s) * (R(i, c) - s);
            x2 += (long double)X(i, c) * X(i, c);
        }
        const long double e = std::sqrt(r2) / (an * std::sqrt(x2));
        worst = e > worst ? e : worst;
    }
    *sb = (double)worst;
}

template <typename M>
static int run_llt(int n, int r)
{
    typedef typename M::Scalar S;
    std::mt19937_64 gen(12345);
    std::normal_distribution<double> nd(0.0, 1.0);
    const M A = gram<M>(n, gen);
    M rhs(n, r);
    for (int j = 0; j < r; ++j)
        for (int i = 0; i < n; ++i) rhs(i, j) = (S)nd(gen);

    // GPU through the LAPACKE plug: LLT::compute -> LAPACKE_?potrf -> MI355X
    Eigen::LLT<M> llt(A);
    const M X = llt.solve(rhs);  // with EIGEN_USE_BLAS: dtrsm_ (forward, then backward)
    const M L = llt.matrixL();
    double fberr = 0.0, sberr = 0.0;
    berrs(A, L, X, rhs, &fberr, &sberr);
    // upper triangle of the storage must still hold A (LLT.h:225-229)
    const M &Mx = llt.matrixLLT();
    double upper_diff = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) upper_diff = std::max(upper_diff, (double)std::fabs(Mx(i, j) - A(i, j)));
    // an indefinite matrix: NumericalIssue
    const M Aneg = -A;
    Eigen::LLT<M> bad(Aneg);
    std::printf("n=%d info=%d fberr=%.3e sberr=%.3e upper_diff=%.3e neg_info=%d", n, (int)llt.info(), fberr, sberr,
                upper_diff, (int)bad.info());
    return 0;
}

#ifdef MI_DROPIN_BLAS
// The two level-3 calls of llt_inplace::blocked (LLT.h:364-365) on a block of the given
// sizes, as Eigen writes them, against plain loops; prints the largest relative differences.
static void run_blocked_calls(int bs, int rs)
{
    using Eigen::MatrixXd;
    std::mt19937_64 gen(777);
    std::normal_distribution<double> nd(0.0, 1.0);
    MatrixXd A11 = gram<MatrixXd>(bs, gen);  // factored below by the LAPACKE plug
    Eigen::LLT<MatrixXd> f(A11);
    MatrixXd L11 = f.matrixL();
    MatrixXd A21(rs, bs), A22(rs, rs);
    for (int j = 0; j < bs; ++j)
        for (int i = 0; i < rs; ++i) A21(i, j) = nd(gen);
    for (int j = 0; j < rs; ++j)
        for (int i = 0; i < rs; ++i) A22(i, j) = nd(gen);
    // reference by loops: P = A21 L11^{-T} (each row solves L11 p^T = a^T), C = A22 - P P^T (lower)
    MatrixXd P = A21;
    for (int i = 0; i < rs; ++i)
        for (int j = 0; j < bs; ++j) {
            long double s = P(i, j);
            for (int k = 0; k < j; ++k) s -= (long double)P(i, k) * L11(j, k);
            P(i, j) = (double)(s / L11(j, j));
        }
    MatrixXd C = A22;
    for (int j = 0; j < rs; ++j)
        for (int i = j; i < rs; ++i) {
            long double s = 0.0L;
            for (int k = 0; k < bs; ++k) s += (long double)P(i, k) * P(j, k);
            C(i, j) = (double)(A22(i, j) - s);
        }
    // LLT.h:364: A21 <- A21 L11^{-T} -> dtrsm_('R', ...)
    MatrixXd G21 = A21;
    L11.adjoint().triangularView<Eigen::Upper>().solveInPlace<Eigen::OnTheRight>(G21);
    // LLT.h:365: A22 -= A21 A21^T on the lower triangle -> dsyrk_('L', 'N', ...)
    MatrixXd G22 = A22;
    G22.selfadjointView<Eigen::Lower>().rankUpdate(P, -1);
    double dp = 0.0, dc = 0.0, up = 0.0, pm = 0.0, cm = 0.0;
    for (int j = 0; j < bs; ++j)
        for (int i = 0; i < rs; ++i) {
            dp = std::max(dp, std::fabs(G21(i, j) - P(i, j)));
            pm = std::max(pm, std::fabs(P(i, j)));
        }
    for (int j = 0; j < rs; ++j)
        for (int i = 0; i < rs; ++i) {
            if (i >= j) {
                dc = std::max(dc, std::fabs(G22(i, j) - C(i, j)));
                cm = std::max(cm, std::fabs(C(i, j)));
            } else {
                up = std::max(up, std::fabs(G22(i, j) - A22(i, j)));
            }
        }
    std::printf(" trsm_rel=%.3e syrk_rel=%.3e syrk_upper_diff=%.3e", dp / pm, dc / cm, up);
}
#endif

int main(int argc, char **argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 500;
    const int r = argc > 2 ? std::atoi(argv[2]) : 4;
    const bool single = argc > 3 && std::strcmp(argv[3], "f") == 0;
#ifdef MI_DROPIN_BLAS
    if (single) {
        std::fprintf(stderr, "the BLAS build binds the double routines only\n");
        return 2;
    }
    run_llt<Eigen::MatrixXd>(n, r);
    run_blocked_calls(128, n > 128 ? n - 128 : 64);
#else
    if (single)
        run_llt<Eigen::MatrixXf>(n, r);
    else
        run_llt<Eigen::MatrixXd>(n, r);
#endif
    std::printf("\n");
    return 0;
}
