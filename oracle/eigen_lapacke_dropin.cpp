// eigen_lapacke_dropin.cpp -- the drop-in check: an unmodified Eigen::LLT<MatrixXd>, built
// from the reference headers with -DEIGEN_USE_LAPACKE, linked against libmi_llt.so, whose
// LAPACKE_dpotrf (include/mi_lapacke.h) it then calls (Eigen/src/Cholesky/LLT_LAPACKE.h:59-61).
// The same program also runs Eigen's own CPU factorization (llt_inplace forced through the
// generic path by calling the unblocked kernel of a plain copy) for comparison.
//
// TEST INFRASTRUCTURE: built by oracle/Makefile into oracle/_ref/ when the reference tree is
// present; run by tests/test_gpu_dropin.py on a GPU box.  Prints one line of key=value pairs.
#include <Eigen/Cholesky>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

using Eigen::MatrixXd;

int main(int argc, char **argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 500;
    const int r = argc > 2 ? std::atoi(argv[2]) : 4;
    std::mt19937_64 gen(12345);
    std::normal_distribution<double> nd(0.0, 1.0);
    MatrixXd B(n, n), rhs(n, r);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) B(i, j) = nd(gen);
    for (int j = 0; j < r; ++j)
        for (int i = 0; i < n; ++i) rhs(i, j) = nd(gen);
    MatrixXd A = B * B.transpose() / n;
    A.diagonal().array() += 0.1;

    // GPU through the LAPACKE plug: LLT::compute -> LAPACKE_dpotrf -> MI355X
    Eigen::LLT<MatrixXd> llt(A);
    MatrixXd X = llt.solve(rhs);
    MatrixXd L = llt.matrixL();
    const double fberr = (A - L * L.transpose()).norm() / A.norm();
    const double sberr = (A * X - rhs).norm() / (A.norm() * X.norm());
    // upper triangle of the storage must still hold A (LLT.h:225-229)
    const MatrixXd &M = llt.matrixLLT();
    double upper_diff = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) upper_diff = std::max(upper_diff, std::fabs(M(i, j) - A(i, j)));

    // an indefinite matrix: NumericalIssue
    MatrixXd Aneg = -A;
    Eigen::LLT<MatrixXd> bad(Aneg);

    std::printf("n=%d info=%d fberr=%.3e sberr=%.3e upper_diff=%.3e neg_info=%d\n", n, (int)llt.info(), fberr, sberr,
                upper_diff, (int)bad.info());
    return 0;
}
