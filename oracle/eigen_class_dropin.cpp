// eigen_class_dropin.cpp -- source-level drop-in check of the C++ host class mi::LLT
// (include/mi_llt.hpp): the same code on Eigen's own MatrixXd with Eigen::LLT (the
// reference, on the CPU) and with mi::LLT (the MI355X), results compared: factor, solve,
// rcond (LLT.h:166-171), rankUpdate (LLT.h:467-480) and a NumericalIssue.
//
// TEST INFRASTRUCTURE: built by oracle/Makefile into oracle/_ref/ when the reference tree is
// present; run by tests/test_gpu_cpp_class.py on a GPU box.  Prints key=value pairs.
#include <Eigen/Cholesky>
#include <cstdio>
#include <random>

#include "mi_llt.hpp"

using Eigen::MatrixXd;
using Eigen::VectorXd;

template <class LLTType>
static void run(const MatrixXd &A, const MatrixXd &B, const VectorXd &v, MatrixXd &L, MatrixXd &X, double &rcond,
                MatrixXd &Lu, int &info, int &neg)
{
    LLTType llt(A);
    info = (int)llt.info();
    X = llt.solve(B);
    L = llt.matrixL();
    rcond = llt.rcond();
    llt.rankUpdate(v, 0.25);
    Lu = llt.matrixL();
    LLTType bad(MatrixXd(-A));
    neg = (int)bad.info();
}

// mi::LLT returns its factors through an output argument; adapt to the expression above
struct MiLLT : mi::LLT<double> {
    explicit MiLLT(const MatrixXd &a) : mi::LLT<double>(a) {}
    MatrixXd matrixL() const
    {
        MatrixXd l;
        mi::LLT<double>::matrixL(l);
        return l;
    }
};

int main()
{
    const int n = 257, r = 3;
    std::mt19937_64 gen(5);
    std::normal_distribution<double> nd(0.0, 1.0);
    MatrixXd Bm(n, n), B(n, r);
    VectorXd v(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) Bm(i, j) = nd(gen);
    for (int j = 0; j < r; ++j)
        for (int i = 0; i < n; ++i) B(i, j) = nd(gen);
    for (int i = 0; i < n; ++i) v(i) = nd(gen);
    MatrixXd A = Bm * Bm.transpose() / n;
    A.diagonal().array() += 0.1;
    MatrixXd Le, Xe, Lue, Lm, Xm, Lum;
    double re = 0, rm = 0;
    int ie = -1, im = -1, ne = -1, nm = -1;
    run<Eigen::LLT<MatrixXd> >(A, B, v, Le, Xe, re, Lue, ie, ne);
    run<MiLLT>(A, B, v, Lm, Xm, rm, Lum, im, nm);
    std::printf("info_eigen=%d info_mi=%d l_diff=%.3e x_diff=%.3e rcond_rel=%.3e update_diff=%.3e neg_eigen=%d "
                "neg_mi=%d\n",
                ie, im, (Le - Lm).cwiseAbs().maxCoeff(), (Xe - Xm).cwiseAbs().maxCoeff() / Xe.cwiseAbs().maxCoeff(),
                std::fabs(re - rm) / re, (Lue - Lum).cwiseAbs().maxCoeff(), ne, nm);
    return 0;
}
