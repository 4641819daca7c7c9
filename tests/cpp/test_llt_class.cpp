// test_llt_class.cpp -- compiled C++ test of the host class mi::LLT (include/mi_llt.hpp), the
// mirror of Eigen::LLT<MatrixXd / MatrixXf> (Eigen/src/Cholesky/LLT.h:66-233), with the
// library's own mi::HostMatrix (no Eigen needed).  Checks, against plain long-double loops:
// compute / info / solve / solveInPlace / matrixL / matrixU / matrixLLT / rcond / rankUpdate /
// l1Norm, the Upper variant, the float variant, a NumericalIssue, and the misuse errors.
// Prints one line of key=value pairs; tests/test_gpu_cpp_class.py runs it on a GPU box.
#include <cmath>
#include <cstdio>
#include <random>

#include "mi_llt.hpp"

using mi::HostMatrix;

template <typename S>
static HostMatrix<S> gram(int n, unsigned seed)
{
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<double> b((size_t)n * n);
    for (auto &x : b) x = nd(gen);
    HostMatrix<S> a(n, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            long double s = 0.0L;
            for (int k = 0; k < n; ++k) s += (long double)b[i + (size_t)k * n] * b[j + (size_t)k * n];
            a(i, j) = (S)(s / n + (i == j ? 0.1L : 0.0L));
        }
    return a;
}

template <typename S>
static double factor_berr(const HostMatrix<S> &a, const HostMatrix<S> &l)
{
    const int n = (int)a.rows();
    long double num = 0.0L, den = 0.0L;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            long double s = 0.0L;
            for (int k = 0; k < n; ++k) s += (long double)l(i, k) * l(j, k);
            num += (a(i, j) - s) * (a(i, j) - s);
            den += (long double)a(i, j) * a(i, j);
        }
    return (double)std::sqrt(num / den);
}

template <typename S>
static double solve_berr(const HostMatrix<S> &a, const HostMatrix<S> &x, const HostMatrix<S> &b)
{
    const int n = (int)a.rows();
    long double den = 0.0L, worst = 0.0L;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) den += (long double)a(i, j) * a(i, j);
    for (int c = 0; c < x.cols(); ++c) {
        long double r2 = 0.0L, x2 = 0.0L;
        for (int i = 0; i < n; ++i) {
            long double s = 0.0L;
            for (int k = 0; k < n; ++k) s += (long double)a(i, k) * x(k, c);
            r2 += (b(i, c) - s) * (b(i, c) - s);
            x2 += (long double)x(i, c) * x(i, c);
        }
        const long double e = std::sqrt(r2) / (std::sqrt(den) * std::sqrt(x2));
        worst = e > worst ? e : worst;
    }
    return (double)worst;
}

int main()
{
    const int n = 300, r = 5;
    HostMatrix<double> a = gram<double>(n, 1);
    HostMatrix<double> b(n, r);
    std::mt19937_64 gen(2);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int c = 0; c < r; ++c)
        for (int i = 0; i < n; ++i) b(i, c) = nd(gen);
    // junk in the triangle LLT never reads
    HostMatrix<double> aj = a;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) aj(i, j) = 1e30;

    mi::LLT<double> llt(aj);
    HostMatrix<double> x = llt.solve(b);
    HostMatrix<double> l, u, m;
    llt.matrixL(l);
    llt.matrixU(u);
    llt.matrixLLT(m);
    double upper_zero = 0.0, lu_transpose = 0.0, junk_kept = 1.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            if (i < j) upper_zero = std::max(upper_zero, std::fabs(l(i, j)));
            lu_transpose = std::max(lu_transpose, std::fabs(l(i, j) - u(j, i)));
            if (i < j && m(i, j) != 1e30) junk_kept = 0.0;
        }
    HostMatrix<double> x2 = b;
    llt.solveInPlace(x2);
    double inplace_diff = 0.0;
    for (int c = 0; c < r; ++c)
        for (int i = 0; i < n; ++i) inplace_diff = std::max(inplace_diff, std::fabs(x2(i, c) - x(i, c)));
    // l1 norm of the symmetric view (LLT.h:442-453)
    double l1 = 0.0;
    for (int c = 0; c < n; ++c) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::fabs(a(i, c));
        l1 = std::max(l1, s);
    }
    const double rcond = llt.rcond();
    // rank update: factor of A + sigma v v^T
    std::vector<double> v(n);
    for (auto &e : v) e = nd(gen);
    struct Vec {
        const std::vector<double> &v;
        int64_t size() const { return (int64_t)v.size(); }
        const double *data() const { return v.data(); }
    };
    llt.rankUpdate(Vec{v}, 0.5);
    HostMatrix<double> lu2, a2 = a;
    llt.matrixL(lu2);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) a2(i, j) += 0.5 * v[i] * v[j];
    const double update_berr = factor_berr(a2, lu2);

    // Upper storage: the upper factor U with A = U^T U
    HostMatrix<double> al = a;
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) al(i, j) = -7.0;  // junk below
    mi::LLT<double, mi::Upper> lltu(al);
    HostMatrix<double> xu = lltu.solve(b), lup;
    lltu.matrixL(lup);

    // float (LLT<MatrixXf>)
    HostMatrix<float> af = gram<float>(n, 1), bf(n, r);
    for (int c = 0; c < r; ++c)
        for (int i = 0; i < n; ++i) bf(i, c) = (float)b(i, c);
    mi::LLT<float> lltf(af);
    HostMatrix<float> xf = lltf.solve(bf), lf;
    lltf.matrixL(lf);

    // NumericalIssue and misuse
    HostMatrix<double> neg = a;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) neg(i, j) = -a(i, j);
    mi::LLT<double> bad(neg);
    int misuse = 0;
    try {
        mi::LLT<double> none;
        (void)none.info();
    } catch (const std::logic_error &) {
        misuse = 1;
    }

    std::printf("info=%d fberr=%.3e sberr=%.3e upper_zero=%.3e lu_transpose=%.3e junk_kept=%.0f inplace_diff=%.3e "
                "l1_diff=%.3e rcond=%.6e update_info=%d update_berr=%.3e upper_info=%d upper_fberr=%.3e "
                "upper_sberr=%.3e float_info=%d float_fberr=%.3e float_sberr=%.3e neg_info=%d misuse=%d\n",
                (int)llt.info(), factor_berr(a, l), solve_berr(a, x, b), upper_zero, lu_transpose, junk_kept,
                inplace_diff, std::fabs(llt.l1Norm() - l1) / l1, rcond, (int)llt.info(), update_berr,
                (int)lltu.info(), factor_berr(a, lup), solve_berr(a, xu, b), (int)lltf.info(), factor_berr(af, lf),
                solve_berr(af, xf, bf), (int)bad.info(), misuse);
    return 0;
}
