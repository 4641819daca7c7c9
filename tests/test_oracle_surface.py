"""Pinning of the oracle's restatements of the rest of the LLT surface (CPU, test
infrastructure): the rank-one update (LLT.h:239-307), the condition estimate
(ConditionEstimator.h:55-169), the level-3 routines of the BLAS route (blas/level3_impl.h) and
the FP32 factorization -- each against closed forms or textbook identities, and against the
genuine Eigen build (oracle/_ref: Eigen's LLT, and Eigen's own BLAS from blas/double.cpp)."""
import numpy as np
import pytest
import scipy.linalg

import oracle
from mi_llt import workloads

needs_eigen = pytest.mark.skipif(not oracle.have_eigen(), reason="genuine Eigen build (oracle/_ref) not present")
needs_blas = pytest.mark.skipif(not oracle.have_eigen_blas(), reason="Eigen BLAS build (oracle/_ref) not present")


def lower_factor(n, seed):
    a = workloads.gram_spd(n, seed=seed)
    return a, np.tril(oracle.llt_blocked(a)[0])


# ------------------------------------------------------------------------ rank update
@pytest.mark.parametrize("sigma", [1.0, 3.5, -0.02])
def test_rank_update_reconstructs_the_updated_matrix(sigma):
    n = 60
    a, l = lower_factor(n, 4)
    v = 0.3 * workloads.rhs(n, 1, seed=5).reshape(-1)
    m, idx = oracle.llt_rank_update(l, v, sigma)
    assert idx == -1
    lu = np.tril(m)
    target = np.tril(a) + np.tril(a, -1).T + sigma * np.outer(v, v)
    assert np.abs(lu @ lu.T - target).max() <= 1e-13 * np.abs(target).max()
    assert np.all(np.diag(lu) > 0)
    # the factor with a positive diagonal is unique: it is numpy's Cholesky of the target
    np.testing.assert_allclose(lu, np.linalg.cholesky(target), rtol=0, atol=1e-12)
    assert np.array_equal(np.triu(m, 1), np.triu(l, 1))  # other triangle untouched


def test_rank_update_one_by_one_closed_form():
    m, idx = oracle.llt_rank_update(np.array([[3.0]]), np.array([4.0]), 1.0)
    assert idx == -1 and m[0, 0] == 5.0  # sqrt(9 + 16)
    m, idx = oracle.llt_rank_update(np.array([[5.0]]), np.array([3.0]), -1.0)
    assert idx == -1 and m[0, 0] == pytest.approx(4.0, rel=1e-15)  # sqrt(25 - 9)
    m, idx = oracle.llt_rank_update(np.array([[2.0]]), np.array([3.0]), -1.0)
    assert idx == 0 and m[0, 0] == 2.0  # 4 - 9 < 0: fails at column 0, untouched


@needs_eigen
@pytest.mark.parametrize("n,sigma", [(33, 2.0), (150, 0.5), (150, -0.05), (129, -1.0), (200, 0.0)])
def test_rank_update_against_genuine_eigen(n, sigma):
    _, l = lower_factor(n, n)
    v = workloads.rhs(n, 1, seed=n + 1).reshape(-1) * (0.1 if sigma < 0 else 1.0)
    mo, io = oracle.llt_rank_update(l, v, sigma)
    me, ie = oracle.eigen_rank_update(l, v, sigma)
    assert io == ie
    np.testing.assert_allclose(mo, me, rtol=0, atol=1e-12 * max(1.0, np.abs(me).max()))


# ------------------------------------------------------------------------ rcond
def test_rcond_exact_on_diagonal_matrices():
    d = np.array([4.0, 0.25, 9.0, 1.0, 16.0])
    l = np.diag(np.sqrt(d))
    anorm = d.max()
    # Hager's method is exact here: ||A^{-1}||_1 = 1 / min d
    assert oracle.llt_rcond(l, anorm) == pytest.approx(d.min() / d.max(), rel=1e-15)


def test_rcond_special_cases():
    assert oracle.llt_rcond(np.zeros((0, 0)), 1.0) == np.inf
    assert oracle.llt_rcond(np.eye(3), 0.0) == 0.0
    assert oracle.llt_rcond(np.array([[7.0]]), 49.0) == 1.0


@pytest.mark.parametrize("n", [2, 40, 150])
def test_rcond_is_an_upper_bound_of_the_true_reciprocal(n):
    a, l = lower_factor(n, 7)
    a = np.tril(a) + np.tril(a, -1).T
    true = 1.0 / (np.linalg.norm(a, 1) * np.linalg.norm(np.linalg.inv(a), 1))
    est = oracle.llt_rcond(l, oracle.l1_norm_lower(a))
    assert true * (1 - 1e-12) <= est <= 3 * true  # a lower bound of ||A^{-1}||_1, within 3x here


@needs_eigen
@pytest.mark.parametrize("n", [10, 97, 300])
def test_rcond_against_genuine_eigen(n):
    a, l = lower_factor(n, 3 * n)
    est = oracle.llt_rcond(l, oracle.l1_norm_lower(a))
    assert est == pytest.approx(oracle.eigen_rcond(a), rel=1e-12)


# ------------------------------------------------------------------------ trsm / syrk
VARIANTS = [(s, u, t, d) for s in "LR" for u in "LU" for t in "NT" for d in "NU"]


def tri_operand(k, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((k, k))
    a[np.diag_indices(k)] = 4.0 + np.abs(a[np.diag_indices(k)])  # well conditioned either triangle
    return np.asfortranarray(a)


@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS)
def test_trsm_against_scipy(side, uplo, trans, diag):
    m, n = 23, 7
    k = m if side == "L" else n
    a = tri_operand(k, 1)
    b = np.asfortranarray(np.random.default_rng(2).standard_normal((m, n)))
    x = oracle.trsm(side, uplo, trans, diag, -0.5, a, b)
    # op(A) as a dense triangular matrix, then the defining equation
    t = np.tril(a) if uplo == "L" else np.triu(a)
    if diag == "U":
        t[np.diag_indices(k)] = 1.0
    op = t.T if trans == "T" else t
    lhs = op @ x if side == "L" else x @ op
    np.testing.assert_allclose(lhs, -0.5 * b, rtol=0, atol=1e-12)
    if side == "L":
        ref = scipy.linalg.solve_triangular(a, b, lower=uplo == "L", trans=trans, unit_diagonal=diag == "U")
        np.testing.assert_allclose(x, -0.5 * ref, rtol=0, atol=1e-12)


@needs_blas
@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS)
def test_trsm_against_eigen_blas(side, uplo, trans, diag):
    m, n = 150, 40
    k = m if side == "L" else n
    a = tri_operand(k, 3)
    b = np.asfortranarray(np.random.default_rng(4).standard_normal((m, n)))
    for alpha in (1.0, 2.5):
        x = oracle.trsm(side, uplo, trans, diag, alpha, a, b)
        e = oracle.eigen_blas_trsm(side, uplo, trans, diag, alpha, a, b)
        np.testing.assert_allclose(x, e, rtol=0, atol=1e-12 * np.abs(e).max())


@pytest.mark.parametrize("uplo,trans", [(u, t) for u in "LU" for t in "NT"])
def test_syrk_against_numpy(uplo, trans):
    n, k = 19, 11
    rng = np.random.default_rng(5)
    a = rng.standard_normal((n, k) if trans == "N" else (k, n))
    c = rng.standard_normal((n, n))
    out = oracle.syrk(uplo, trans, 0.75, a, -2.0, c)
    op = a if trans == "N" else a.T
    full = 0.75 * op @ op.T - 2.0 * c
    tri = np.tril if uplo == "L" else np.triu
    other = np.triu if uplo == "L" else np.tril
    np.testing.assert_allclose(tri(out), tri(full), rtol=0, atol=1e-13)
    assert np.array_equal(other(out, 1 if uplo == "L" else -1), other(c, 1 if uplo == "L" else -1))
    c_nan = np.full((n, n), np.nan)
    out0 = oracle.syrk(uplo, trans, 1.0, a, 0.0, c_nan)  # beta == 0 never reads C
    np.testing.assert_allclose(tri(out0), tri(op @ op.T), rtol=0, atol=1e-13)


@needs_blas
@pytest.mark.parametrize("uplo,trans", [(u, t) for u in "LU" for t in "NT"])
def test_syrk_against_eigen_blas(uplo, trans):
    n, k = 140, 67
    rng = np.random.default_rng(6)
    a = rng.standard_normal((n, k) if trans == "N" else (k, n))
    c = rng.standard_normal((n, n))
    for alpha, beta in ((1.0, 1.0), (-1.0, 0.5), (2.0, 0.0)):
        np.testing.assert_allclose(oracle.syrk(uplo, trans, alpha, a, beta, c),
                                   oracle.eigen_blas_syrk(uplo, trans, alpha, a, beta, c), rtol=0, atol=1e-12)


# ------------------------------------------------------------------------ FP32 factor
def test_f32_factor_closed_forms():
    m, idx = oracle.llt_blocked_f32(np.array([[4.0, -1.0, 2.0], [-1.0, 6.0, 0.0], [2.0, 0.0, 5.0]]))
    assert idx == -1
    s = np.sqrt(5.75)
    np.testing.assert_allclose(np.tril(m), [[2, 0, 0], [-0.5, s, 0], [1, 0.5 / s, np.sqrt(4 - 0.25 / 5.75)]],
                               rtol=2e-7)
    assert oracle.llt_blocked_f32(workloads.indefinite_at(70, 33, seed=2))[1] == 33


@needs_eigen
@pytest.mark.parametrize("n", [31, 100, 257])
def test_f32_factor_against_genuine_eigen(n):
    a = workloads.gram_spd(n, seed=n + 9)
    mo, io = oracle.llt_blocked_f32(a)
    me, ie = oracle.eigen_llt_inplace_f32(a)
    assert io == ie == -1
    # two FP32 factorizations of a matrix with kappa ~ 40: agree to a few FP32 ulps of kappa
    np.testing.assert_allclose(np.tril(mo), np.tril(me), rtol=0, atol=2e-5)
    assert np.array_equal(np.triu(mo, 1), np.triu(me, 1))


# ------------------------------------------------------------------------ GP gradient
@needs_eigen
def test_eigen_gp_gradient_closed_form():
    n = 40
    k = workloads.gram_spd(n, seed=12)
    y = workloads.rhs(n, 1, seed=13).reshape(-1)
    g = oracle.eigen_gp_grad(k, y)
    ksym = np.tril(k) + np.tril(k, -1).T
    alpha = np.linalg.solve(ksym, y)
    np.testing.assert_allclose(g, 0.5 * (np.outer(alpha, alpha) - np.linalg.inv(ksym)), rtol=0, atol=1e-12)
