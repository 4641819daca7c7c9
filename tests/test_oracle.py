"""Pins the CPU oracle (oracle/llt_oracle.c) before it is trusted as the GPU checker.

Checks, per oracle function:
  * llt_blocked / llt_unblocked: the reference's 3x3 documentation example
    (doc/snippets/LLT_example.cpp:1-2) against its closed form; n=1; the 2x2 closed form;
    diagonal matrices; the golden fixtures written from genuine Eigen (tests/golden);
    the genuine Eigen build itself (oracle/_ref) when present; failing-column indices on
    planted non-PD leading minors; the block-size table of LLT.h:346-348.
  * llt_solve: golden X; A x = b residual; identity and diagonal closed forms.
  * l1_norm_lower: brute-force symmetric column sums.
  * factor_berr / solve_berr: exact values on small integer cases.
Each check would fail on a plausible slip (transposed operand, dropped term, wrong index).
"""
import math
import os

import numpy as np
import pytest

import oracle
from mi_llt import workloads

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "eigen_llt_golden.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def test_block_size_rule():
    # LLT.h:346-348: bs = clamp(16*floor((n/8)/16), 8, 128)
    table = {32: 8, 127: 8, 128: 16, 255: 16, 256: 32, 1023: 112, 1024: 128, 8192: 128}
    for n, bs in table.items():
        assert oracle.block_size(n) == bs, n


def test_kat_3x3_closed_form():
    a = np.array([[4.0, -1.0, 2.0], [-1.0, 6.0, 0.0], [2.0, 0.0, 5.0]])
    l22 = math.sqrt(6.0 - 0.25)
    l32 = (0.0 - 1.0 * -0.5) / l22  # (a32 - l31*l21) / l22 with l31 = 1, l21 = -0.5
    l33 = math.sqrt(5.0 - 1.0 - l32 * l32)
    want = np.array([[2.0, 0, 0], [-0.5, l22, 0], [1.0, l32, l33]])
    for fn in (oracle.llt_blocked, oracle.llt_unblocked):
        m, idx = fn(a)
        assert idx == -1
        np.testing.assert_allclose(np.tril(m), want, rtol=0, atol=1e-15)
        assert np.array_equal(np.triu(m, 1), np.triu(a, 1))


def test_small_closed_forms():
    m, idx = oracle.llt_blocked(np.array([[9.0]]))
    assert idx == -1 and m[0, 0] == 3.0
    a = np.array([[4.0, 2.0], [2.0, 5.0]])
    m, _ = oracle.llt_blocked(a)
    np.testing.assert_allclose(np.tril(m), [[2.0, 0.0], [1.0, 2.0]], atol=0)
    d = np.diag(np.arange(1.0, 41.0) ** 2)
    m, idx = oracle.llt_blocked(d)
    assert idx == -1
    np.testing.assert_array_equal(np.tril(m), np.diag(np.arange(1.0, 41.0)))


def test_empty():
    m, idx = oracle.llt_blocked(np.zeros((0, 0)))
    assert idx == -1 and m.shape == (0, 0)


@pytest.mark.parametrize("tag", ["kat3"] + [f"spd{n}" for n in
                                            (1, 2, 3, 7, 8, 16, 31, 32, 33, 64, 100, 127, 128, 129, 200)])
def test_against_golden_factor_and_solve(golden, tag):
    a = golden[f"{tag}/A"]
    m, idx = oracle.llt_blocked(a)
    assert idx == int(golden[f"{tag}/IDX"]) == -1
    le = golden[f"{tag}/LLT"]
    # same blocked algorithm, different summation order (Eigen's GEBP/FMA): agreement to
    # within a few ulps scaled by the size
    np.testing.assert_allclose(np.tril(m), np.tril(le), rtol=0, atol=1e-13)
    assert np.array_equal(np.triu(le, 1), np.triu(a, 1))  # Eigen leaves the upper triangle
    assert np.array_equal(np.triu(m, 1), np.triu(a, 1))
    b = golden[f"{tag}/B"]
    x = oracle.llt_solve(np.tril(m), b)
    np.testing.assert_allclose(x, golden[f"{tag}/X"], rtol=0, atol=1e-11 * max(1.0, np.abs(x).max()))
    assert oracle.factor_berr(a, np.tril(m)) < 1e-15
    assert oracle.solve_berr(a, x, b) < 1e-15


@pytest.mark.parametrize("n,k", [(1, 0), (5, 3), (33, 0), (33, 32), (100, 17), (129, 128), (200, 150)])
def test_failing_column_matches_golden(golden, n, k):
    tag = f"fail{n}_{k}"
    a = golden[f"{tag}/A"]
    np.testing.assert_array_equal(a, workloads.indefinite_at(n, k, seed=n))
    _, idx = oracle.llt_blocked(a)
    assert idx == int(golden[f"{tag}/IDX"]) == k
    assert int(golden[f"{tag}/INFO"]) == 1


@pytest.mark.parametrize("n", [33, 129])
def test_upper_against_golden(golden, n):
    tag = f"upper{n}"
    a = golden[f"{tag}/A"]
    m, idx = oracle.llt_blocked(a, upper=True)
    assert idx == int(golden[f"{tag}/IDX"]) == -1
    np.testing.assert_allclose(np.triu(m), np.triu(golden[f"{tag}/LLT"]), rtol=0, atol=1e-13)
    assert np.array_equal(np.tril(m, -1), np.tril(a, -1))


def test_inputs_regenerate_golden(golden):
    np.testing.assert_array_equal(golden["spd100/A"], workloads.gram_spd(100, seed=100))
    np.testing.assert_array_equal(golden["spd100/B"], workloads.rhs(100, 3, seed=107))


def test_negative_definite_fails_at_zero():
    a = -workloads.gram_spd(20, seed=3)
    assert oracle.llt_blocked(a)[1] == 0


def test_solve_closed_forms():
    x = oracle.llt_solve(np.eye(6), np.arange(12.0).reshape(6, 2))
    np.testing.assert_array_equal(x, np.arange(12.0).reshape(6, 2))
    l = np.diag([2.0, 4.0])
    x = oracle.llt_solve(l, np.array([[8.0], [32.0]]))  # A = diag(4, 16)
    np.testing.assert_array_equal(x, [[2.0], [2.0]])
    # non-symmetric triangle: a transposed operand would change the answer
    l = np.array([[1.0, 0.0], [3.0, 1.0]])  # A = [[1,3],[3,10]]
    a = l @ l.T
    b = np.array([[1.0], [2.0]])
    np.testing.assert_allclose(a @ oracle.llt_solve(l, b), b, atol=1e-14)


def test_l1_norm_brute_force():
    a = workloads.gram_spd(37, seed=5)
    a_sym = np.tril(a) + np.tril(a, -1).T
    assert oracle.l1_norm_lower(np.tril(a) + 7.0 * np.triu(np.ones_like(a), 1)) == pytest.approx(
        np.abs(a_sym).sum(axis=0).max(), rel=1e-15)


def test_berr_exact_small():
    a = np.array([[4.0, 2.0], [2.0, 5.0]])
    l = np.array([[2.0, 0.0], [1.0, 3.0]])  # L L^T = [[4,2],[2,10]] -> error 5 at (1,1)
    assert oracle.factor_berr(a, l) == pytest.approx(5.0 / math.sqrt(16 + 4 + 4 + 25), rel=1e-15)
    x = np.array([[1.0], [0.0]])
    b = np.array([[4.0], [3.0]])  # A x = [4, 2] -> residual [0, 1]
    assert oracle.solve_berr(a, x, b) == pytest.approx(1.0 / math.sqrt(49.0), rel=1e-15)


@pytest.mark.skipif(not oracle.have_eigen(), reason="genuine Eigen build (oracle/_ref) not present")
@pytest.mark.parametrize("n", [40, 160, 300, 513])
def test_against_genuine_eigen(n):
    a = workloads.gram_spd(n, seed=11 * n)
    b = workloads.rhs(n, 5, seed=n)
    m, idx = oracle.llt_blocked(a)
    e = oracle.eigen_llt(a, b)
    assert idx == -1 and e["info"] == 0
    np.testing.assert_allclose(np.tril(m), np.tril(e["L"]), rtol=0, atol=1e-13)
    x = oracle.llt_solve(np.tril(m), b)
    np.testing.assert_allclose(x, e["X"], rtol=0, atol=1e-10)
    ei, eidx = oracle.eigen_llt_inplace(workloads.indefinite_at(n, n // 3, seed=n))
    assert eidx == oracle.llt_blocked(workloads.indefinite_at(n, n // 3, seed=n))[1] == n // 3
