"""GPU parity of the FP64 factor + solve against the oracle and Eigen's golden outputs.

Bar (floating point): normwise backward errors ||A - L L^T||_F / ||A||_F and
max_c ||A x_c - b_c|| / (||A||_F ||x_c||) at most 1e-14 (long-double metric from the
oracle at n <= 1000); element-wise agreement of L with the oracle and with Eigen's L
within 1e-12 absolute (entries are O(1)); identical failing column (LAPACK info) and an
untouched other triangle, bit for bit.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

import mi_llt
import oracle
from mi_llt import workloads

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "eigen_llt_golden.npz")
TOL_BERR = 1e-14
TOL_L = 1e-12


def say(*a):
    print(*a, flush=True)


def to_dev(x, dev):
    return mi_llt.colmajor(torch.from_numpy(np.asarray(x, dtype=np.float64)).to(dev))


def assert_status_ok():
    assert mi_llt.default_handle().status() == 0, "a solve sweep reported a timeout"


def gpu_factor(a, dev, uplo="L"):
    da = to_dev(a, dev)
    info = mi_llt.potrf_(da, uplo)
    return da.cpu().numpy(), info


@pytest.mark.parametrize("n", [1, 2, 3, 8, 16, 31, 32, 33, 100, 127, 128, 129, 255, 256, 257, 300, 640, 1000])
def test_factor_matches_oracle(dev, n):
    say(f"[gpu] factor n={n}")
    a = workloads.gram_spd(n, seed=n)
    m, info = gpu_factor(a, dev)
    assert info == 0
    lo, idx = oracle.llt_blocked(a)
    assert idx == -1
    assert np.array_equal(np.triu(m, 1), np.triu(a, 1)), "upper triangle modified"
    np.testing.assert_allclose(np.tril(m), np.tril(lo), rtol=0, atol=TOL_L)
    assert oracle.factor_berr(a, np.tril(m)) <= TOL_BERR


@pytest.mark.parametrize("tag", ["kat3"] + [f"spd{n}" for n in (1, 2, 3, 7, 8, 16, 31, 32, 33, 64, 100, 127,
                                                                128, 129, 200)])
def test_factor_solve_matches_eigen_golden(dev, tag):
    say(f"[gpu] golden {tag}")
    g = np.load(GOLDEN)
    a, b = g[f"{tag}/A"], g[f"{tag}/B"]
    da, db = to_dev(a, dev), to_dev(b, dev)
    d_info = torch.zeros(1, dtype=torch.int32, device=dev)
    mi_llt.posv_async_(da, db, d_info)
    torch.cuda.synchronize()
    assert_status_ok()
    assert int(d_info.item()) == 0
    m, x = da.cpu().numpy(), db.cpu().numpy()
    np.testing.assert_array_equal(np.triu(m, 1), np.triu(g[f"{tag}/LLT"], 1))
    np.testing.assert_allclose(np.tril(m), np.tril(g[f"{tag}/LLT"]), rtol=0, atol=TOL_L)
    np.testing.assert_allclose(x, g[f"{tag}/X"], rtol=0, atol=1e-10 * max(1.0, np.abs(x).max()))
    assert oracle.solve_berr(a, x, b) <= TOL_BERR


@pytest.mark.parametrize("n,k", [(1, 0), (5, 3), (33, 0), (33, 32), (100, 17), (129, 128), (200, 150)])
def test_failing_column_matches_eigen(dev, n, k):
    say(f"[gpu] fail n={n} k={k}")
    g = np.load(GOLDEN)
    a = g[f"fail{n}_{k}/A"]
    _, info = gpu_factor(a, dev)
    assert info == int(g[f"fail{n}_{k}/IDX"]) + 1 == k + 1


@pytest.mark.parametrize("n,k", [(300, 0), (300, 129), (513, 511), (1000, 640)])
def test_failing_column_matches_oracle(dev, n, k):
    say(f"[gpu] fail n={n} k={k}")
    a = workloads.indefinite_at(n, k, seed=n)
    _, info = gpu_factor(a, dev)
    assert info == oracle.llt_blocked(a)[1] + 1 == k + 1


def test_negative_definite_and_nan(dev):
    say("[gpu] negative definite / nan")
    a = -workloads.gram_spd(50, seed=2)
    assert gpu_factor(a, dev)[1] == 1
    # Eigen's pivot test is `x <= 0` (LLT.h:327): a NaN pivot does not stop the factorization
    a = workloads.gram_spd(40, seed=2)
    a[5, 5] = np.nan
    assert gpu_factor(a, dev)[1] == oracle.llt_blocked(a)[1] + 1 == 0


@pytest.mark.parametrize("n", [3, 33, 129, 300])
def test_upper_matches_golden_and_oracle(dev, n):
    say(f"[gpu] upper n={n}")
    a = workloads.gram_spd(n, seed=100 + n)
    m, info = gpu_factor(a, dev, "U")
    assert info == 0
    mo, _ = oracle.llt_blocked(a, upper=True)
    assert np.array_equal(np.tril(m, -1), np.tril(a, -1)), "lower triangle modified"
    np.testing.assert_allclose(np.triu(m), np.triu(mo), rtol=0, atol=TOL_L)
    if n in (33, 129):
        g = np.load(GOLDEN)
        np.testing.assert_allclose(np.triu(m), np.triu(g[f"upper{n}/LLT"]), rtol=0, atol=TOL_L)


# The solve's chain takes 0, 2 or 3 coupled predecessors by its block count (below 8 blocks,
# 8-15, 16 and more): 1000 (8 blocks, last one 104 rows), 2049 (17 blocks, last one a single
# row, so one 64-row half of it is empty) and 2176 (17 whole blocks) reach the coupled paths.
@pytest.mark.parametrize("n,r", [(1, 1), (100, 1), (129, 3), (300, 16), (513, 64), (640, 65), (257, 130),
                                 (1000, 17), (2049, 40), (2176, 1)])
def test_solve_matches_oracle(dev, n, r):
    say(f"[gpu] solve n={n} r={r}")
    a = workloads.gram_spd(n, seed=n + 1)
    b = workloads.rhs(n, r, seed=n)
    da = to_dev(a, dev)
    assert mi_llt.potrf_(da) == 0
    db = to_dev(b, dev)
    mi_llt.potrs_(da, db)
    torch.cuda.synchronize()
    assert_status_ok()
    x = db.cpu().numpy()
    xo = oracle.llt_solve(np.tril(oracle.llt_blocked(a)[0]), b)
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-10 * max(1.0, np.abs(xo).max()))
    assert oracle.solve_berr(a, x, b) <= TOL_BERR


def test_upper_solve(dev):
    say("[gpu] upper solve")
    n, r = 200, 5
    a, b = workloads.gram_spd(n, seed=9), workloads.rhs(n, r, seed=9)
    da = to_dev(a, dev)
    assert mi_llt.potrf_(da, "U") == 0
    db = to_dev(b, dev)
    mi_llt.potrs_(da, db, "U")
    torch.cuda.synchronize()
    assert oracle.solve_berr(a, db.cpu().numpy(), b) <= TOL_BERR


def test_leading_dimension_and_empty(dev):
    say("[gpu] lda > n, n = 0")
    n, lda = 150, 173
    a = workloads.gram_spd(n, seed=4)
    big = np.full((lda, n), 7.25)
    big[:n, :] = a
    dbig = to_dev(big, dev)
    sub = dbig[:n, :]
    assert sub.stride() == (1, lda)
    assert mi_llt.potrf_(sub) == 0
    out = dbig.cpu().numpy()
    assert np.all(out[n:, :] == 7.25), "padding rows below n were written"
    np.testing.assert_allclose(np.tril(out[:n]), np.tril(oracle.llt_blocked(a)[0]), rtol=0, atol=TOL_L)
    e = torch.empty_strided((0, 0), (1, 1), dtype=torch.float64, device=dev)
    assert mi_llt.potrf_(e) == 0


def test_l1_norm(dev):
    say("[gpu] l1 norm")
    a = workloads.gram_spd(333, seed=8)
    junk = np.tril(a) + 5.0 * np.triu(np.ones_like(a), 1)
    v = float(mi_llt.sy_l1norm(to_dev(junk, dev), "L"))
    assert v == pytest.approx(oracle.l1_norm_lower(junk), rel=1e-14)
    v = float(mi_llt.sy_l1norm(to_dev(junk.T.copy(), dev), "U"))
    assert v == pytest.approx(oracle.l1_norm_lower(junk), rel=1e-14)


def test_llt_object_mirror(dev):
    say("[gpu] LLT object")
    n = 260
    a = workloads.gram_spd(n, seed=21)
    b = workloads.rhs(n, 2, seed=22)
    llt = mi_llt.LLT(torch.from_numpy(a).to(dev))
    assert llt.info() == mi_llt.ComputationInfo.Success
    x = llt.solve(torch.from_numpy(b).to(dev)).cpu().numpy()
    assert oracle.solve_berr(a, x, b) <= TOL_BERR
    lm = llt.matrixL().cpu().numpy()
    assert not np.triu(lm, 1).any()  # matrixL() materialises zeros above (TriangularMatrix.h:118)
    np.testing.assert_allclose(llt.matrixU().cpu().numpy(), lm.T, atol=0)  # cholesky.cpp:122-125
    np.testing.assert_array_equal(np.triu(llt.matrixLLT().cpu().numpy(), 1), np.triu(a, 1))
    assert llt.l1_norm() == pytest.approx(oracle.l1_norm_lower(a), rel=1e-14)
    bad = mi_llt.LLT(torch.from_numpy(-a).to(dev))
    assert bad.info() == mi_llt.ComputationInfo.NumericalIssue
    with pytest.raises(RuntimeError):
        mi_llt.LLT(5).info()


def test_lapacke_shim_host_pointers(dev):
    say("[gpu] LAPACKE / Fortran shims")
    lib = mi_llt.lib()
    n, r = 190, 4
    a = workloads.gram_spd(n, seed=31)
    b = workloads.rhs(n, r, seed=32)
    lo = np.tril(oracle.llt_blocked(a)[0])
    # column-major, lower (the EIGEN_USE_LAPACKE call of LLT_LAPACKE.h:59-61)
    m = np.asfortranarray(a.copy())
    assert lib.LAPACKE_dpotrf(102, b"L", n, m.ctypes.data, n) == 0
    np.testing.assert_allclose(np.tril(m), lo, rtol=0, atol=TOL_L)
    assert np.array_equal(np.triu(m, 1), np.triu(a, 1))
    x = np.asfortranarray(b.copy())
    assert lib.LAPACKE_dpotrs(102, b"L", n, r, m.ctypes.data, n, x.ctypes.data, n) == 0
    assert oracle.solve_berr(a, x, b) <= TOL_BERR
    # row-major lower == column-major upper of the same bytes
    mr = np.ascontiguousarray(a.copy())
    assert lib.LAPACKE_dpotrf(101, b"L", n, mr.ctypes.data, n) == 0
    np.testing.assert_allclose(np.tril(mr), lo, rtol=0, atol=TOL_L)
    xr = np.ascontiguousarray(b.copy())
    assert lib.LAPACKE_dpotrs(101, b"L", n, r, mr.ctypes.data, n, xr.ctypes.data, r) == 0
    assert oracle.solve_berr(a, xr, b) <= TOL_BERR
    # Fortran ABI (lapack/cholesky.cpp:14-72): info = k+1 on a failing pivot
    f = np.asfortranarray(workloads.indefinite_at(n, 77, seed=n))
    info, nn, ld = ctypes.c_int(0), ctypes.c_int(n), ctypes.c_int(n)
    lib.dpotrf_(b"L", ctypes.byref(nn), f.ctypes.data, ctypes.byref(ld), ctypes.byref(info))
    assert info.value == 78


@pytest.mark.parametrize("n,r", [(8192, 64), (12288, 16), (16384, 16)])
def test_full_size_backward_error(dev, n, r):
    """The bench workload (n=8192, 64 right-hand sides) and the larger orders of the mixed
    workload: backward errors from FP64 GPU products (torch, test-side only) -- size-
    independent bar of 1e-14.  (Many factorization steps with many workgroups in flight:
    a publication race between the diagonal block and the panel solve showed here.)"""
    say(f"[gpu] full size n={n} r={r}")
    a = workloads.gram_spd_torch(n, seed=0, device=dev)
    b = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, r, seed=1)).to(dev))
    da, db = a.clone(), b.clone()
    d_info = torch.zeros(1, dtype=torch.int32, device=dev)
    mi_llt.posv_async_(da, db, d_info)
    torch.cuda.synchronize()
    assert_status_ok()
    assert int(d_info.item()) == 0
    l = torch.tril(da)
    asym = torch.tril(a) + torch.tril(a, -1).t()
    fb = float(torch.linalg.matrix_norm(asym - l @ l.t()) / torch.linalg.matrix_norm(asym))
    res = asym @ db - b
    sb = float((res.norm(dim=0) / (torch.linalg.matrix_norm(asym) * db.norm(dim=0))).max())
    say(f"  factor berr {fb:.3e} solve berr {sb:.3e}")
    assert fb <= TOL_BERR and sb <= TOL_BERR
    assert torch.equal(torch.triu(da, 1), torch.triu(a, 1))
