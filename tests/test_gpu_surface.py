"""GPU parity of the rest of the LLT surface and of its BLAS/LAPACK bindings, against the
oracle's restatements (pinned to genuine Eigen in tests/test_oracle_surface.py):

* single precision (LLT<MatrixXf>: mi_spotrf / mi_spotrs, LAPACKE_spotrf / spotrs, spotrf_ /
  spotrs_): FP32 factor within FP32 bounds of the oracle's FP32 restatement and of the FP64
  factor, identical failing columns, untouched other triangle;
* mi_dtrsm / dtrsm_ for all 16 (side, uplo, trans, diag) cases and mi_dsyrk / dsyrk_ for all
  4 (uplo, trans) cases, against the oracle's substitution and product (FP64: 1e-12 relative);
* rcond (mi_dpocon): the same estimate as the oracle (and Eigen) to 1e-10 relative, exact on
  diagonal matrices;
* rank-one update / downdate (mi_dllt_rank_update): the oracle's factor to 1e-12 relative,
  L L^T = A + sigma v v^T with a backward error <= 1e-14, identical failing column.
"""
import ctypes

import numpy as np
import pytest
import torch

import mi_llt
import oracle
from mi_llt import workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dev_t(x, dev, dtype=torch.float64):
    return mi_llt.colmajor(torch.from_numpy(np.asarray(x)).to(device=dev, dtype=dtype))


# ------------------------------------------------------------------ single precision
@pytest.mark.parametrize("n", [1, 7, 31, 32, 100, 128, 129, 300, 1000])
def test_spotrf_matches_fp32_oracle(dev, n):
    say(f"[gpu] spotrf n={n}")
    a = workloads.gram_spd(n, seed=n + 3)
    da = dev_t(a, dev, torch.float32)
    assert mi_llt.spotrf_(da) == 0
    m = da.cpu().numpy()
    lo32, idx = oracle.llt_blocked_f32(a)
    assert idx == -1
    a32 = a.astype(np.float32)
    assert np.array_equal(np.triu(m, 1), np.triu(a32, 1)), "upper triangle modified"
    # two FP32 factorizations of a matrix with kappa ~ 40 agree to a few FP32 ulps of kappa
    np.testing.assert_allclose(np.tril(m), np.tril(lo32), rtol=0, atol=2e-5)
    lf = np.tril(m).astype(np.float64)
    assert oracle.factor_berr(a32.astype(np.float64), lf) <= 2e-7 * max(1.0, np.sqrt(n) / 4)


@pytest.mark.parametrize("n,k", [(50, 0), (100, 17), (300, 129), (513, 511)])
def test_spotrf_failing_column(dev, n, k):
    a = workloads.indefinite_at(n, k, seed=n)
    da = dev_t(a, dev, torch.float32)
    assert mi_llt.spotrf_(da) == oracle.llt_blocked_f32(a)[1] + 1 == k + 1


@pytest.mark.parametrize("n,r", [(64, 1), (300, 5), (1025, 20)])
def test_spotrs_and_upper(dev, n, r):
    say(f"[gpu] spotrs n={n} r={r}")
    a = workloads.gram_spd(n, seed=2 * n)
    b = workloads.rhs(n, r, seed=n)
    for uplo in "LU":
        da = dev_t(a, dev, torch.float32)
        assert mi_llt.spotrf_(da, uplo) == 0
        db = dev_t(b, dev, torch.float32)
        mi_llt.spotrs_(da, db, uplo)
        assert mi_llt.default_handle().status() == 0
        x = db.cpu().numpy().astype(np.float64)
        assert oracle.solve_berr(a.astype(np.float32).astype(np.float64), x,
                                 b.astype(np.float32).astype(np.float64)) <= 1e-6
        if uplo == "U":
            assert np.array_equal(np.tril(da.cpu().numpy(), -1), np.tril(a.astype(np.float32), -1))


def test_float_lapacke_and_fortran_shims(dev):
    say("[gpu] LAPACKE_spotrf / spotrs, spotrf_ / spotrs_")
    lib = mi_llt.lib()
    n, r = 190, 3
    a = workloads.gram_spd(n, seed=41).astype(np.float32)
    b = workloads.rhs(n, r, seed=42).astype(np.float32)
    lo32 = np.tril(oracle.llt_blocked_f32(a)[0])
    m = np.asfortranarray(a.copy())
    assert lib.LAPACKE_spotrf(102, b"L", n, m.ctypes.data, n) == 0
    np.testing.assert_allclose(np.tril(m), lo32, rtol=0, atol=2e-5)
    x = np.asfortranarray(b.copy())
    assert lib.LAPACKE_spotrs(102, b"L", n, r, m.ctypes.data, n, x.ctypes.data, n) == 0
    assert oracle.solve_berr(a.astype(np.float64), x.astype(np.float64), b.astype(np.float64)) <= 1e-6
    mr = np.ascontiguousarray(a.copy())  # row-major lower == column-major upper
    assert lib.LAPACKE_spotrf(101, b"L", n, mr.ctypes.data, n) == 0
    np.testing.assert_allclose(np.tril(mr), lo32, rtol=0, atol=2e-5)
    f = np.asfortranarray(workloads.indefinite_at(n, 77, seed=n).astype(np.float32))
    info, nn, ld = ctypes.c_int(0), ctypes.c_int(n), ctypes.c_int(n)
    lib.spotrf_(b"L", ctypes.byref(nn), f.ctypes.data, ctypes.byref(ld), ctypes.byref(info))
    assert info.value == 78
    m2 = np.asfortranarray(a.copy())
    lib.spotrf_(b"L", ctypes.byref(nn), m2.ctypes.data, ctypes.byref(ld), ctypes.byref(info))
    assert info.value == 0
    x2 = np.asfortranarray(b.copy())
    rr = ctypes.c_int(r)
    lib.spotrs_(b"L", ctypes.byref(nn), ctypes.byref(rr), m2.ctypes.data, ctypes.byref(ld), x2.ctypes.data,
                ctypes.byref(ld), ctypes.byref(info))
    assert info.value == 0
    np.testing.assert_allclose(x2, x, rtol=0, atol=1e-5 * np.abs(x).max())


def test_llt_float_mirror(dev):
    n = 200
    a = workloads.gram_spd(n, seed=51)
    b = workloads.rhs(n, 2, seed=52)
    llt = mi_llt.LLT(torch.from_numpy(a).to(dev), dtype=torch.float32)
    assert llt.info() == mi_llt.ComputationInfo.Success
    x = llt.solve(torch.from_numpy(b).to(dev)).cpu().numpy()
    assert x.dtype == np.float32
    assert oracle.solve_berr(a, x.astype(np.float64), b) <= 1e-6


# ------------------------------------------------------------------ trsm / syrk
VARIANTS = [(s, u, t, d) for s in "LR" for u in "LU" for t in "NT" for d in "NU"]


def tri_operand(k, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((k, k))
    a[np.diag_indices(k)] = 4.0 + np.abs(a[np.diag_indices(k)])
    return np.asfortranarray(a)


@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS)
def test_trsm_all_variants(dev, side, uplo, trans, diag):
    say(f"[gpu] trsm {side}{uplo}{trans}{diag}")
    for (m, n), alpha in (((300, 20), 1.0), ((129, 257), -0.5), ((5, 3), 2.0)):
        k = m if side == "L" else n
        a = tri_operand(k, k)
        b = np.asfortranarray(np.random.default_rng(m + n).standard_normal((m, n)))
        da, db = dev_t(a, dev), dev_t(b, dev)
        mi_llt.trsm_(da, db, side, uplo, trans, diag, alpha)
        assert mi_llt.default_handle().status() == 0
        ref = oracle.trsm(side, uplo, trans, diag, alpha, a, b)
        got = db.cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        assert torch.equal(da.cpu(), torch.from_numpy(a)), "A modified"


def test_trsm_alpha_zero_and_empty(dev):
    a = tri_operand(40, 1)
    b = np.full((40, 3), np.nan)
    db = dev_t(b, dev)
    mi_llt.trsm_(dev_t(a, dev), db, "L", "L", "N", "N", 0.0)
    assert torch.all(db == 0.0)
    e = torch.empty_strided((0, 5), (1, 1), dtype=torch.float64, device=dev)
    mi_llt.trsm_(dev_t(a[:0, :0], dev), e, "L", "L", "N", "N", 1.0)


@pytest.mark.parametrize("uplo,trans", [(u, t) for u in "LU" for t in "NT"])
def test_syrk_all_variants(dev, uplo, trans):
    say(f"[gpu] syrk {uplo}{trans}")
    rng = np.random.default_rng(7)
    for n, k, alpha, beta in ((300, 128, -1.0, 1.0), (257, 17, 0.5, 2.0), (130, 300, 1.0, 0.0), (9, 1, 2.0, -1.0)):
        a = np.asfortranarray(rng.standard_normal((n, k) if trans == "N" else (k, n)))
        c = np.asfortranarray(rng.standard_normal((n, n)))
        if beta == 0.0:
            tri = np.tril if uplo == "L" else np.triu
            c = c + np.where(tri(np.ones((n, n))) > 0, np.nan, 0.0)  # never read when beta == 0
        dc = dev_t(c, dev)
        mi_llt.syrk_(dev_t(a, dev), dc, uplo, trans, alpha, beta)
        ref = oracle.syrk(uplo, trans, alpha, a, beta, c)
        got = dc.cpu().numpy()
        scale = np.nanmax(np.abs(ref))
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * scale)


def test_blas_host_shims_match_eigen_blas(dev):
    say("[gpu] dtrsm_ / dsyrk_ host shims")
    lib = mi_llt.lib()
    I = ctypes.c_int  # noqa: E741
    m, n = 200, 30
    a = tri_operand(m, 3)
    b = np.asfortranarray(np.random.default_rng(8).standard_normal((m, n)))
    x = b.copy(order="F")
    assert lib.dtrsm_(b"L", b"U", b"T", b"N", ctypes.byref(I(m)), ctypes.byref(I(n)),
                      ctypes.byref(ctypes.c_double(1.5)), a.ctypes.data, ctypes.byref(I(m)), x.ctypes.data,
                      ctypes.byref(I(m))) == 0
    ref = oracle.trsm("L", "U", "T", "N", 1.5, a, b)
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    if oracle.have_eigen_blas():
        np.testing.assert_allclose(x, oracle.eigen_blas_trsm("L", "U", "T", "N", 1.5, a, b), rtol=0,
                                   atol=1e-12 * np.abs(ref).max())
    k = 64
    p = np.asfortranarray(np.random.default_rng(9).standard_normal((m, k)))
    c = np.asfortranarray(np.random.default_rng(10).standard_normal((m, m)))
    cc = c.copy(order="F")
    assert lib.dsyrk_(b"L", b"N", ctypes.byref(I(m)), ctypes.byref(I(k)), ctypes.byref(ctypes.c_double(-1.0)),
                      p.ctypes.data, ctypes.byref(I(m)), ctypes.byref(ctypes.c_double(1.0)), cc.ctypes.data,
                      ctypes.byref(I(m))) == 0
    ref = oracle.syrk("L", "N", -1.0, p, 1.0, c)
    np.testing.assert_allclose(cc, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # an illegal argument: reported, nothing computed (blas/level3_impl.h:139-149)
    before = cc.copy()
    assert lib.dsyrk_(b"X", b"N", ctypes.byref(I(m)), ctypes.byref(I(k)), ctypes.byref(ctypes.c_double(-1.0)),
                      p.ctypes.data, ctypes.byref(I(m)), ctypes.byref(ctypes.c_double(1.0)), cc.ctypes.data,
                      ctypes.byref(I(m))) == 0
    assert np.array_equal(cc, before)


# ------------------------------------------------------------------ rcond
@pytest.mark.parametrize("n", [2, 40, 300, 1000])
def test_rcond_matches_oracle(dev, n):
    say(f"[gpu] rcond n={n}")
    a, = (workloads.gram_spd(n, seed=5 * n),)
    lo = np.tril(oracle.llt_blocked(a)[0])
    anorm = oracle.l1_norm_lower(a)
    ref = oracle.llt_rcond(lo, anorm)
    for uplo in "LU":
        da = dev_t(a, dev)
        assert mi_llt.potrf_(da, uplo) == 0
        got = mi_llt.pocon(da, anorm, uplo)
        assert got == pytest.approx(ref, rel=1e-10)
    if oracle.have_eigen() and n <= 300:
        assert got == pytest.approx(oracle.eigen_rcond(a), rel=1e-10)


def test_rcond_diagonal_exact_and_special_cases(dev):
    d = np.array([4.0, 0.25, 9.0, 1.0, 16.0, 2.0])
    l = dev_t(np.diag(np.sqrt(d)), dev)
    assert mi_llt.pocon(l, d.max()) == pytest.approx(d.min() / d.max(), rel=1e-14)
    assert mi_llt.pocon(dev_t(np.eye(4), dev), 0.0) == 0.0
    assert mi_llt.pocon(dev_t(np.array([[3.0]]), dev), 9.0) == 1.0
    e = torch.empty_strided((0, 0), (1, 1), dtype=torch.float64, device=dev)
    assert mi_llt.pocon(e, 1.0) == float("inf")


def test_llt_mirror_rcond_and_rank_update(dev):
    n = 260
    a = workloads.gram_spd(n, seed=61)
    v = workloads.rhs(n, 1, seed=62).reshape(-1)
    llt = mi_llt.LLT(torch.from_numpy(a).to(dev))
    lo = np.tril(oracle.llt_blocked(a)[0])
    assert llt.rcond() == pytest.approx(oracle.llt_rcond(lo, oracle.l1_norm_lower(a)), rel=1e-10)
    llt.rankUpdate(torch.from_numpy(v).to(dev), 0.7)
    assert llt.info() == mi_llt.ComputationInfo.Success
    ref, _ = oracle.llt_rank_update(lo, v, 0.7)
    np.testing.assert_allclose(llt.matrixL().cpu().numpy(), np.tril(ref), rtol=0, atol=1e-12)
    llt.rankUpdate(torch.from_numpy(v).to(dev), -50.0)  # makes it indefinite
    assert llt.info() == mi_llt.ComputationInfo.NumericalIssue


# ------------------------------------------------------------------ rank update
@pytest.mark.parametrize("n,sigma", [(1, 2.0), (5, 1.0), (64, 0.5), (129, 3.0), (300, -0.05), (700, 1.0),
                                     (1000, -0.2), (257, 0.0)])
def test_rank_update_matches_oracle(dev, n, sigma):
    say(f"[gpu] rank update n={n} sigma={sigma}")
    a = workloads.gram_spd(n, seed=n + 17)
    lo = np.tril(oracle.llt_blocked(a)[0])
    # a downdate keeps A + sigma v v^T positive definite when |sigma| ||v||^2 is below the
    # smallest eigenvalue (~0.1 here): ||v||^2 ~ 0.09 for the downdates
    v = workloads.rhs(n, 1, seed=n).reshape(-1) * (0.3 / np.sqrt(n) if sigma < 0 else 1.0)
    ref, idx = oracle.llt_rank_update(lo, v, sigma)
    assert idx == -1
    junk = lo + 3.0 * np.triu(np.ones((n, n)), 1)  # the other triangle must stay as it is
    dl = dev_t(junk, dev)
    info = mi_llt.rank_update_(dl, torch.from_numpy(v).to(dev), sigma)
    assert info == 0
    assert mi_llt.default_handle().status() == 0
    got = dl.cpu().numpy()
    np.testing.assert_array_equal(np.triu(got, 1), np.triu(junk, 1))
    np.testing.assert_allclose(np.tril(got), np.tril(ref), rtol=0, atol=1e-12 * np.abs(ref).max())
    target = np.tril(a) + np.tril(a, -1).T + sigma * np.outer(v, v)
    lu = np.tril(got)
    assert np.linalg.norm(lu @ lu.T - target) / np.linalg.norm(target) <= 1e-14


@pytest.mark.parametrize("n,k0,scale", [(60, 40, 30.0), (300, 200, 3.0), (1000, 700, 2.0)])
def test_rank_downdate_failing_column(dev, n, k0, scale):
    say(f"[gpu] rank downdate failure n={n}")
    a = workloads.gram_spd(n, seed=n)
    lo = np.tril(oracle.llt_blocked(a)[0])
    # v is zero above row k0: the first k0 columns go through unchanged, the failure comes later
    v = workloads.rhs(n, 1, seed=n + 1).reshape(-1) * scale
    v[:k0] = 0.0
    ref, idx = oracle.llt_rank_update(lo, v, -1.0)
    assert idx >= k0
    dl = dev_t(lo, dev)
    assert mi_llt.rank_update_(dl, torch.from_numpy(v).to(dev), -1.0) == idx + 1
    got = dl.cpu().numpy()
    # columns before the failing one updated, the rest left as they were
    np.testing.assert_allclose(got[:, :idx], ref[:, :idx], rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))
    np.testing.assert_array_equal(got[:, idx:], lo[:, idx:])


def test_rank_update_upper_and_strided_vector(dev):
    n = 300
    a = workloads.gram_spd(n, seed=71)
    lo = np.tril(oracle.llt_blocked(a)[0])
    v = workloads.rhs(n, 1, seed=72).reshape(-1)
    ref, _ = oracle.llt_rank_update(lo, v, 1.5)
    du = dev_t(lo.T.copy(), dev)  # upper factor U = L^T
    vv = torch.zeros(2 * n, dtype=torch.float64, device=dev)
    vv[::2] = torch.from_numpy(v).to(dev)
    h = mi_llt.default_handle()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    h.check(mi_llt.lib().mi_dllt_rank_update(h.ptr, b"U", n, ctypes.c_void_p(du.data_ptr()), n,
                                             ctypes.c_void_p(vv.data_ptr()), 2, 1.5,
                                             ctypes.c_void_p(info.data_ptr())), "rank update")
    assert int(info.item()) == 0
    np.testing.assert_allclose(np.triu(du.cpu().numpy()), np.tril(ref).T, rtol=0, atol=1e-12 * np.abs(ref).max())
