"""Mixed-precision factorization + FP64 iterative refinement (mi_dposv_mixed).

No reference counterpart; pinned by the FP64 acceptance test computed here independently
in torch FP64 (||A x - b|| / (||A||_F ||x||) <= 1e-14 per right-hand side), by agreement
with the oracle's FP64 solve, and by the closed-form spectrum of the test matrices
(A = Q diag(lambda) Q^T with lambda log-uniform in [1, kappa])."""
import numpy as np
import pytest
import torch

import mi_llt
import oracle
from mi_llt import workloads

pytestmark = pytest.mark.gpu

TOL = 1e-14


def berr_fp64(a, x, b):
    """max_j ||A x_j - b_j||_2 / (||A||_F ||x_j||_2), A symmetric from its lower triangle."""
    asym = torch.tril(a) + torch.tril(a, -1).t()
    r = asym @ x - b
    return float((r.norm(dim=0) / (torch.linalg.matrix_norm(asym) * x.norm(dim=0))).max())


@pytest.mark.parametrize("prec", ["fp32", "bf16", "bf16x3", "bf16x6", "fp64"])
@pytest.mark.parametrize("n,kappa", [(300, 1e2), (1000, 1e4), (1000, 1e6), (2049, 1e6)])
def test_mixed_reaches_fp64_backward_error(dev, prec, n, kappa):
    print(f"[gpu] mixed prec={prec} n={n} kappa={kappa:g}", flush=True)
    a = workloads.spectrum_spd_torch(n, kappa, seed=3, device=dev)
    b = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, 5, seed=4)).to(dev))
    a0, b0 = a.clone(), b.clone()
    x, iters, berr, info = mi_llt.posv_mixed(a, b, prec=prec, tol=TOL, maxit=30)
    assert info == 0
    assert torch.equal(a, a0) and torch.equal(b, b0)  # inputs untouched
    assert berr <= TOL
    assert berr_fp64(a, x, b) <= 2 * TOL  # the library's own measure, recomputed independently
    if prec == "fp64":
        assert iters == 0
    if prec == "fp32":
        # kappa * u_fp32 <= 0.06: classical refinement converges without the FP64 fallback
        assert iters >= 1
    if prec == "bf16x6":
        # the 3-part split carries all 24 bits of the FP32 working copy, and its products are
        # accumulated in FP32 over slabs of up to 512 columns: it converges like the FP32
        # factor of the same matrix, within one more correction
        _, iters32, _, info32 = mi_llt.posv_mixed(a, b, prec="fp32", tol=TOL, maxit=30)
        assert info32 == 0 and iters32 >= 1
        assert 1 <= iters <= iters32 + 1, (iters, iters32)
    if prec == "bf16x3" and kappa <= 1e4:
        # split BF16 (~16 significant bits): a few corrections, no fallback -- a factor from
        # wrongly split operands would not converge this fast
        assert 1 <= iters <= 3, iters
    assert iters >= -2


def test_mixed_agrees_with_oracle_solution(dev):
    n, r = 257, 3
    a_h = workloads.spectrum_spd(n, 1e3, seed=5)
    b_h = workloads.rhs(n, r, seed=6)
    x, iters, berr, info = mi_llt.posv_mixed(mi_llt.colmajor(torch.from_numpy(a_h).to(dev)),
                                            mi_llt.colmajor(torch.from_numpy(b_h).to(dev)), prec="fp32")
    assert info == 0 and iters >= 0 and berr <= TOL
    m, _ = oracle.llt_blocked(a_h)
    x_ref = oracle.llt_solve(m, b_h)
    # both are backward stable to ~1e-15; forward error is bounded by kappa * berr
    rel = np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
    assert rel <= 1e3 * 1e-13


def test_mixed_upper_storage(dev):
    n = 700
    a = workloads.spectrum_spd_torch(n, 1e4, seed=7, device=dev)
    b = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, 2, seed=8)).to(dev))
    up = mi_llt.colmajor(torch.triu(a) + 7.0 * torch.tril(torch.ones_like(a), -1))  # junk below
    x, iters, berr, info = mi_llt.posv_mixed(up, b, prec="fp32", uplo="U")
    assert info == 0 and iters >= 0 and berr <= TOL
    assert berr_fp64(a, x, b) <= 2 * TOL


def test_mixed_not_positive_definite(dev):
    n, k = 300, 129
    a = mi_llt.colmajor(torch.from_numpy(workloads.indefinite_at(n, k)).to(dev))
    b = mi_llt.colmajor(torch.ones(n, 1, dtype=torch.float64, device=dev))
    x = mi_llt.colmajor(torch.full((n, 1), 3.0, dtype=torch.float64, device=dev))
    _, iters, berr, info = mi_llt.posv_mixed(a, b, x=x, prec="fp32")
    assert info == k + 1      # the FP64 fallback reports LAPACK's k+1 (Eigen's column k)
    assert iters == -1        # the FP32 factor already failed
    assert torch.all(x == 3.0)  # X unchanged on failure


def test_mixed_logdet_closed_form(dev):
    # the FP32 factor's diagonal is not returned; the FP64 path's is, and must give
    # logdet(A) = sum log lambda_i for the planted spectrum
    n, kappa = 1000, 1e6
    a = workloads.spectrum_spd_torch(n, kappa, seed=9, device=dev)
    lam = workloads.spectrum(n, kappa, seed=9)
    l = a.clone()
    assert mi_llt.potrf_(l) == 0
    logdet = 2.0 * float(torch.log(torch.diagonal(l)).sum())
    assert abs(logdet - float(np.log(lam).sum())) <= 1e-10 * abs(float(np.log(lam).sum()))


def test_mixed_empty_and_arguments(dev):
    a = mi_llt.colmajor(torch.eye(4, dtype=torch.float64, device=dev))
    b = mi_llt.colmajor(torch.ones(4, 0, dtype=torch.float64, device=dev))
    x, iters, berr, info = mi_llt.posv_mixed(a, b)
    assert (iters, berr, info) == (0, 0.0, 0)
    with pytest.raises(mi_llt.MiLLTError):
        mi_llt.posv_mixed(a, mi_llt.colmajor(torch.ones(4, 1, dtype=torch.float64, device=dev)), tol=-1.0)
