"""Reverse mode: mi_dpotrf_adjoint / mi_dpotrs_adjoint behind torch.autograd.Function.

No reference counterpart; pinned by (1) central finite differences in FP64 of the forward
path itself, (2) the closed-form Gaussian-process gradient dL/dK = 1/2 (a a^T - K^{-1}),
a = K^{-1} y, and (3) torch.linalg.cholesky's own backward (library routine used as a
checker only)."""
import numpy as np
import pytest
import torch

import mi_llt
from mi_llt import autograd as mad
from mi_llt import workloads

pytestmark = pytest.mark.gpu


def spd(n, seed, dev):
    return torch.from_numpy(workloads.gram_spd(n, seed=seed)).to(dev)


@pytest.mark.parametrize("n", [1, 5, 37, 130, 300, 384, 1000])
def test_cholesky_adjoint_matches_torch(dev, n):
    print(f"[gpu] adjoint vs torch n={n}", flush=True)
    a = spd(n, 7, dev)
    w = torch.from_numpy(np.random.default_rng(n).standard_normal((n, n))).to(dev)
    a1 = a.clone().requires_grad_(True)
    (mad.cholesky(a1) * w).sum().backward()
    a2 = a.clone().cpu().requires_grad_(True)
    (torch.linalg.cholesky(a2) * w.cpu()).sum().backward()
    g1, g2 = a1.grad.cpu(), a2.grad
    assert torch.allclose(g1, g1.t(), rtol=0, atol=1e-12 * g2.abs().max())
    assert (g1 - g2).abs().max() <= 1e-10 * g2.abs().max()


@pytest.mark.parametrize("n,samples", [(24, 60), (130, 24), (300, 12)])
def test_cholesky_adjoint_finite_differences(dev, n, samples):
    print(f"[gpu] adjoint finite differences n={n}", flush=True)
    a = spd(n, 3, dev)
    w = torch.from_numpy(np.tril(np.random.default_rng(1).standard_normal((n, n)))).to(dev)
    a1 = a.clone().requires_grad_(True)
    (mad.cholesky(a1) * w).sum().backward()
    g = a1.grad
    f = lambda m: float((mad.cholesky(m) * w).sum())  # noqa: E731
    eps = 1e-6
    rng = np.random.default_rng(2)
    for _ in range(samples):
        i, j = sorted(rng.integers(0, n, size=2))[::-1]
        e = torch.zeros_like(a)
        e[i, j] = 1.0
        e[j, i] = 1.0
        fd = (f(a + eps * e) - f(a - eps * e)) / (2 * eps)
        an = float(g[i, j] + g[j, i]) if i != j else float(g[i, i])
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (i, j, fd, an)


def test_solve_adjoint_matches_torch(dev):
    n, r = 150, 3
    a = spd(n, 11, dev)
    b = torch.from_numpy(np.random.default_rng(5).standard_normal((n, r))).to(dev)
    wx = torch.from_numpy(np.random.default_rng(6).standard_normal((n, r))).to(dev)
    a1, b1 = a.clone().requires_grad_(True), b.clone().requires_grad_(True)
    (mad.cholesky_solve(b1, mad.cholesky(a1)) * wx).sum().backward()
    a2, b2 = a.clone().cpu().requires_grad_(True), b.clone().cpu().requires_grad_(True)
    (torch.cholesky_solve(b2, torch.linalg.cholesky(a2)) * wx.cpu()).sum().backward()
    assert (a1.grad.cpu() - a2.grad).abs().max() <= 1e-10 * a2.grad.abs().max()
    assert (b1.grad.cpu() - b2.grad).abs().max() <= 1e-10 * b2.grad.abs().max()


@pytest.mark.parametrize("n", [512, 4096])
def test_gp_gradient_closed_form(dev, n):
    print(f"[gpu] GP gradient n={n}", flush=True)
    k, y = workloads.gp_kernel_torch(n, seed=0, device=dev)
    k1 = k.clone().requires_grad_(True)
    ll = mad.gp_log_marginal(k1, y)
    ll.backward()
    # closed form from torch's FP64 inverse (an independent library routine, checker only)
    ksym = torch.tril(k) + torch.tril(k, -1).t()
    kinv = torch.linalg.inv(ksym)
    alpha = (kinv @ y).reshape(-1, 1)
    ref = 0.5 * (alpha @ alpha.t() - kinv)
    err = float((k1.grad - ref).abs().max() / ref.abs().max())
    assert err <= 1e-9, err
    # log marginal value against torch's slogdet
    ll_ref = -0.5 * float(y @ alpha.reshape(-1)) - 0.5 * float(torch.linalg.slogdet(k)[1]) \
        - 0.5 * n * np.log(2 * np.pi)
    assert abs(float(ll) - ll_ref) <= 1e-9 * abs(ll_ref)
