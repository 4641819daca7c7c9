"""Seeded synthetic inputs of the LLT workloads (host numpy; optional GPU assembly).

Random numbers come from numpy's PCG64 ``Generator(seed).standard_normal``, so the same
seed gives the same matrix on every machine.  (Eigen's own ``Random`` is ``std::rand``,
Eigen/src/Core/MathFunctions.h:650-660, uniform and libc-specific, so it cannot serve.)

* ``gram_spd``     A = B B^T / n + 0.1 I, B iid N(0,1)           (factor+solve workloads)
* ``spectrum_spd`` A = Q diag(lambda) Q^T, Q a product of seeded Householder reflectors,
                   lambda log-uniform in [1, kappa] with 1 and kappa planted
* ``rhs``          B iid N(0,1)
* ``indefinite_at`` SPD matrix whose leading minor of order k+1 is not positive definite
"""
from __future__ import annotations

import numpy as np


def gram_spd(n: int, seed: int = 0, shift: float = 0.1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, n))
    a = b @ b.T
    a /= max(n, 1)
    a[np.diag_indices(n)] += shift
    return np.asfortranarray(a)


def gram_spd_torch(n: int, seed: int = 0, shift: float = 0.1, device="cuda"):
    """Same matrix as :func:`gram_spd`, with the product formed on the device (setup only)."""
    import torch

    rng = np.random.default_rng(seed)
    b = torch.from_numpy(rng.standard_normal((n, n))).to(device)
    a = torch.matmul(b, b.t().contiguous())
    a /= max(n, 1)
    a.diagonal().add_(shift)
    del b
    out = torch.empty_strided((n, n), (1, n), dtype=torch.float64, device=device)
    out.copy_(a)
    return out


def rhs(n: int, r: int, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((n, r)))


def spectrum(n: int, kappa: float, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed + 1000)
    u = rng.random(n)
    if n >= 2:
        u[0], u[1] = 0.0, 1.0
    return kappa ** u


def householder_q_apply(x: np.ndarray, n_reflectors: int, seed: int) -> np.ndarray:
    """Q x with Q = H_1 ... H_m, H_i = I - 2 v v^T / v^T v, v_i iid N(0,1) from ``seed``."""
    rng = np.random.default_rng(seed + 2000)
    vs = rng.standard_normal((n_reflectors, x.shape[0]))
    y = np.array(x, dtype=np.float64, copy=True)
    for v in vs[::-1]:
        y -= np.outer(v, (2.0 / (v @ v)) * (v @ y))
    return y


def spectrum_spd(n: int, kappa: float, seed: int = 0, n_reflectors: int = 64) -> np.ndarray:
    lam = spectrum(n, kappa, seed)
    q = householder_q_apply(np.eye(n), n_reflectors, seed)
    a = (q * lam) @ q.T
    a = 0.5 * (a + a.T)
    return np.asfortranarray(a)


def spectrum_spd_torch(n: int, kappa: float, seed: int = 0, n_reflectors: int = 64, device="cuda"):
    """Same matrix as :func:`spectrum_spd`, formed as H_1 (... (H_m diag(lambda) H_m) ...) H_1 with
    two-sided reflector updates (O(m n^2), no dense product), column-major float64 on ``device``:
    for H = I - tau v v^T, p = tau A v, w = p - (tau/2)(v^T p) v, A <- A - v w^T - w v^T."""
    import torch

    lam = torch.from_numpy(spectrum(n, kappa, seed)).to(device)
    rng = np.random.default_rng(seed + 2000)
    vs = torch.from_numpy(rng.standard_normal((n_reflectors, n))).to(device)
    a = torch.diag(lam)
    for i in range(n_reflectors - 1, -1, -1):
        v = vs[i]
        tau = 2.0 / torch.dot(v, v)
        p = tau * (a @ v)
        w = p - (0.5 * tau * torch.dot(v, p)) * v
        a -= torch.outer(v, w)
        a -= torch.outer(w, v)
    a = 0.5 * (a + a.t())
    out = torch.empty_strided((n, n), (1, n), dtype=torch.float64, device=device)
    out.copy_(a)
    return out


def indefinite_at(n: int, k: int, seed: int = 0) -> np.ndarray:
    """A copy of gram_spd(n) whose pivot k is made negative: a_kk is replaced by
    ||L[k, :k]||^2 - 1, so the leading minors of order <= k are PD and order k+1 is not
    (Eigen returns column k, LAPACK info k+1)."""
    a = gram_spd(n, seed)
    if k == 0:
        a[0, 0] = -1.0
        return a
    l = np.linalg.cholesky(a[:k, :k])
    y = np.linalg.solve(l, a[:k, k])
    a[k, k] = y @ y - 1.0
    return np.asfortranarray(a)


def gp_kernel_torch(n: int, d: int = 8, ell: float = 1.5, sigma2: float = 1e-2, seed: int = 0, device="cuda"):
    """The GP workload: X n x d iid N(0,1) (PCG64 ``seed``), K = exp(-||x_i - x_j||^2 / (2 ell^2))
    + sigma2 I, y ~ N(0,1) (seed + 1).  Returns (K column-major float64, y)."""
    import torch

    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.standard_normal((n, d))).to(device)
    y = torch.from_numpy(np.random.default_rng(seed + 1).standard_normal(n)).to(device)
    sq = (x * x).sum(1)
    d2 = (sq[:, None] + sq[None, :] - 2.0 * (x @ x.t())).clamp_min(0.0)
    k = torch.exp(-d2 / (2.0 * ell * ell))
    k.diagonal().add_(sigma2)
    out = torch.empty_strided((n, n), (1, n), dtype=torch.float64, device=device)
    out.copy_(0.5 * (k + k.t()))
    return out, y
