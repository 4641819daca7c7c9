"""torch.autograd.Functions over the factorization, the solve and their hand-written adjoints
(``mi_dpotrf_adjoint`` / ``mi_dpotrs_adjoint`` in libmi_llt.so).

No reference counterpart: Eigen's LLT (Eigen/src/Cholesky/LLT.h) has no derivatives.  The
gradient convention is that of ``torch.linalg.cholesky``: the gradient with respect to the
symmetric input is returned as a full symmetric matrix.

* :class:`Cholesky`      A -> L (lower, upper triangle zero), backward = mi_dpotrf_adjoint
* :class:`CholeskySolve` (B, L) -> X = (L L^T)^{-1} B, backward = mi_dpotrs_adjoint
* :func:`gp_log_marginal` the Gaussian-process log marginal likelihood
  -1/2 y^T K^{-1} y - sum log L_ii - n/2 log(2 pi), whose gradient is 1/2 (a a^T - K^{-1}),
  a = K^{-1} y.
All tensors are float64 on the GPU; inputs are copied to column-major storage.
"""
from __future__ import annotations

import math

import torch

from . import _VP, _ld, colmajor, default_handle, lib, potrf_, potrs_


def _cm(t: torch.Tensor) -> torch.Tensor:
    return colmajor(t.to(torch.float64))


class Cholesky(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a: torch.Tensor) -> torch.Tensor:
        l = _cm(a)
        info = potrf_(l)
        if info != 0:
            raise torch.linalg.LinAlgError(f"Cholesky: the leading minor of order {info} is not positive definite")
        l = torch.tril(l)
        l = colmajor(l)
        ctx.save_for_backward(l)
        return l

    @staticmethod
    def backward(ctx, lbar: torch.Tensor) -> torch.Tensor:
        (l,) = ctx.saved_tensors
        g = _cm(torch.tril(lbar))
        h = default_handle()
        n = int(l.shape[0])
        h.check(lib().mi_dpotrf_adjoint(h.ptr, b"L", n, _VP(l.data_ptr()), _ld(l), _VP(g.data_ptr()), _ld(g)),
                "mi_dpotrf_adjoint")
        return g


class CholeskySolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b: torch.Tensor, l: torch.Tensor) -> torch.Tensor:
        vec = b.dim() == 1
        x = _cm(b.reshape(b.shape[0], -1))
        lc = _cm(l)
        potrs_(lc, x)
        ctx.save_for_backward(lc, x)
        ctx.vec = vec
        return x.reshape(-1) if vec else x

    @staticmethod
    def backward(ctx, xbar: torch.Tensor):
        lc, x = ctx.saved_tensors
        n, r = int(x.shape[0]), int(x.shape[1])
        bbar = _cm(xbar.reshape(n, r))
        lbar = torch.zeros_like(lc)
        lbar = colmajor(lbar)
        h = default_handle()
        h.check(lib().mi_dpotrs_adjoint(h.ptr, b"L", n, r, _VP(lc.data_ptr()), _ld(lc), _VP(x.data_ptr()), _ld(x),
                                        _VP(bbar.data_ptr()), _ld(bbar), _VP(lbar.data_ptr()), _ld(lbar)),
                "mi_dpotrs_adjoint")
        bb = bbar.reshape(-1) if ctx.vec else bbar
        return bb, lbar


def cholesky(a: torch.Tensor) -> torch.Tensor:
    return Cholesky.apply(a)


def cholesky_solve(b: torch.Tensor, l: torch.Tensor) -> torch.Tensor:
    return CholeskySolve.apply(b, l)


def gp_log_marginal(k: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """log N(y | 0, K) through the library's factorization and solve (differentiable in K)."""
    n = k.shape[0]
    l = cholesky(k)
    alpha = cholesky_solve(y, l)
    return -0.5 * torch.dot(y, alpha) - torch.log(torch.diagonal(l)).sum() - 0.5 * n * math.log(2.0 * math.pi)
