"""torch.autograd.Functions over the factorization, the solve and their hand-written adjoints
(``mi_dpotrf_adjoint`` / ``mi_dpotrs_adjoint`` in libmi_llt.so).

No reference counterpart: Eigen's LLT (Eigen/src/Cholesky/LLT.h) has no derivatives.  The
gradient convention is that of ``torch.linalg.cholesky``: the gradient with respect to the
symmetric input is returned as a full symmetric matrix.

* :class:`Cholesky`      A -> L (lower, upper triangle zero), backward = mi_dpotrf_adjoint
* :class:`CholeskySolve` (B, L) -> X = (L L^T)^{-1} B, backward = mi_dpotrs_adjoint
* :class:`CholeskyBatched`, :class:`CholeskySolveBatched`: the same for a (batch, n, n) stack
  of systems of one order, through the batched entry points (every kernel launch covers the
  whole batch); :class:`GPLogMarginalBatched` / :func:`gp_log_marginal_batched` the fused
  log marginal likelihood of a batch (mi_dposv_batched_async forward, mi_dposv_adjoint_batched
  backward)
* :func:`gp_log_marginal` the Gaussian-process log marginal likelihood
  -1/2 y^T K^{-1} y - sum log L_ii - n/2 log(2 pi), whose gradient is 1/2 (a a^T - K^{-1}),
  a = K^{-1} y.
All tensors are float64 on the GPU; inputs are copied to column-major storage.
"""
from __future__ import annotations

import math

import torch

from . import (_VP, _batched_ld, _ld, colmajor, colmajor_batched, default_handle, lib, posv_batched_async_, potrf_,
               potrf_batched_async_, potrs_, potrs_batched_)


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


class CholeskyBatched(torch.autograd.Function):
    """A (batch, n, n) -> L (batch, n, n) lower; backward = mi_dpotrf_adjoint_batched."""

    @staticmethod
    def forward(ctx, a: torch.Tensor) -> torch.Tensor:
        l = colmajor_batched(a.to(torch.float64))
        info = torch.zeros(l.shape[0], dtype=torch.int32, device=l.device)
        potrf_batched_async_(l, info)
        bad = torch.nonzero(info).flatten()
        if bad.numel():
            z = int(bad[0])
            raise torch.linalg.LinAlgError(f"Cholesky: system {z}: the leading minor of order {int(info[z])} is not "
                                           "positive definite")
        l = colmajor_batched(torch.tril(l))
        ctx.save_for_backward(l)
        return l

    @staticmethod
    def backward(ctx, lbar: torch.Tensor) -> torch.Tensor:
        (l,) = ctx.saved_tensors
        g = colmajor_batched(torch.tril(lbar.to(torch.float64)))
        h = default_handle()
        ldl, sl = _batched_ld(l)
        ldg, sg = _batched_ld(g)
        h.check(lib().mi_dpotrf_adjoint_batched(h.ptr, b"L", int(l.shape[1]), int(l.shape[0]), _VP(l.data_ptr()), ldl,
                                                sl, _VP(g.data_ptr()), ldg, sg), "mi_dpotrf_adjoint_batched")
        return g


class CholeskySolveBatched(torch.autograd.Function):
    """(B (batch, n, r), L) -> X = (L L^T)^{-1} B per system; backward = mi_dpotrs_adjoint_batched."""

    @staticmethod
    def forward(ctx, b: torch.Tensor, l: torch.Tensor) -> torch.Tensor:
        vec = b.dim() == 2
        x = colmajor_batched(b.to(torch.float64))
        lc = colmajor_batched(l.to(torch.float64))
        potrs_batched_(lc, x)
        ctx.save_for_backward(lc, x)
        ctx.vec = vec
        return x.squeeze(-1) if vec else x

    @staticmethod
    def backward(ctx, xbar: torch.Tensor):
        lc, x = ctx.saved_tensors
        bbar = colmajor_batched(xbar.to(torch.float64).reshape(x.shape))
        lbar = colmajor_batched(torch.zeros_like(lc))
        h = default_handle()
        ldl, sl = _batched_ld(lc)
        ldx, sx = _batched_ld(x)
        ldb, sb = _batched_ld(bbar)
        ldlb, slb = _batched_ld(lbar)
        h.check(lib().mi_dpotrs_adjoint_batched(h.ptr, b"L", int(lc.shape[1]), int(x.shape[2]), int(lc.shape[0]),
                                                _VP(lc.data_ptr()), ldl, sl, _VP(x.data_ptr()), ldx, sx,
                                                _VP(bbar.data_ptr()), ldb, sb, _VP(lbar.data_ptr()), ldlb, slb),
                "mi_dpotrs_adjoint_batched")
        bb = bbar.squeeze(-1) if ctx.vec else bbar
        return bb, lbar


def cholesky_batched(a: torch.Tensor) -> torch.Tensor:
    return CholeskyBatched.apply(a)


def cholesky_solve_batched(b: torch.Tensor, l: torch.Tensor) -> torch.Tensor:
    return CholeskySolveBatched.apply(b, l)


class GPLogMarginalBatched(torch.autograd.Function):
    """(K (batch, n, n), y (batch, n)) -> log N(y_z | 0, K_z) per system, fused: the forward is
    one factor + solve per system (mi_dposv_batched_async: L and a = K^{-1} y), the backward one
    call of mi_dposv_adjoint_batched with the gradients the log-likelihood sends into a and L
    (abar = -g/2 y, Lbar = -g diag(1/L_ii)), which returns dL/dK = g/2 (a a^T - K^{-1})."""

    @staticmethod
    def forward(ctx, k: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        n = int(k.shape[-1])
        l = colmajor_batched(k.detach().to(torch.float64))  # the lower triangle becomes L
        a = colmajor_batched(y.detach().to(torch.float64))
        info = torch.zeros(l.shape[0], dtype=torch.int32, device=l.device)
        posv_batched_async_(l, a, info)
        bad = torch.nonzero(info).flatten()
        if bad.numel():
            z = int(bad[0])
            raise torch.linalg.LinAlgError(f"Cholesky: system {z}: the leading minor of order {int(info[z])} is not "
                                           "positive definite")
        d = torch.diagonal(l, dim1=-2, dim2=-1)
        ya = a.squeeze(-1)
        ctx.save_for_backward(l, a, y.detach().to(torch.float64))
        return -0.5 * (y.detach() * ya).sum(-1) - torch.log(d).sum(-1) - 0.5 * n * math.log(2.0 * math.pi)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        l, a, y = ctx.saved_tensors
        b, n = int(l.shape[0]), int(l.shape[1])
        g = g.to(torch.float64)
        xbar = colmajor_batched((-0.5 * g)[:, None] * y)
        lbar = torch.zeros_like(l)  # same (column-major) strides as l
        torch.diagonal(lbar, dim1=-2, dim2=-1).copy_(-g[:, None] / torch.diagonal(l, dim1=-2, dim2=-1))
        h = default_handle()
        ldl, sl = _batched_ld(l)
        ldx, sx = _batched_ld(a)
        ldb, sb = _batched_ld(xbar)
        ldlb, slb = _batched_ld(lbar)
        h.check(lib().mi_dposv_adjoint_batched(h.ptr, b"L", n, 1, b, _VP(l.data_ptr()), ldl, sl, _VP(a.data_ptr()), ldx,
                                               sx, _VP(xbar.data_ptr()), ldb, sb, _VP(lbar.data_ptr()), ldlb, slb),
                "mi_dposv_adjoint_batched")
        return lbar, None


def gp_log_marginal_batched(k: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """log N(y_z | 0, K_z) for a (batch, n, n) stack K and (batch, n) targets y: a (batch,)
    tensor, differentiable in K (one batched factor + solve forward, one batched adjoint
    backward)."""
    return GPLogMarginalBatched.apply(k, y)
