"""``mi_llt.LLT`` -- object mirror of ``Eigen::LLT<MatrixXd, UpLo>`` on the MI355X.

Follows Eigen/src/Cholesky/LLT.h:
  * ``LLT(matrix)`` / ``compute(matrix)`` (:93-110, :430-460): copies the matrix into the
    object's own storage (``m_matrix``), records the L1 norm of the symmetric view
    (``m_l1_norm``, :442-453), factors in place, sets ``info()``;
  * ``solve(b)`` (:483-498) returns A^{-1} b, ``solve_in_place(b)`` (:510-522);
  * ``matrixL()`` / ``matrixU()`` (:128-139, :401-419): the triangular factor, materialised
    with zeros in the other triangle as assigning Eigen's TriangularView to a dense matrix
    does (TriangularMatrix.h:118);
  * ``matrixLLT()`` (:178-183): the whole storage, whose other triangle still holds the
    input (only one triangle is written, :225-229);
  * ``info()`` (:191-195) with Eigen's ComputationInfo values (Core/util/Constants.h:430-440);
  * ``rcond()`` (:166-171): the reciprocal condition estimate from ``m_l1_norm`` and the
    factor (``mi_dpocon``);
  * ``rankUpdate(v, sigma)`` (:467-480): the factor of A + sigma v v^T in place
    (``mi_dllt_rank_update``), ``info()`` NumericalIssue when that is not positive definite.

``dtype=torch.float32`` mirrors ``LLT<MatrixXf>`` (``mi_spotrf`` / ``mi_spotrs``); ``rcond`` and
``rankUpdate`` are FP64 only.

All storage lives in device memory (torch tensors); the factor and solve run in the
library's HIP kernels.  Misuse (use before compute) raises, as ``eigen_assert`` does.
"""
from __future__ import annotations

import enum

import torch

from . import (_uplo, colmajor, default_handle, pocon, posv_async_, potrf_, potrs_, rank_update_,  # noqa: F401
               spotrf_, spotrs_, sy_l1norm)


class ComputationInfo(enum.IntEnum):
    Success = 0
    NumericalIssue = 1
    NoConvergence = 2
    InvalidInput = 3


class LLT:
    """Cholesky decomposition A = L L^T of a symmetric positive-definite float64 matrix."""

    def __init__(self, matrix: torch.Tensor | int | None = None, uplo: str = "L", device="cuda",
                 dtype: torch.dtype = torch.float64):
        _uplo(uplo)
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("LLT: float64 (MatrixXd) or float32 (MatrixXf)")
        self.uplo = uplo.upper()
        self.dtype = dtype
        self._m = None
        self._l1 = None
        self._info = None
        self._device = device
        if isinstance(matrix, int):  # LLT(Index size): preallocation (LLT.h:93-95)
            self._m = torch.empty_strided((matrix, matrix), (1, max(matrix, 1)), dtype=dtype, device=device)
        elif matrix is not None:
            self.compute(matrix)

    # -- LLT.h:430-460
    def compute(self, a: torch.Tensor) -> "LLT":
        if a.dim() != 2 or a.shape[0] != a.shape[1]:
            raise ValueError("LLT::compute: square matrix required")
        n = int(a.shape[0])
        dev = a.device if a.is_cuda else torch.device(self._device)
        if self._m is None or tuple(self._m.shape) != (n, n) or self._m.device != dev:
            self._m = torch.empty_strided((n, n), (1, max(n, 1)), dtype=self.dtype, device=dev)
        self._m.copy_(a)
        if self.dtype == torch.float64:
            self._l1 = sy_l1norm(self._m, self.uplo)
            info = potrf_(self._m, self.uplo)
        else:
            self._l1 = sy_l1norm(colmajor(self._m.double()), self.uplo).float()
            info = spotrf_(self._m, self.uplo)
        self._info = ComputationInfo.Success if info == 0 else ComputationInfo.NumericalIssue
        return self

    def _check(self):
        if self._info is None:
            raise RuntimeError("LLT is not initialized.")

    # -- LLT.h:191-195
    def info(self) -> ComputationInfo:
        self._check()
        return self._info

    def rows(self) -> int:
        return 0 if self._m is None else int(self._m.shape[0])

    cols = rows

    # -- LLT.h:178-183
    def matrixLLT(self) -> torch.Tensor:
        self._check()
        return self._m

    # -- LLT.h:128-139, 401-419
    def matrixL(self) -> torch.Tensor:
        self._check()
        return torch.tril(self._m) if self.uplo == "L" else torch.tril(self._m.t())

    def matrixU(self) -> torch.Tensor:
        self._check()
        return self.matrixL().t()

    def l1_norm(self) -> float:
        self._check()
        return float(self._l1)

    def _potrs(self, x: torch.Tensor):
        (potrs_ if self.dtype == torch.float64 else spotrs_)(self._m, x, self.uplo)

    # -- LLT.h:483-498
    def solve(self, b: torch.Tensor) -> torch.Tensor:
        self._check()
        vec = b.dim() == 1
        x = colmajor(b.to(device=self._m.device, dtype=self.dtype))
        self._potrs(x)
        return x.reshape(-1) if vec else x

    # -- LLT.h:510-522
    def solve_in_place(self, b: torch.Tensor) -> torch.Tensor:
        self._check()
        if b.dim() != 2 or b.stride(0) != 1 or b.dtype != self.dtype or b.device != self._m.device:
            raise ValueError("solve_in_place needs a column-major tensor of the factor's dtype and device")
        self._potrs(b)
        return b

    # -- LLT.h:166-171
    def rcond(self) -> float:
        self._check()
        if self._info != ComputationInfo.Success:
            raise RuntimeError("LLT failed because matrix appears to be negative")
        if self.dtype != torch.float64:
            raise NotImplementedError("rcond: float64 factors only")
        return pocon(self._m, float(self._l1), self.uplo)

    # -- LLT.h:467-480
    def rankUpdate(self, v: torch.Tensor, sigma: float = 1.0) -> "LLT":
        if self._m is None or self._info is None:
            raise RuntimeError("LLT is not initialized.")
        if self.dtype != torch.float64:
            raise NotImplementedError("rankUpdate: float64 factors only")
        if v.numel() != self.rows():
            raise ValueError("rankUpdate: vector of size rows() required")
        vv = v.reshape(-1).to(device=self._m.device, dtype=torch.float64).contiguous()
        info = rank_update_(self._m, vv, float(sigma), self.uplo)
        self._info = ComputationInfo.Success if info == 0 else ComputationInfo.NumericalIssue
        return self

    rank_update = rankUpdate

    def adjoint(self) -> "LLT":
        return self
