"""Python binding of the MI355X dense Cholesky (LLT) library.

A thin ctypes layer over the C ABI in ``include/mi_llt.h`` (library
``eigen-git-mirror_amd/lib/libmi_llt.so``).  PyTorch provides device memory and the
stream; every computation runs in the HIP kernels of the library.  There is no CPU or
PyTorch fallback: if the library is missing, :func:`lib` raises.

Matrices are column-major, as Eigen's ``MatrixXd``: a torch tensor ``M`` of shape
``(n, m)`` with ``M.stride(0) == 1`` holds element (i, j) at ``i + j * ld``,
``ld = M.stride(1)``.  :func:`colmajor` makes such a tensor.

The object-level mirror of ``Eigen::LLT<MatrixXd>`` (Eigen/src/Cholesky/LLT.h:66-233) is
:class:`LLT` in ``mi_llt.llt``.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  -- must load torch's HIP runtime before the library binds to one

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmi_llt.so")

_lib = None

_I64 = ctypes.c_int64
_VP = ctypes.c_void_p
_INT = ctypes.c_int


class MiLLTError(RuntimeError):
    """A status code other than 0 from the C ABI (``code`` holds it)."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


E_DEVICE = -1005  # MI_LLT_E_DEVICE: the device reported a failure of queued work
STATUS_SOLVE_TIMEOUT, STATUS_PANEL_TIMEOUT, STATUS_FACTOR_TICKET = 1, 2, 4


def lib() -> ctypes.CDLL:
    """Load ``libmi_llt.so`` (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MiLLTError(f"{LIB_PATH} not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.mi_llt_create.argtypes = [ctypes.POINTER(_VP)]
        L.mi_llt_destroy.argtypes = [_VP]
        L.mi_llt_set_stream.argtypes = [_VP, _VP]
        L.mi_llt_last_error.argtypes = [_VP]
        L.mi_llt_last_error.restype = ctypes.c_char_p
        L.mi_llt_workspace_bytes.argtypes = [_VP]
        L.mi_llt_workspace_bytes.restype = _I64
        L.mi_dpotrf.argtypes = [_VP, ctypes.c_char, _I64, _VP, _I64, ctypes.POINTER(_INT)]
        L.mi_dpotrf_async.argtypes = [_VP, ctypes.c_char, _I64, _VP, _I64, _VP]
        L.mi_dpotrs.argtypes = [_VP, ctypes.c_char, _I64, _I64, _VP, _I64, _VP, _I64]
        L.mi_dposv_async.argtypes = [_VP, ctypes.c_char, _I64, _I64, _VP, _I64, _VP, _I64, _VP]
        L.mi_dsy_l1norm.argtypes = [_VP, ctypes.c_char, _I64, _VP, _I64, _VP]
        L.mi_llt_trace.argtypes = [_VP, _INT]
        L.mi_llt_trace_read.argtypes = [_VP, _INT, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I64)]
        L.mi_llt_trace_reset.argtypes = [_VP]
        L.mi_llt_status.argtypes = [_VP, ctypes.POINTER(_INT)]
        L.mi_llt_inject_status.argtypes = [_VP, _INT]
        L.mi_dposv_vbatched.argtypes = [_VP, ctypes.c_char, _I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP]
        L.mi_sposv_vbatched.argtypes = [_VP, ctypes.c_char, _I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP]
        L.mi_llt_set_vbatched_config.argtypes = [_VP, _INT, _INT, _INT]
        L.mi_llt_get_vbatched_config.argtypes = [_VP, ctypes.POINTER(_INT), ctypes.POINTER(_INT),
                                                 ctypes.POINTER(_INT)]
        L.mi_dpotrf_adjoint.argtypes = [_VP, ctypes.c_char, _I64, _VP, _I64, _VP, _I64]
        L.mi_dpotrs_adjoint.argtypes = [_VP, ctypes.c_char, _I64, _I64, _VP, _I64, _VP, _I64, _VP, _I64, _VP, _I64]
        L.mi_dposv_mixed.argtypes = [_VP, ctypes.c_char, _I64, _I64, _VP, _I64, _VP, _I64, _VP, _I64, _INT,
                                     ctypes.c_double, _INT, ctypes.POINTER(_INT), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(_INT)]
        L.mi_spotrf.argtypes = [_VP, ctypes.c_char, _I64, _VP, _I64, ctypes.POINTER(_INT)]
        L.mi_spotrf_async.argtypes = [_VP, ctypes.c_char, _I64, _VP, _I64, _VP]
        L.mi_spotrs.argtypes = [_VP, ctypes.c_char, _I64, _I64, _VP, _I64, _VP, _I64]
        L.mi_dtrsm.argtypes = [_VP, ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_char, _I64, _I64,
                               ctypes.c_double, _VP, _I64, _VP, _I64]
        L.mi_dsyrk.argtypes = [_VP, ctypes.c_char, ctypes.c_char, _I64, _I64, ctypes.c_double, _VP, _I64,
                               ctypes.c_double, _VP, _I64]
        L.mi_dpocon.argtypes = [_VP, ctypes.c_char, _I64, _VP, _I64, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.mi_dllt_rank_update.argtypes = [_VP, ctypes.c_char, _I64, _VP, _I64, _VP, _I64, ctypes.c_double, _VP]
        L.mi_dpotrf_batched_async.argtypes = [_VP, ctypes.c_char, _I64, _I64, _VP, _I64, _I64, _VP]
        L.mi_dpotrs_batched.argtypes = [_VP, ctypes.c_char, _I64, _I64, _I64, _VP, _I64, _I64, _VP, _I64, _I64]
        L.mi_dpotrf_adjoint_batched.argtypes = [_VP, ctypes.c_char, _I64, _I64, _VP, _I64, _I64, _VP, _I64, _I64]
        L.mi_dpotrs_adjoint_batched.argtypes = [_VP, ctypes.c_char, _I64, _I64, _I64, _VP, _I64, _I64, _VP, _I64,
                                                _I64, _VP, _I64, _I64, _VP, _I64, _I64]
        L.mi_dposv_noalloc.argtypes = [ctypes.c_char, _I64, _I64, _VP, _I64, _VP, _I64, _VP, _VP]
        L.mi_dposv_batched_async.argtypes = [_VP, ctypes.c_char, _I64, _I64, _I64, _VP, _I64, _I64, _VP, _I64, _I64,
                                             _VP]
        L.mi_dposv_adjoint_batched.argtypes = [_VP, ctypes.c_char, _I64, _I64, _I64, _VP, _I64, _I64, _VP, _I64,
                                               _I64, _VP, _I64, _I64, _VP, _I64, _I64]
        L.LAPACKE_spotrf.argtypes = [_INT, ctypes.c_char, _INT, _VP, _INT]
        L.LAPACKE_spotrs.argtypes = [_INT, ctypes.c_char, _INT, _INT, _VP, _INT, _VP, _INT]
        _PI = ctypes.POINTER(_INT)
        _PD = ctypes.POINTER(ctypes.c_double)
        L.spotrf_.argtypes = [ctypes.c_char_p, _PI, _VP, _PI, _PI]
        L.spotrs_.argtypes = [ctypes.c_char_p, _PI, _PI, _VP, _PI, _VP, _PI, _PI]
        L.dtrsm_.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _PI, _PI, _PD,
                             _VP, _PI, _VP, _PI]
        L.dsyrk_.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _PI, _PI, _PD, _VP, _PI, _PD, _VP, _PI]
        L.spotrf_.restype = None
        L.spotrs_.restype = None
        L.dtrsm_.restype = _INT
        L.dsyrk_.restype = _INT
        L.LAPACKE_dpotrf.argtypes = [_INT, ctypes.c_char, _INT, _VP, _INT]
        L.LAPACKE_dpotrs.argtypes = [_INT, ctypes.c_char, _INT, _INT, _VP, _INT, _VP, _INT]
        L.dpotrf_.argtypes = [ctypes.c_char_p, ctypes.POINTER(_INT), _VP, ctypes.POINTER(_INT), ctypes.POINTER(_INT)]
        L.dpotrs_.argtypes = [ctypes.c_char_p, ctypes.POINTER(_INT), ctypes.POINTER(_INT), _VP,
                              ctypes.POINTER(_INT), _VP, ctypes.POINTER(_INT), ctypes.POINTER(_INT)]
        for name in ("mi_llt_create", "mi_llt_destroy", "mi_llt_set_stream", "mi_dpotrf", "mi_dpotrf_async",
                     "mi_dpotrs", "mi_dposv_async", "mi_dsy_l1norm", "LAPACKE_dpotrf", "LAPACKE_dpotrs",
                     "mi_llt_trace", "mi_llt_trace_read", "mi_llt_trace_reset", "mi_llt_status", "mi_llt_inject_status",
                     "mi_dposv_mixed", "mi_dposv_vbatched", "mi_sposv_vbatched", "mi_llt_set_vbatched_config",
                     "mi_llt_get_vbatched_config", "mi_dpotrf_adjoint", "mi_dpotrs_adjoint", "mi_spotrf",
                     "mi_spotrf_async", "mi_spotrs", "mi_dtrsm", "mi_dsyrk", "mi_dpocon", "mi_dllt_rank_update",
                     "LAPACKE_spotrf", "LAPACKE_spotrs", "mi_dpotrf_batched_async", "mi_dpotrs_batched",
                     "mi_dpotrf_adjoint_batched", "mi_dpotrs_adjoint_batched", "mi_dposv_noalloc",
                     "mi_dposv_batched_async", "mi_dposv_adjoint_batched"):
            getattr(L, name).restype = _INT
        L.dpotrf_.restype = None
        L.dpotrs_.restype = None
        _lib = L
    return _lib


def colmajor(t: torch.Tensor) -> torch.Tensor:
    """Column-major copy (stride(0) == 1) of a 2-D tensor on its device."""
    if t.dim() == 1:
        t = t.reshape(-1, 1)
    out = torch.empty_strided(tuple(t.shape), (1, max(t.shape[0], 1)), dtype=t.dtype, device=t.device)
    out.copy_(t)
    return out


def colmajor_batched(t: torch.Tensor) -> torch.Tensor:
    """Copy of a (batch, rows, cols) tensor with every matrix column-major and packed:
    strides (rows * cols, 1, rows), the layout of the batched entry points."""
    if t.dim() == 2:
        t = t.unsqueeze(-1)
    b, r, c = (int(x) for x in t.shape)
    out = torch.empty_strided((b, r, c), (r * c, 1, max(r, 1)), dtype=t.dtype, device=t.device)
    out.copy_(t)
    return out


def _batched_ld(t: torch.Tensor):
    """(ld, stride) of a batched column-major tensor (see colmajor_batched)."""
    if t.dim() != 3 or (t.shape[1] > 1 and t.stride(1) != 1):
        raise ValueError("expected a (batch, rows, cols) tensor of column-major matrices; use mi_llt.colmajor_batched()")
    ld = max(int(t.stride(2)), int(t.shape[1]), 1)
    stride = int(t.stride(0)) if t.shape[0] > 1 else ld * int(t.shape[2])
    return ld, stride


def _ld(t: torch.Tensor) -> int:
    if t.dim() != 2 or (t.shape[0] > 1 and t.stride(0) != 1):
        raise ValueError("expected a column-major 2-D tensor (stride(0) == 1); use mi_llt.colmajor()")
    return max(int(t.stride(1)), int(t.shape[0]), 1)


class Handle:
    """Owns a ``mi_llt_handle`` bound to torch's current stream at call time."""

    def __init__(self):
        h = _VP()
        rc = lib().mi_llt_create(ctypes.byref(h))
        if rc != 0:
            raise MiLLTError(f"mi_llt_create failed ({rc})")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.mi_llt_destroy(self._h)
            self._h = None

    @property
    def ptr(self):
        return self._h

    def bind_stream(self, stream: torch.cuda.Stream | None = None):
        s = stream if stream is not None else torch.cuda.current_stream()
        lib().mi_llt_set_stream(self._h, _VP(s.cuda_stream))
        return self

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = lib().mi_llt_last_error(self._h).decode(errors="replace")
            raise MiLLTError(f"{what} returned {rc}: {msg}", rc)

    def workspace_bytes(self) -> int:
        return int(lib().mi_llt_workspace_bytes(self._h))

    def status(self) -> int:
        """Waits for the stream; device status bits since the last call (0 = all good)."""
        v = ctypes.c_int(0)
        self.check(lib().mi_llt_status(self._h, ctypes.byref(v)), "mi_llt_status")
        return int(v.value)

    def inject_status(self, bits: int):
        """Sets device status bits on the handle's stream (tests of the failure path)."""
        self.check(lib().mi_llt_inject_status(self._h, int(bits)), "mi_llt_inject_status")

    TRACE_KINDS = {"factor_step": 0, "solve": 3, "trtri": 4, "residual": 5, "batched": 6}

    def trace(self, enable: bool = True):
        self.check(lib().mi_llt_trace(self._h, int(enable)), "mi_llt_trace")

    def trace_reset(self):
        self.check(lib().mi_llt_trace_reset(self._h), "mi_llt_trace_reset")

    def trace_read(self) -> dict:
        """{kind: (total_ms, algorithmic_flops, launches)} since the last reset."""
        out = {}
        for name, k in self.TRACE_KINDS.items():
            ms, fl, nl = ctypes.c_double(), ctypes.c_double(), _I64()
            self.check(lib().mi_llt_trace_read(self._h, k, ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(nl)),
                       "mi_llt_trace_read")
            out[name] = (ms.value, fl.value, int(nl.value))
        return out


_handles: dict[int, Handle] = {}


def default_handle() -> Handle:
    """The handle of torch's current stream (one per stream: a handle's workspace, tickets and
    flags serve one stream's calls in order, so work on several streams needs several)."""
    s = torch.cuda.current_stream()
    h = _handles.get(s.cuda_stream)
    if h is None:
        h = _handles[s.cuda_stream] = Handle()
    return h.bind_stream(s)


def _uplo(u: str) -> bytes:
    if u not in ("L", "U", "l", "u"):
        raise ValueError("uplo must be 'L' or 'U'")
    return u.encode()


def potrf_(a: torch.Tensor, uplo: str = "L", handle: Handle | None = None) -> int:
    """In-place Cholesky of the ``uplo`` triangle of column-major ``a`` (float64, on the GPU).

    Returns LAPACK ``info`` (0, or k+1 for a failing pivot at column k).  Synchronizes.
    """
    h = handle.bind_stream() if handle else default_handle()
    n = int(a.shape[0])
    info = ctypes.c_int(0)
    h.check(lib().mi_dpotrf(h.ptr, _uplo(uplo), n, _VP(a.data_ptr()), _ld(a), ctypes.byref(info)), "mi_dpotrf")
    return int(info.value)


def potrf_async_(a: torch.Tensor, d_info: torch.Tensor, uplo: str = "L", handle: Handle | None = None):
    """Asynchronous :func:`potrf_`; ``d_info`` is a 1-element int32 device tensor."""
    h = handle.bind_stream() if handle else default_handle()
    n = int(a.shape[0])
    h.check(lib().mi_dpotrf_async(h.ptr, _uplo(uplo), n, _VP(a.data_ptr()), _ld(a), _VP(d_info.data_ptr())),
            "mi_dpotrf_async")


def potrs_(l: torch.Tensor, b: torch.Tensor, uplo: str = "L", handle: Handle | None = None):
    """In-place solve ``b <- A^{-1} b`` with the factor in ``l`` (as left by :func:`potrf_`)."""
    h = handle.bind_stream() if handle else default_handle()
    n = int(l.shape[0])
    nrhs = int(b.shape[1])
    h.check(lib().mi_dpotrs(h.ptr, _uplo(uplo), n, nrhs, _VP(l.data_ptr()), _ld(l), _VP(b.data_ptr()), _ld(b)),
            "mi_dpotrs")


def posv_async_(a: torch.Tensor, b: torch.Tensor, d_info: torch.Tensor, uplo: str = "L",
                handle: Handle | None = None):
    """Factor ``a`` in place and solve ``b`` in place, asynchronously (``d_info`` int32 device)."""
    h = handle.bind_stream() if handle else default_handle()
    n = int(a.shape[0])
    nrhs = int(b.shape[1])
    h.check(lib().mi_dposv_async(h.ptr, _uplo(uplo), n, nrhs, _VP(a.data_ptr()), _ld(a), _VP(b.data_ptr()),
                                 _ld(b), _VP(d_info.data_ptr())), "mi_dposv_async")


def spotrf_(a: torch.Tensor, uplo: str = "L", handle: Handle | None = None) -> int:
    """In-place Cholesky of a column-major float32 ``a`` (``mi_spotrf``); returns LAPACK info."""
    h = handle.bind_stream() if handle else default_handle()
    if a.dtype != torch.float32:
        raise ValueError("spotrf_ needs float32")
    info = ctypes.c_int(0)
    h.check(lib().mi_spotrf(h.ptr, _uplo(uplo), int(a.shape[0]), _VP(a.data_ptr()), _ld(a), ctypes.byref(info)),
            "mi_spotrf")
    return int(info.value)


def spotrs_(l: torch.Tensor, b: torch.Tensor, uplo: str = "L", handle: Handle | None = None):
    """In-place float32 solve ``b <- A^{-1} b`` with the factor from :func:`spotrf_`."""
    h = handle.bind_stream() if handle else default_handle()
    if l.dtype != torch.float32 or b.dtype != torch.float32:
        raise ValueError("spotrs_ needs float32")
    h.check(lib().mi_spotrs(h.ptr, _uplo(uplo), int(l.shape[0]), int(b.shape[1]), _VP(l.data_ptr()), _ld(l),
                            _VP(b.data_ptr()), _ld(b)), "mi_spotrs")


def trsm_(a: torch.Tensor, b: torch.Tensor, side: str = "L", uplo: str = "L", trans: str = "N", diag: str = "N",
          alpha: float = 1.0, handle: Handle | None = None):
    """``b <- alpha op(a)^{-1} b`` (side 'L') or ``alpha b op(a)^{-1}`` (side 'R'), float64 (``mi_dtrsm``)."""
    h = handle.bind_stream() if handle else default_handle()
    m, n = int(b.shape[0]), int(b.shape[1])
    h.check(lib().mi_dtrsm(h.ptr, side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n, float(alpha),
                           _VP(a.data_ptr()), _ld(a), _VP(b.data_ptr()), _ld(b)), "mi_dtrsm")


def syrk_(a: torch.Tensor, c: torch.Tensor, uplo: str = "L", trans: str = "N", alpha: float = 1.0,
          beta: float = 1.0, handle: Handle | None = None):
    """``c <- alpha op(a) op(a)^T + beta c`` on the ``uplo`` triangle, float64 (``mi_dsyrk``)."""
    h = handle.bind_stream() if handle else default_handle()
    n = int(c.shape[0])
    k = int(a.shape[1]) if trans in ("N", "n") else int(a.shape[0])
    h.check(lib().mi_dsyrk(h.ptr, uplo.encode(), trans.encode(), n, k, float(alpha), _VP(a.data_ptr()), _ld(a),
                           float(beta), _VP(c.data_ptr()), _ld(c)), "mi_dsyrk")


def pocon(l: torch.Tensor, anorm: float, uplo: str = "L", handle: Handle | None = None) -> float:
    """Reciprocal condition estimate from the factor in ``l`` and ``anorm`` = ||A||_1 (``mi_dpocon``)."""
    h = handle.bind_stream() if handle else default_handle()
    out = ctypes.c_double(0.0)
    h.check(lib().mi_dpocon(h.ptr, _uplo(uplo), int(l.shape[0]), _VP(l.data_ptr()), _ld(l), float(anorm),
                            ctypes.byref(out)), "mi_dpocon")
    return float(out.value)


def rank_update_(l: torch.Tensor, v: torch.Tensor, sigma: float = 1.0, uplo: str = "L",
                 handle: Handle | None = None) -> int:
    """The factor of ``L L^T + sigma v v^T`` in place of ``l`` (``mi_dllt_rank_update``); returns
    LAPACK-style info (0, or j+1 when the result is not positive definite at column j)."""
    h = handle.bind_stream() if handle else default_handle()
    n = int(l.shape[0])
    if v.numel() != n or v.dtype != torch.float64 or v.device != l.device:
        raise ValueError("rank_update_ needs a float64 vector of n entries on the factor's device")
    inc = 1 if v.dim() == 1 else (int(v.stride(0)) if v.shape[0] == n else int(v.stride(1)))
    d_info = torch.zeros(1, dtype=torch.int32, device=l.device)
    h.check(lib().mi_dllt_rank_update(h.ptr, _uplo(uplo), n, _VP(l.data_ptr()), _ld(l), _VP(v.data_ptr()),
                                      max(inc, 1), float(sigma), _VP(d_info.data_ptr())), "mi_dllt_rank_update")
    return int(d_info.item())


def potrf_batched_async_(a: torch.Tensor, d_info: torch.Tensor, handle: Handle | None = None):
    """In-place Cholesky of every matrix of a (batch, n, n) float64 tensor (column-major per
    matrix, lower triangle; ``mi_dpotrf_batched_async``); ``d_info``: int32 device tensor of
    ``batch`` entries (LAPACK info per system)."""
    h = handle.bind_stream() if handle else default_handle()
    ld, st = _batched_ld(a)
    h.check(lib().mi_dpotrf_batched_async(h.ptr, b"L", int(a.shape[1]), int(a.shape[0]), _VP(a.data_ptr()), ld, st,
                                          _VP(d_info.data_ptr())), "mi_dpotrf_batched_async")


def posv_batched_async_(a: torch.Tensor, b: torch.Tensor, d_info: torch.Tensor, handle: Handle | None = None):
    """Factor every matrix of ``a`` (batch, n, n) in place and solve ``b`` (batch, n, r) in
    place (``mi_dposv_batched_async``); ``d_info``: int32 device tensor of ``batch`` entries."""
    h = handle.bind_stream() if handle else default_handle()
    lda, sa = _batched_ld(a)
    ldb, sb = _batched_ld(b)
    h.check(lib().mi_dposv_batched_async(h.ptr, b"L", int(a.shape[1]), int(b.shape[2]), int(a.shape[0]),
                                         _VP(a.data_ptr()), lda, sa, _VP(b.data_ptr()), ldb, sb,
                                         _VP(d_info.data_ptr())), "mi_dposv_batched_async")


def potrs_batched_(l: torch.Tensor, b: torch.Tensor, handle: Handle | None = None):
    """In-place ``b[z] <- A[z]^{-1} b[z]`` with the factors in ``l`` (``mi_dpotrs_batched``)."""
    h = handle.bind_stream() if handle else default_handle()
    ldl, sl = _batched_ld(l)
    ldb, sb = _batched_ld(b)
    h.check(lib().mi_dpotrs_batched(h.ptr, b"L", int(l.shape[1]), int(b.shape[2]), int(l.shape[0]),
                                    _VP(l.data_ptr()), ldl, sl, _VP(b.data_ptr()), ldb, sb), "mi_dpotrs_batched")


PREC = {"fp64": 0, "fp32": 1, "bf16": 2, "bf16x3": 3, "bf16x6": 4}


def posv_mixed(a: torch.Tensor, b: torch.Tensor, x: torch.Tensor | None = None, prec: str = "fp32",
               tol: float = 1e-14, maxit: int = 10, uplo: str = "L", handle: Handle | None = None):
    """Mixed-precision solve with FP64 iterative refinement (``mi_dposv_mixed``).

    ``a`` (float64, column-major, ``uplo`` triangle) is not modified; returns
    ``(x, iters, berr, info)``: the FP64 solution (written into ``x`` when given), the number
    of refinement corrections (-1 / -2: the FP64 factorization was used because the
    low-precision factor was not positive definite / refinement did not reach ``tol``), the
    final normwise backward error and LAPACK ``info``.  Synchronizes.
    """
    h = handle.bind_stream() if handle else default_handle()
    n = int(a.shape[0])
    nrhs = int(b.shape[1])
    if x is None:
        x = torch.empty_strided((n, nrhs), (1, max(n, 1)), dtype=torch.float64, device=a.device)
    it, be, inf = ctypes.c_int(0), ctypes.c_double(0.0), ctypes.c_int(0)
    h.check(lib().mi_dposv_mixed(h.ptr, _uplo(uplo), n, nrhs, _VP(a.data_ptr()), _ld(a), _VP(b.data_ptr()), _ld(b),
                                 _VP(x.data_ptr()), _ld(x), PREC[prec], float(tol), int(maxit), ctypes.byref(it),
                                 ctypes.byref(be), ctypes.byref(inf)), "mi_dposv_mixed")
    return x, int(it.value), float(be.value), int(inf.value)


def sy_l1norm(a: torch.Tensor, uplo: str = "L", handle: Handle | None = None) -> torch.Tensor:
    """``m_l1_norm`` of LLT::compute (LLT.h:442-453) as a 0-d float64 device tensor."""
    h = handle.bind_stream() if handle else default_handle()
    out = torch.empty((), dtype=torch.float64, device=a.device)
    h.check(lib().mi_dsy_l1norm(h.ptr, _uplo(uplo), int(a.shape[0]), _VP(a.data_ptr()), _ld(a),
                                _VP(out.data_ptr())), "mi_dsy_l1norm")
    return out


from .llt import LLT, ComputationInfo  # noqa: E402

__all__ = ["lib", "colmajor", "colmajor_batched", "potrf_batched_async_", "potrs_batched_", "posv_batched_async_",
           "Handle", "default_handle", "potrf_", "potrf_async_", "potrs_", "posv_async_",
           "posv_mixed", "PREC", "sy_l1norm", "spotrf_", "spotrs_", "trsm_", "syrk_", "pocon", "rank_update_",
           "LLT", "ComputationInfo", "MiLLTError", "LIB_PATH", "E_DEVICE"]
