"""CPU oracle for the dense LLT path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import
this module, and only as the checker or the timed CPU baseline; the product path never
routes through it.

Two back ends, both loaded with ctypes:

* ``build/liboracle_llt.so`` -- the plain-C restatement in ``llt_oracle.c`` of
  Eigen/src/Cholesky/LLT.h:309-368 (blocked/unblocked), :442-453 (L1 norm) and :483-498
  (solve).  Pinned by tests/test_oracle.py against closed forms, the reference's 3x3
  documentation example, planted failing minors and the genuine Eigen build below.
* ``_ref/libeigen_ref_v{3,4}.so`` -- genuine Eigen 3.3.90 ``LLT<MatrixXd>`` compiled from
  the reference's headers by ``oracle/Makefile`` (absent when the reference tree was not
  available at build time; ``have_eigen()`` says which).
* ``_ref/libeigen_blas_d.so`` -- Eigen's own BLAS (blas/double.cpp), whose ``dtrsm_`` and
  ``dsyrk_`` pin the restatements of the level-3 routines (``have_eigen_blas()``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_I64 = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)

_oracle = None
_eigen = None


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def _colmajor(a: np.ndarray) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _load_oracle():
    global _oracle
    if _oracle is None:
        path = os.path.join(_HERE, "build", "liboracle_llt.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        lib.oracle_llt_unblocked.restype = _I64
        lib.oracle_llt_unblocked.argtypes = [_DP, _I64, _I64]
        lib.oracle_llt_blocked.restype = _I64
        lib.oracle_llt_blocked.argtypes = [_DP, _I64, _I64]
        lib.oracle_llt_blocked_upper.restype = _I64
        lib.oracle_llt_blocked_upper.argtypes = [_DP, _I64, _I64, _DP]
        lib.oracle_llt_block_size.restype = _I64
        lib.oracle_llt_block_size.argtypes = [_I64]
        lib.oracle_llt_solve.restype = None
        lib.oracle_llt_solve.argtypes = [_DP, _I64, _I64, _DP, _I64, _I64]
        lib.oracle_l1_norm_lower.restype = ctypes.c_double
        lib.oracle_l1_norm_lower.argtypes = [_DP, _I64, _I64]
        lib.oracle_factor_berr.restype = ctypes.c_double
        lib.oracle_factor_berr.argtypes = [_DP, _I64, _DP, _I64, _I64]
        lib.oracle_solve_berr.restype = ctypes.c_double
        lib.oracle_solve_berr.argtypes = [_DP, _I64, _I64, _DP, _I64, _DP, _I64, _I64]
        _FP = ctypes.POINTER(ctypes.c_float)
        lib.oracle_llt_blocked_f32.restype = _I64
        lib.oracle_llt_blocked_f32.argtypes = [_FP, _I64, _I64]
        lib.oracle_llt_rank_update.restype = _I64
        lib.oracle_llt_rank_update.argtypes = [_DP, _I64, _I64, _DP, ctypes.c_double, _DP]
        lib.oracle_llt_rcond.restype = ctypes.c_double
        lib.oracle_llt_rcond.argtypes = [_DP, _I64, _I64, ctypes.c_double, _DP]
        lib.oracle_trsm.restype = None
        lib.oracle_trsm.argtypes = [ctypes.c_char] * 4 + [_I64, _I64, ctypes.c_double, _DP, _I64, _DP, _I64]
        lib.oracle_syrk.restype = None
        lib.oracle_syrk.argtypes = [ctypes.c_char, ctypes.c_char, _I64, _I64, ctypes.c_double, _DP, _I64,
                                    ctypes.c_double, _DP, _I64]
        _oracle = lib
    return _oracle


def _host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = line.split()
                    return all(x in flags for x in ("avx512f", "avx512bw", "avx512dq", "avx512vl"))
    except OSError:
        pass
    return False


def eigen_path() -> str | None:
    name = "libeigen_ref_v4.so" if _host_has_avx512() else "libeigen_ref_v3.so"
    path = os.path.join(_HERE, "_ref", name)
    return path if os.path.exists(path) else None


def have_eigen() -> bool:
    return eigen_path() is not None


def _load_eigen():
    global _eigen
    if _eigen is None:
        path = eigen_path()
        if path is None:
            raise RuntimeError("genuine Eigen build (oracle/_ref) not present")
        lib = ctypes.CDLL(path)
        lib.eigen_ref_version.restype = ctypes.c_int
        lib.eigen_ref_llt_inplace.restype = _I64
        lib.eigen_ref_llt_inplace.argtypes = [_I64, _DP, _I64]
        lib.eigen_ref_llt_inplace_upper.restype = _I64
        lib.eigen_ref_llt_inplace_upper.argtypes = [_I64, _DP, _I64]
        lib.eigen_ref_llt_factor_solve.restype = ctypes.c_int
        lib.eigen_ref_llt_factor_solve.argtypes = [_I64, _DP, _I64, _DP, _I64, _DP, _I64, _I64,
                                                   _DP, _I64, _DP, _DP, _DP, _DP]
        lib.eigen_ref_llt_time.restype = ctypes.c_double
        lib.eigen_ref_llt_time.argtypes = [_I64, _DP, _I64, _DP, _I64, _I64, _DP, _I64]
        lib.eigen_ref_set_threads.restype = ctypes.c_int
        lib.eigen_ref_set_threads.argtypes = [ctypes.c_int]
        lib.eigen_ref_llt_inplace_f32.restype = _I64
        lib.eigen_ref_llt_inplace_f32.argtypes = [_I64, ctypes.POINTER(ctypes.c_float), _I64]
        lib.eigen_ref_llt_rank_update.restype = _I64
        lib.eigen_ref_llt_rank_update.argtypes = [_I64, _DP, _I64, _DP, ctypes.c_double]
        lib.eigen_ref_llt_rcond.restype = ctypes.c_double
        lib.eigen_ref_llt_rcond.argtypes = [_I64, _DP, _I64]
        lib.eigen_ref_gp_grad.restype = ctypes.c_int
        lib.eigen_ref_gp_grad.argtypes = [_I64, _DP, _I64, _DP, _DP, _I64]
        _eigen = lib
    return _eigen


_eigen_blas = None


def have_eigen_blas() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libeigen_blas_d.so"))


def _load_eigen_blas():
    global _eigen_blas
    if _eigen_blas is None:
        path = os.path.join(_HERE, "_ref", "libeigen_blas_d.so")
        if not os.path.exists(path):
            raise RuntimeError("Eigen's BLAS build (oracle/_ref) not present")
        lib = ctypes.CDLL(path)  # RTLD_LOCAL: its dtrsm_/dsyrk_ are reached through this handle only
        _PI = ctypes.POINTER(ctypes.c_int)
        _CP = ctypes.c_char_p
        lib.dtrsm_.argtypes = [_CP, _CP, _CP, _CP, _PI, _PI, _DP, _DP, _PI, _DP, _PI]
        lib.dtrsm_.restype = ctypes.c_int
        lib.dsyrk_.argtypes = [_CP, _CP, _PI, _PI, _DP, _DP, _PI, _DP, _DP, _PI]
        lib.dsyrk_.restype = ctypes.c_int
        _eigen_blas = lib
    return _eigen_blas


# ----------------------------------------------------------------------------- restatement
def block_size(n: int) -> int:
    """LLT.h:346-348 block size rule."""
    return int(_load_oracle().oracle_llt_block_size(n))


def llt_unblocked(a: np.ndarray):
    """LLT.h:312-336.  Returns (factored copy, index) with index -1 on success."""
    m = _colmajor(a).copy(order="F")
    n = m.shape[0]
    r = _load_oracle().oracle_llt_unblocked(_dp(m), n, max(n, 1))
    return m, int(r)


def llt_blocked(a: np.ndarray, upper: bool = False):
    """LLT.h:338-368 (Lower) or :377-399 (Upper).  Returns (factored copy, index)."""
    m = _colmajor(a).copy(order="F")
    n = m.shape[0]
    lib = _load_oracle()
    if upper:
        work = np.zeros((max(n, 1), max(n, 1)), order="F")
        r = lib.oracle_llt_blocked_upper(_dp(m), n, max(n, 1), _dp(work))
    else:
        r = lib.oracle_llt_blocked(_dp(m), n, max(n, 1))
    return m, int(r)


def llt_solve(l: np.ndarray, b: np.ndarray) -> np.ndarray:
    """LLT.h:483-498: X = L^-T L^-1 B using the lower triangle of ``l``."""
    lf = _colmajor(l)
    x = _colmajor(b).copy(order="F")
    vec = x.ndim == 1
    if vec:
        x = x.reshape(-1, 1, order="F")
    n, r = x.shape
    _load_oracle().oracle_llt_solve(_dp(lf), n, max(n, 1), _dp(x), r, max(n, 1))
    return x.reshape(-1) if vec else x


def l1_norm_lower(a: np.ndarray) -> float:
    """LLT.h:442-453 m_l1_norm."""
    af = _colmajor(a)
    n = af.shape[0]
    return float(_load_oracle().oracle_l1_norm_lower(_dp(af), n, max(n, 1)))


def factor_berr(a: np.ndarray, l: np.ndarray) -> float:
    """||A - L L^T||_F / ||A||_F (lower triangles of a and l), long double accumulation."""
    af, lf = _colmajor(a), _colmajor(l)
    n = af.shape[0]
    return float(_load_oracle().oracle_factor_berr(_dp(af), n, _dp(lf), n, n))


def solve_berr(a: np.ndarray, x: np.ndarray, b: np.ndarray) -> float:
    """max_c ||A x_c - b_c||_2 / (||A||_F ||x_c||_2), long double accumulation."""
    af = _colmajor(a)
    xf = _colmajor(x).reshape(af.shape[0], -1, order="F")
    bf = _colmajor(b).reshape(af.shape[0], -1, order="F")
    n, r = xf.shape
    return float(_load_oracle().oracle_solve_berr(_dp(af), n, n, _dp(xf), n, _dp(bf), n, r))


# ----------------------------------------------------------------------------- genuine Eigen
def eigen_llt_inplace(a: np.ndarray, upper: bool = False):
    """Eigen's own llt_inplace<double,Lower|Upper>::blocked on a copy: (matrix, index)."""
    m = _colmajor(a).copy(order="F")
    n = m.shape[0]
    lib = _load_eigen()
    fn = lib.eigen_ref_llt_inplace_upper if upper else lib.eigen_ref_llt_inplace
    r = fn(n, _dp(m), max(n, 1))
    return m, int(r)


def eigen_llt(a: np.ndarray, b: np.ndarray):
    """Eigen LLT<MatrixXd>(A).solve(B): returns dict with L storage, X, info, times, rcond."""
    af = _colmajor(a)
    bf = _colmajor(b)
    vec = bf.ndim == 1
    if vec:
        bf = bf.reshape(-1, 1, order="F")
    n = af.shape[0]
    r = bf.shape[1]
    l = np.zeros((n, n), order="F")
    x = np.zeros((n, r), order="F")
    tf = ctypes.c_double()
    ts = ctypes.c_double()
    rc = ctypes.c_double()
    info = _load_eigen().eigen_ref_llt_factor_solve(
        n, _dp(af), max(n, 1), _dp(l), max(n, 1), _dp(bf), r, max(n, 1), _dp(x), max(n, 1),
        ctypes.byref(tf), ctypes.byref(ts), None, ctypes.byref(rc))
    return {"L": l, "X": x.reshape(-1) if vec else x, "info": int(info),
            "t_factor": tf.value, "t_solve": ts.value, "rcond": rc.value}


def eigen_set_threads(n: int) -> int:
    """Eigen::setNbThreads(n); returns Eigen::nbThreads()."""
    return int(_load_eigen().eigen_ref_set_threads(int(n)))


def eigen_llt_time(a: np.ndarray, b: np.ndarray) -> float:
    """Wall seconds of Eigen compute()+solve() (copies outside the timed region)."""
    af = _colmajor(a)
    bf = _colmajor(b)
    n, r = bf.shape
    return float(_load_eigen().eigen_ref_llt_time(n, _dp(af), n, _dp(bf), r, n, None, n))


def eigen_llt_inplace_f32(a: np.ndarray):
    """Eigen's llt_inplace<float,Lower>::blocked on a float32 copy: (matrix, index)."""
    m = np.asfortranarray(np.asarray(a, dtype=np.float32)).copy(order="F")
    n = m.shape[0]
    r = _load_eigen().eigen_ref_llt_inplace_f32(n, m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), max(n, 1))
    return m, int(r)


def eigen_rank_update(l: np.ndarray, v: np.ndarray, sigma: float):
    """Eigen's llt_inplace<double,Lower>::rankUpdate on a copy of the lower factor: (L, index)."""
    m = _colmajor(l).copy(order="F")
    vv = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    n = m.shape[0]
    r = _load_eigen().eigen_ref_llt_rank_update(n, _dp(m), max(n, 1), _dp(vv), float(sigma))
    return m, int(r)


def eigen_rcond(a: np.ndarray) -> float:
    """LLT<MatrixXd>(A).rcond() with Eigen itself."""
    af = _colmajor(a)
    n = af.shape[0]
    return float(_load_eigen().eigen_ref_llt_rcond(n, _dp(af), max(n, 1)))


def eigen_gp_grad(k: np.ndarray, y: np.ndarray) -> np.ndarray:
    """dL/dK = 1/2 (alpha alpha^T - K^{-1}), alpha = K^{-1} y, from Eigen's LLT on the CPU."""
    kf = _colmajor(k)
    yy = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
    n = kf.shape[0]
    g = np.zeros((n, n), order="F")
    rc = _load_eigen().eigen_ref_gp_grad(n, _dp(kf), max(n, 1), _dp(yy), _dp(g), max(n, 1))
    if rc != 0:
        raise RuntimeError("Eigen LLT failed on the GP kernel matrix")
    return g


def eigen_blas_trsm(side, uplo, transa, diag, alpha, a, b) -> np.ndarray:
    """Eigen's own dtrsm_ (blas/level3_impl.h:78-178) on a copy of b."""
    af = _colmajor(a)
    x = _colmajor(b).copy(order="F")
    m, n = x.shape
    I = ctypes.c_int  # noqa: E741
    _load_eigen_blas().dtrsm_(side.encode(), uplo.encode(), transa.encode(), diag.encode(), ctypes.byref(I(m)),
                              ctypes.byref(I(n)), ctypes.byref(ctypes.c_double(alpha)), _dp(af),
                              ctypes.byref(I(max(af.shape[0], 1))), _dp(x), ctypes.byref(I(max(m, 1))))
    return x


def eigen_blas_syrk(uplo, trans, alpha, a, beta, c) -> np.ndarray:
    """Eigen's own dsyrk_ (blas/level3_impl.h:357-420) on a copy of c."""
    af = _colmajor(a)
    cf = _colmajor(c).copy(order="F")
    n = cf.shape[0]
    k = af.shape[1] if trans in ("N", "n") else af.shape[0]
    I = ctypes.c_int  # noqa: E741
    _load_eigen_blas().dsyrk_(uplo.encode(), trans.encode(), ctypes.byref(I(n)), ctypes.byref(I(k)),
                              ctypes.byref(ctypes.c_double(alpha)), _dp(af), ctypes.byref(I(max(af.shape[0], 1))),
                              ctypes.byref(ctypes.c_double(beta)), _dp(cf), ctypes.byref(I(max(n, 1))))
    return cf


# ----------------------------------------------------------------------------- more restatements
def llt_blocked_f32(a: np.ndarray):
    """LLT.h:338-368 in FP32 arithmetic (LLT<MatrixXf>): (factored float32 copy, index)."""
    m = np.asfortranarray(np.asarray(a, dtype=np.float32)).copy(order="F")
    n = m.shape[0]
    r = _load_oracle().oracle_llt_blocked_f32(m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, max(n, 1))
    return m, int(r)


def llt_rank_update(l: np.ndarray, v: np.ndarray, sigma: float):
    """LLT.h:239-307 on a copy of the lower factor: (updated copy, index -1 or failing column)."""
    m = _colmajor(l).copy(order="F")
    vv = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    n = m.shape[0]
    work = np.zeros(max(n, 1))
    r = _load_oracle().oracle_llt_rank_update(_dp(m), n, max(n, 1), _dp(vv), float(sigma), _dp(work))
    return m, int(r)


def llt_rcond(l: np.ndarray, anorm: float) -> float:
    """ConditionEstimator.h:55-169 with the lower factor l and anorm = ||A||_1."""
    lf = _colmajor(l)
    n = lf.shape[0]
    work = np.zeros(max(4 * n, 1))
    return float(_load_oracle().oracle_llt_rcond(_dp(lf), n, max(n, 1), float(anorm), _dp(work)))


def trsm(side, uplo, transa, diag, alpha, a, b) -> np.ndarray:
    """blas/level3_impl.h:78-178 semantics by plain substitution, on a copy of b."""
    af = _colmajor(a)
    x = _colmajor(b).copy(order="F")
    m, n = x.shape
    _load_oracle().oracle_trsm(side.encode(), uplo.encode(), transa.encode(), diag.encode(), m, n, float(alpha),
                               _dp(af), max(af.shape[0], 1), _dp(x), max(m, 1))
    return x


def syrk(uplo, trans, alpha, a, beta, c) -> np.ndarray:
    """blas/level3_impl.h:357-420 semantics, on a copy of c."""
    af = _colmajor(a)
    cf = _colmajor(c).copy(order="F")
    n = cf.shape[0]
    k = af.shape[1] if trans in ("N", "n") else af.shape[0]
    _load_oracle().oracle_syrk(uplo.encode(), trans.encode(), n, k, float(alpha), _dp(af), max(af.shape[0], 1),
                               float(beta), _dp(cf), max(n, 1))
    return cf
