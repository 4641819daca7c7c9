"""Drop-in check: an unmodified Eigen::LLT built from the reference headers calls this library.

* -DEIGEN_USE_LAPACKE: LLT<MatrixXd>::compute calls LAPACKE_dpotrf and LLT<MatrixXf>::compute
  LAPACKE_spotrf (Eigen/src/Cholesky/LLT_LAPACKE.h:59-61, :90), so the factor runs on the MI355X.
* -DEIGEN_USE_LAPACKE -DEIGEN_USE_BLAS as well: LLT::solve's triangular solves go to dtrsm_
  (TriangularSolverMatrix_BLAS.h:41-99), and the panel solve and rank update of
  llt_inplace::blocked (LLT.h:364-365), written as Eigen writes them, to dtrsm_ (side 'R') and
  dsyrk_ (GeneralMatrixMatrixTriangular_BLAS.h:47-134).
The drivers (oracle/eigen_lapacke_dropin.cpp) are built by oracle/Makefile when the reference
tree is available; each prints info, backward errors (long-double loops), whether the upper
triangle of LLT's storage still holds A, and the info of an indefinite matrix."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
EXE = os.path.join(REF, "eigen_lapacke_dropin")
BLAS_EXE = os.path.join(REF, "eigen_blas_dropin")


def run(exe, *args):
    print(f"[gpu] {os.path.basename(exe)} {' '.join(map(str, args))}", flush=True)
    out = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    return dict(tok.split("=") for tok in out.stdout.split())


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in driver not built (needs the reference tree)")
@pytest.mark.parametrize("n,r", [(97, 1), (500, 4), (1300, 8)])
def test_unmodified_eigen_llt_runs_on_gpu(n, r):
    kv = run(EXE, n, r)
    assert int(kv["info"]) == 0  # Eigen::Success
    assert float(kv["fberr"]) <= 1e-14 and float(kv["sberr"]) <= 1e-14
    assert float(kv["upper_diff"]) == 0.0
    assert int(kv["neg_info"]) == 1  # Eigen::NumericalIssue


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in driver not built (needs the reference tree)")
@pytest.mark.parametrize("n,r", [(97, 1), (700, 4)])
def test_unmodified_eigen_llt_float_runs_on_gpu(n, r):
    kv = run(EXE, n, r, "f")
    assert int(kv["info"]) == 0
    assert float(kv["fberr"]) <= 2e-7 and float(kv["sberr"]) <= 2e-6  # FP32 backward errors
    assert float(kv["upper_diff"]) == 0.0
    assert int(kv["neg_info"]) == 1


@pytest.mark.skipif(not os.path.exists(BLAS_EXE), reason="BLAS drop-in driver not built (needs the reference tree)")
@pytest.mark.parametrize("n,r", [(200, 3), (900, 16)])
def test_unmodified_eigen_blas_route_runs_on_gpu(n, r):
    kv = run(BLAS_EXE, n, r)
    assert int(kv["info"]) == 0
    assert float(kv["fberr"]) <= 1e-14 and float(kv["sberr"]) <= 1e-14  # solve through dtrsm_
    assert float(kv["upper_diff"]) == 0.0 and int(kv["neg_info"]) == 1
    assert float(kv["trsm_rel"]) <= 1e-12   # LLT.h:364 through dtrsm_ ('R')
    assert float(kv["syrk_rel"]) <= 1e-12   # LLT.h:365 through dsyrk_
    assert float(kv["syrk_upper_diff"]) == 0.0
