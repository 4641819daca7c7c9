"""Drop-in check: an unmodified Eigen::LLT<MatrixXd> built from the reference headers with
-DEIGEN_USE_LAPACKE calls this library's LAPACKE_dpotrf (Eigen/src/Cholesky/LLT_LAPACKE.h:59-61)
and factors on the MI355X.  The driver (oracle/eigen_lapacke_dropin.cpp) is built by
oracle/Makefile when the reference tree is available; it prints info, backward errors, whether
the upper triangle of LLT's storage still holds A, and the info of an indefinite matrix."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "eigen_lapacke_dropin")


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in driver not built (needs the reference tree)")
@pytest.mark.parametrize("n,r", [(97, 1), (500, 4), (1300, 8)])
def test_unmodified_eigen_llt_runs_on_gpu(n, r):
    print(f"[gpu] eigen drop-in n={n} r={r}", flush=True)
    out = subprocess.run([EXE, str(n), str(r)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    kv = dict(tok.split("=") for tok in out.stdout.split())
    assert int(kv["info"]) == 0  # Eigen::Success
    assert float(kv["fberr"]) <= 1e-14 and float(kv["sberr"]) <= 1e-14
    assert float(kv["upper_diff"]) == 0.0
    assert int(kv["neg_info"]) == 1  # Eigen::NumericalIssue
