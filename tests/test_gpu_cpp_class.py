"""The C++ host class mi::LLT (include/mi_llt.hpp), the mirror of Eigen::LLT's surface
(Eigen/src/Cholesky/LLT.h:66-233), from compiled C++:

* tests/cpp/test_llt_class.cpp with the library's own HostMatrix: compute, info, solve,
  solveInPlace, matrixL/U/LLT, rcond, rankUpdate, l1Norm, Upper, float, NumericalIssue and
  misuse, against long-double loops;
* oracle/eigen_class_dropin.cpp (built when the reference tree is present) with Eigen's own
  MatrixXd: the same program text with Eigen::LLT and with mi::LLT, results compared."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "build", "test_llt_class")
EIGEN_EXE = os.path.join(ROOT, "oracle", "_ref", "eigen_class_dropin")


def run(exe):
    print(f"[gpu] {os.path.basename(exe)}", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    return {k: float(v) for k, v in (tok.split("=") for tok in out.stdout.split())}


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ test not built (run __graft_entry__.build())")
def test_cpp_llt_class():
    kv = run(EXE)
    assert kv["info"] == 0 and kv["fberr"] <= 1e-14 and kv["sberr"] <= 1e-14
    assert kv["upper_zero"] == 0.0 and kv["lu_transpose"] == 0.0 and kv["junk_kept"] == 1.0
    assert kv["inplace_diff"] == 0.0
    assert kv["l1_diff"] <= 1e-14
    assert 0.0 < kv["rcond"] <= 1.0
    assert kv["update_info"] == 0 and kv["update_berr"] <= 1e-14
    assert kv["upper_info"] == 0 and kv["upper_fberr"] <= 1e-14 and kv["upper_sberr"] <= 1e-14
    assert kv["float_info"] == 0 and kv["float_fberr"] <= 2e-7 and kv["float_sberr"] <= 2e-6
    assert kv["neg_info"] == 1 and kv["misuse"] == 1


@pytest.mark.skipif(not os.path.exists(EIGEN_EXE), reason="Eigen class drop-in not built (needs the reference tree)")
def test_cpp_llt_class_with_eigen_types():
    kv = run(EIGEN_EXE)
    assert kv["info_eigen"] == kv["info_mi"] == 0
    assert kv["l_diff"] <= 1e-12 and kv["x_diff"] <= 1e-10
    assert kv["rcond_rel"] <= 1e-10
    assert kv["update_diff"] <= 1e-12
    assert kv["neg_eigen"] == kv["neg_mi"] == 1
