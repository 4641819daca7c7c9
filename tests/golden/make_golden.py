"""Write tests/golden/eigen_llt_golden.npz from the genuine Eigen LLT.

Run in the container that holds the reference tree, after ``make -C oracle`` has compiled
``oracle/_ref/libeigen_ref_*.so`` from the reference's own headers (oracle/eigen_ref.cpp):

    python tests/golden/make_golden.py

The fixture holds inputs and Eigen's outputs only (data, no reference source):
  case "<tag>": A (input, column-major), B (right-hand sides), LLT (Eigen's matrixLLT():
  factor in the lower triangle, untouched input above), X (Eigen's solve), IDX (Eigen's
  llt_inplace::blocked return: -1 or the failing column), INFO (0 Success / 1 NumericalIssue).
Inputs come from mi_llt.workloads (seeded PCG64), so tests can regenerate and compare them.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "eigen-git-mirror_amd"))

import oracle  # noqa: E402
from mi_llt import workloads  # noqa: E402  (numpy only; the library itself is not loaded)

OUT = os.path.join(HERE, "eigen_llt_golden.npz")

SPD_SIZES = [1, 2, 3, 7, 8, 16, 31, 32, 33, 64, 100, 127, 128, 129, 200]
FAIL_CASES = [(1, 0), (5, 3), (33, 0), (33, 32), (100, 17), (129, 128), (200, 150)]


def case(store: dict, tag: str, a: np.ndarray, b: np.ndarray, upper: bool = False):
    m, idx = oracle.eigen_llt_inplace(a, upper=upper)
    store[f"{tag}/A"] = a
    store[f"{tag}/IDX"] = np.array(idx)
    if upper:
        store[f"{tag}/LLT"] = m
        return
    res = oracle.eigen_llt(a, b)
    store[f"{tag}/B"] = b
    store[f"{tag}/LLT"] = res["L"]
    store[f"{tag}/X"] = res["X"]
    store[f"{tag}/INFO"] = np.array(res["info"])
    assert np.array_equal(res["L"], m), "LLT::compute and llt_inplace::blocked disagree"


def main():
    if not oracle.have_eigen():
        raise SystemExit("oracle/_ref is not built (needs the reference tree): run make -C oracle")
    store: dict[str, np.ndarray] = {}
    # doc/snippets/LLT_example.cpp:1-2
    kat = np.asfortranarray(np.array([[4.0, -1.0, 2.0], [-1.0, 6.0, 0.0], [2.0, 0.0, 5.0]]))
    case(store, "kat3", kat, np.asfortranarray(np.eye(3)))
    for n in SPD_SIZES:
        a = workloads.gram_spd(n, seed=n)
        b = workloads.rhs(n, 3, seed=n + 7)
        case(store, f"spd{n}", a, b)
    for n in (33, 129):
        case(store, f"upper{n}", workloads.gram_spd(n, seed=100 + n), None, upper=True)
    for n, k in FAIL_CASES:
        a = workloads.indefinite_at(n, k, seed=n)
        case(store, f"fail{n}_{k}", a, workloads.rhs(n, 1, seed=1))
    store["eigen_version"] = np.array(oracle._load_eigen().eigen_ref_version())
    np.savez_compressed(OUT, **store)
    print(f"wrote {OUT}: {len(store)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
