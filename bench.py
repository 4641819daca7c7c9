#!/usr/bin/env python3
"""SPD factor + solve throughput on one MI355X (FP64 LLT drop-in).

Workload: the plain FP64 LLT at n=8192 with 64 right-hand sides,
A = B B^T / n + 0.1 I (B iid N(0,1), PCG64 seed 0), right-hand sides iid N(0,1) (seed 1).
A step is what Eigen's ``LLT<MatrixXd>(A).solve(B)`` does: copy A into the object's own
storage (LLT.h:438-440), factor it in place, copy B and solve -- here
``mi_dposv_async`` on device-resident copies.  Inputs are resident in HBM before the
timed region.

TFLOP/s = [F(n) + 2 n^2 r] / time, F(n) = n(n+1)(2n+1)/6 (LAPACK / Eigen bench count,
bench/benchCholesky.cpp:34-39).

Also reported: normwise backward errors of the last step, the dominant kernel's roofline
(the trailing SYRK, from live HIP events around each launch in separate traced steps), and
the genuine Eigen LLT timed on this host's cores (``cpu_baseline``).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "eigen-git-mirror_amd"))

METRIC = "SPD factor+solve TFLOP/s (% MFMA peak); mixed vs FP64 at equal backward err"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 MFMA, vendor figure
PEAK_HBM_GBS = 8000.0


def flops_factor(n: int) -> float:
    return n * (n + 1) * (2 * n + 1) / 6.0


def flops_solve(n: int, r: int) -> float:
    return 2.0 * n * n * r


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(a_host, b_host, n, r):
    """Genuine Eigen LLT<MatrixXd>::compute + solve (oracle/_ref) on this host, one run of
    the full workload; falls back to the oracle's C restatement when the Eigen build is
    absent.  Eigen's LLT never reaches its OpenMP GEMM (SelfadjointProduct.h:106,
    SolveTriangular.h:98), so it uses one core."""
    import oracle

    if oracle.have_eigen():
        t = oracle.eigen_llt_time(a_host, b_host)
        kind = "reference"
        what = f"Eigen {oracle._load_eigen().eigen_ref_version()} LLT<MatrixXd> compute()+solve(), " \
               f"{os.path.basename(oracle.eigen_path())}"
    else:
        t0 = time.perf_counter()
        m, _ = oracle.llt_blocked(a_host)
        oracle.llt_solve(m, b_host)
        t = time.perf_counter() - t0
        kind = "port"
        what = "oracle C restatement of llt_inplace::blocked + solve"
    fl = flops_factor(n) + flops_solve(n, r)
    return {"value": fl / t / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": kind,
            "sample": f"one full n={n}, {r}-RHS factor+solve ({what}); wall {t:.3f} s; "
                      f"host threads available {os.cpu_count()}, used by LLT 1",
            "seconds": t}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X and the path is built for one GPU "
              f"(got --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import mi_llt
    from mi_llt import workloads

    n, r = args.n, args.nrhs
    dev = torch.device("cuda:0")
    log(f"[bench] generating n={n} r={r}")
    a = workloads.gram_spd_torch(n, seed=0, device=dev)
    b = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, r, seed=1)).to(dev))
    da = torch.empty_strided((n, n), (1, n), dtype=torch.float64, device=dev)
    db = torch.empty_strided((n, r), (1, n), dtype=torch.float64, device=dev)
    d_info = torch.zeros(1, dtype=torch.int32, device=dev)
    h = mi_llt.Handle().bind_stream()

    def step():
        da.copy_(a)
        db.copy_(b)
        mi_llt.posv_async_(da, db, d_info, handle=h)

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    log(f"[bench] timing {args.steps} steps")
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1) / args.steps
    info = int(d_info.item())
    if h.status() != 0:
        raise SystemExit("a solve sweep reported a timeout")
    if info != 0:
        raise SystemExit(f"factorization reported info={info} on an SPD matrix")
    total_flops = flops_factor(n) + flops_solve(n, r)
    tflops = total_flops / (ms * 1e-3) / 1e12

    # backward errors of the last step (FP64 products on the GPU, outside the timed region)
    l = torch.tril(da)
    asym = torch.tril(a) + torch.tril(a, -1).t()
    anorm = torch.linalg.matrix_norm(asym)
    fberr = float(torch.linalg.matrix_norm(asym - l @ l.t()) / anorm)
    res = asym @ db - b
    sberr = float((res.norm(dim=0) / (anorm * db.norm(dim=0))).max())
    del l, asym, res

    # per-kernel durations from HIP events around each launch (separate traced steps)
    h.trace(True)
    h.trace_reset()
    n_traced = 2
    for _ in range(n_traced):
        step()
    tr = h.trace_read()
    h.trace(False)
    kern = {}
    for k, (tms, fl, nl) in tr.items():
        if nl:
            kern[k] = {"ms_per_step": tms / n_traced, "launches_per_step": nl // n_traced,
                       "tflops": fl / (tms * 1e-3) / 1e12 if tms > 0 else None}
    # dominant kernel: factor_step_kernel (one launch per block step: panel k+1 + trailing
    # update by panel k), FP64 MFMA; achieved = algorithmic flops of the launches / their time
    st_ms, st_fl, st_n = tr["factor_step"]
    achieved = st_fl / (st_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "kernel": "factor_step_kernel (panel POTRF/TRSM + trailing SYRK, FP64 MFMA)",
                "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                "per_launch": {"avg_ms": st_ms / st_n, "avg_algorithmic_flops": st_fl / st_n,
                               "launches": st_n}}

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (Eigen LLT on the host)")
        cpu = cpu_baseline(np.asfortranarray(a.cpu().numpy()), np.asfortranarray(b.cpu().numpy()), n, r)

    out = {
        "metric": METRIC, "value": round(tflops, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"FP64 LLT factor+solve, n={n}, {r} RHS, A=B*B^T/n+0.1I (seed 0), RHS N(0,1)",
                   "n": n, "nrhs": r, "uplo": "L", "block": 128,
                   "step": "copy A and B into working storage + mi_dposv_async (factor + solve)"},
        "pct_fp64_mfma_peak": round(100.0 * tflops / PEAK_FP64_TFLOPS, 2),
        "backward_error": {"factor": fberr, "solve": sberr},
        "kernels": kern,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "wall_s_timed": round(t_wall, 4),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
