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
(factor_step_kernel, from live HIP events around each launch in separate traced steps), and
the genuine Eigen LLT timed on this host's cores (``cpu_baseline``).

``legs`` (skipped with --no-legs) measures the other workloads of the path, each on its own
configuration: the mixed-precision solve with FP64 refinement against the plain FP64 path
at n=16384 (kappa 1e2, 1e4, 1e6, 16 right-hand sides), the batched uneven solve of 20,000
systems (graph replay), and the Gaussian-process adjoint (n=4096, and 256 systems of 512).
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
    """Genuine Eigen LLT<MatrixXd>::compute + solve (oracle/_ref) on this host, one run of the
    full workload with Eigen allowed 1 thread and one with every host thread (Eigen's LLT never
    reaches its parallel GEMM -- SelfadjointProduct.h:106, SolveTriangular.h:98 -- so both use
    one core; the faster run is reported, with its thread setting); falls back to the oracle's
    C restatement when the Eigen build is absent."""
    import oracle

    fl = flops_factor(n) + flops_solve(n, r)
    if oracle.have_eigen():
        runs = {}
        for threads in (1, os.cpu_count() or 1):
            got = oracle.eigen_set_threads(threads)
            runs[got] = oracle.eigen_llt_time(a_host, b_host)
        best = min(runs, key=runs.get)
        t = runs[best]
        what = f"Eigen {oracle._load_eigen().eigen_ref_version()} LLT<MatrixXd> compute()+solve(), " \
               f"{os.path.basename(oracle.eigen_path())}"
        return {"value": fl / t / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "reference",
                "sample": f"one full n={n}, {r}-RHS factor+solve ({what}); wall "
                          + ", ".join(f"{v:.3f} s with Eigen::setNbThreads({k})" for k, v in sorted(runs.items()))
                          + f" -- the LLT path runs on one core either way; host threads {os.cpu_count()}",
                "seconds": t, "seconds_by_threads": {str(k): v for k, v in runs.items()}}
    t0 = time.perf_counter()
    m, _ = oracle.llt_blocked(a_host)
    oracle.llt_solve(m, b_host)
    t = time.perf_counter() - t0
    return {"value": fl / t / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"one full n={n}, {r}-RHS factor+solve (oracle C restatement of llt_inplace::blocked + "
                      f"solve, one thread); wall {t:.3f} s", "seconds": t}


def host_api_rate(a_host, b_host, n, r):
    """The drop-in path end to end: LAPACKE_dpotrf then LAPACKE_dpotrs on host (pageable)
    arrays, each staging its operands over PCIe (what an unmodified Eigen built with
    EIGEN_USE_LAPACKE pays per LLT::compute); one warm call, then one timed.  Reported beside
    the device-resident value, never as it."""
    import numpy as np

    import mi_llt

    lib = mi_llt.lib()
    fl = flops_factor(n) + flops_solve(n, r)
    times = []
    for _ in range(2):
        a = np.asfortranarray(a_host.copy())
        b = np.asfortranarray(b_host.copy())
        t0 = time.perf_counter()
        i1 = lib.LAPACKE_dpotrf(102, b"L", n, a.ctypes.data, n)
        i2 = lib.LAPACKE_dpotrs(102, b"L", n, r, a.ctypes.data, n, b.ctypes.data, n)
        times.append(time.perf_counter() - t0)
        if i1 != 0 or i2 != 0:
            raise SystemExit(f"LAPACKE shims returned {i1}, {i2}")
    t = times[-1]
    return {"value": round(fl / t / 1e12, 4), "unit": "TFLOP/s", "ms": round(t * 1e3, 2),
            "what": "LAPACKE_dpotrf + LAPACKE_dpotrs on host arrays (H2D/D2H of A twice, B once each way)"}


def _time_ms(fn, reps, setup=None):
    """Median of ``reps`` event-timed calls (``setup`` runs before each, untimed)."""
    import torch

    ts = []
    for _ in range(reps):
        if setup:
            setup()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    ts.sort()
    return ts[len(ts) // 2], ts[0]


def leg_mixed(n=16384, r=16, kappas=(1e2, 1e4, 1e6), reps=3):
    """Mixed-precision LLT + FP64 refinement vs the plain FP64 factor+solve, same inputs.
    Throughput counts FP64-equivalent flops F(n) + 2 n^2 r for every precision; the mixed
    time includes the FP32 copy, refinement steps and their FP64 residuals."""
    import torch

    import mi_llt
    from mi_llt import workloads

    dev = torch.device("cuda:0")
    fl = flops_factor(n) + flops_solve(n, r)
    b = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, r, seed=1)).to(dev))
    out = {"n": n, "nrhs": r, "tol": 1e-14, "cases": []}
    h = mi_llt.Handle().bind_stream()
    for kappa in kappas:
        a = workloads.spectrum_spd_torch(n, kappa, seed=0, device=dev)
        # plain FP64: copy into working storage + factor + solve (mi_dposv_async)
        da = torch.empty_strided((n, n), (1, n), dtype=torch.float64, device=dev)
        xb = torch.empty_strided((n, r), (1, n), dtype=torch.float64, device=dev)
        d_info = torch.zeros(1, dtype=torch.int32, device=dev)

        def plain():
            da.copy_(a)
            xb.copy_(b)
            mi_llt.posv_async_(da, xb, d_info, handle=h)

        plain()
        t64, _ = _time_ms(plain, reps)
        asym = torch.tril(a) + torch.tril(a, -1).t()
        be64 = float(((asym @ xb - b).norm(dim=0) / (torch.linalg.matrix_norm(asym) * xb.norm(dim=0))).max())
        del da, asym
        case = {"kappa": kappa, "fp64": {"ms": round(t64, 3), "tflops": round(fl / t64 / 1e9, 3), "berr": be64}}
        best = None
        for prec in ("fp32", "bf16x6", "bf16x3", "bf16"):
            res = {}

            def mixed():
                res["r"] = mi_llt.posv_mixed(a, b, prec=prec, tol=1e-14, maxit=30, handle=h)

            mixed()
            tm, _ = _time_ms(mixed, reps)
            x, iters, berr, info = res["r"]
            case[prec] = {"ms": round(tm, 3), "tflops": round(fl / tm / 1e9, 3), "iters": iters, "berr": berr,
                          "speedup_vs_fp64": round(t64 / tm, 3), "info": info}
            if info == 0 and berr <= 1e-14 and (best is None or tm < case[best]["ms"]):
                best = prec
        case["best"] = best
        out["cases"].append(case)
        del a
        torch.cuda.empty_cache()
    out["speedup_best"] = {str(c["kappa"]): (c[c["best"]]["speedup_vs_fp64"] if c["best"] else None)
                           for c in out["cases"]}
    return out


def leg_batched(count=20000, reps=5):
    """20,000 uneven SPD systems (n ~ U{8..256}, 1..8 right-hand sides) factored and solved in
    one persistent launch, replayed from a HIP graph; inputs restored between replays
    (untimed).  Roofline: per system max(bytes / 8 TB/s, flops / peak), summed."""
    import numpy as np
    import torch

    import mi_llt
    from mi_llt import batched

    dev = torch.device("cuda:0")
    out = {"count": count}
    for dt, peak in ((torch.float64, PEAK_FP64_TFLOPS), (torch.float32, 157.3)):
        vb = batched.gram_batch(count, seed=0, dtype=dt, device=dev)
        a0, b0 = vb.a_flat.clone(), vb.b_flat.clone()
        h = mi_llt.Handle().bind_stream()

        def restore():
            vb.a_flat.copy_(a0)
            vb.b_flat.copy_(b0)

        tune = batched.autotune(vb, h, candidates=[(2, 4, True), (3, 2, True), (2, 2, True), (4, 2, True)],
                                reps=2, cache=None, )
        restore()
        batched.posv_vbatched_(vb, h)
        torch.cuda.synchronize()
        restore()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s):
            batched.posv_vbatched_(vb, h)
        tm, tb = _time_ms(g.replay, reps, setup=restore)
        bad = int((vb.d_info != 0).sum())
        n, r = vb.ns.astype(np.float64), vb.rs.astype(np.float64)
        es = 8 if dt == torch.float64 else 4
        sys_bytes = ((n * (n + 1) / 2) * 2 + 2 * n * r) * es
        sys_flops = n * (n + 1) * (2 * n + 1) / 6 + 2 * n * n * r
        t_roof = float(np.maximum(sys_bytes / (PEAK_HBM_GBS * 1e9), sys_flops / (peak * 1e12)).sum()) * 1e3
        out["fp64" if dt == torch.float64 else "fp32"] = {
            "ms": round(tm, 4), "best_ms": round(tb, 4), "systems_per_s": round(count / (tm * 1e-3), 1),
            "tflops": round(vb.flops() / tm / 1e9, 3), "algorithmic_GBps": round(vb.bytes() / tm / 1e6, 1),
            "roofline_ms": round(t_roof, 4), "frac_of_roofline": round(t_roof / tm, 4), "failed": bad,
            "config": {"wg_per_cu": tune["config"][0], "group": tune["config"][1], "longest_first": tune["config"][2]},
            "graph": True}
        del vb, a0, b0, g
        torch.cuda.empty_cache()
    return out


def leg_adjoint(n=4096, nb=512, batch=256, reps=3):
    """Gaussian-process log marginal likelihood and its gradient dL/dK through the library's
    factorization, solve and hand-written adjoints (torch.autograd.Function)."""
    import torch

    from mi_llt import autograd as mad
    from mi_llt import workloads

    dev = torch.device("cuda:0")
    out = {}
    k, y = workloads.gp_kernel_torch(n, seed=0, device=dev)

    def fwd_bwd():
        k1 = k.detach().requires_grad_(True)
        mad.gp_log_marginal(k1, y).backward()
        return k1.grad

    fwd_bwd()
    t_all, _ = _time_ms(fwd_bwd, reps)

    def fwd():
        with torch.no_grad():
            mad.gp_log_marginal(k, y)

    t_f, _ = _time_ms(fwd, reps)
    # adjoint flops: tril(L^T Lbar) n^3/3 + two triangular solves with n right-hand sides 2 n^3
    # (the textbook count); the library's second solve computes only the block-upper half of
    # the symmetric result, n^3/3 + n^3 + n^3/2 performed
    adj_fl = n ** 3 / 3 + 2.0 * n ** 3
    adj_done = n ** 3 / 3 + 1.5 * n ** 3
    t_adj = max(t_all - t_f, 1e-6)
    out["single"] = {"n": n, "forward_ms": round(t_f, 3), "forward_backward_ms": round(t_all, 3),
                     "adjoint_ms": round(t_adj, 3), "adjoint_tflops": round(adj_fl / t_adj / 1e9, 3),
                     "adjoint_frac_fp64_peak": round(adj_fl / t_adj / 1e9 / PEAK_FP64_TFLOPS, 4),
                     "adjoint_tflops_performed": round(adj_done / t_adj / 1e9, 3),
                     "adjoint_frac_fp64_peak_performed": round(adj_done / t_adj / 1e9 / PEAK_FP64_TFLOPS, 4)}
    ks, ys = zip(*[workloads.gp_kernel_torch(nb, seed=100 + i, device=dev) for i in range(batch)])
    kb = torch.stack(ks)
    yb = torch.stack(ys)

    def batch_all():
        # one batched forward + backward: every kernel launch covers all systems
        k1 = kb.detach().requires_grad_(True)
        mad.gp_log_marginal_batched(k1, yb).sum().backward()
        return k1.grad

    def batch_fwd():
        with torch.no_grad():
            mad.gp_log_marginal_batched(kb, yb)

    batch_all()
    tb, _ = _time_ms(batch_all, reps)
    tbf, _ = _time_ms(batch_fwd, reps)
    tba = max(tb - tbf, 1e-6)
    # forward F(n) + 2 n^2 (solve) and adjoint n^3/3 + 2 n^3 per system (LAPACK / textbook counts)
    fl_f = batch * (flops_factor(nb) + 2.0 * nb * nb)
    fl_a = batch * (nb ** 3 / 3 + 2.0 * nb ** 3)
    out["batch"] = {"count": batch, "n": nb, "forward_backward_ms": round(tb, 3), "forward_ms": round(tbf, 3),
                    "adjoint_ms": round(tba, 3), "tflops": round((fl_f + fl_a) / tb / 1e9, 3),
                    "frac_fp64_peak": round((fl_f + fl_a) / tb / 1e9 / PEAK_FP64_TFLOPS, 4),
                    "adjoint_tflops": round(fl_a / tba / 1e9, 3),
                    "adjoint_frac_fp64_peak": round(fl_a / tba / 1e9 / PEAK_FP64_TFLOPS, 4),
                    "path": "mi_dposv_batched_async forward, "
                            "mi_dposv_adjoint_batched backward"}
    return out


def counter_workload(name: str, reps: int):
    """A short, fixed workload for hardware-counter runs (rocprofv3 --pmc ... -- python3 bench.py
    --counter-workload NAME): first two 256 MiB device copies (the calibration kernel of the
    FETCH_SIZE / WRITE_SIZE scale factors), then `reps` calls of one path.  Prints nothing on
    stdout (no bench line)."""
    import torch

    import mi_llt
    from mi_llt import autograd as mad
    from mi_llt import batched, workloads

    dev = torch.device("cuda:0")
    src = torch.ones(32 * 2 ** 20, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    h = mi_llt.Handle().bind_stream()
    if name in ("factor_fp64", "solve_fp64"):  # the headline workload: n = 8192, 64 RHS
        n, r = 8192, 64
        a = workloads.gram_spd_torch(n, seed=0, device=dev)
        b = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, r, seed=1)).to(dev))
        da, db = a.clone(), b.clone()
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        for _ in range(reps):
            da.copy_(a)
            db.copy_(b)
            mi_llt.posv_async_(da, db, info, handle=h)
    elif name in ("mixed_bf16x3", "mixed_bf16x6", "mixed_fp32"):  # n = 16384, kappa 1e2, 16 RHS
        n, r = 16384, 16
        a = workloads.spectrum_spd_torch(n, 1e2, seed=0, device=dev)
        b = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, r, seed=1)).to(dev))
        for _ in range(reps):
            mi_llt.posv_mixed(a, b, prec=name.split("_")[1], tol=1e-14, maxit=30, handle=h)
    elif name in ("vbatched_fp64", "vbatched_fp32"):  # 20,000 uneven systems
        dt = torch.float64 if name.endswith("64") else torch.float32
        vb = batched.gram_batch(20000, seed=0, dtype=dt, device=dev)
        a0, b0 = vb.a_flat.clone(), vb.b_flat.clone()
        for _ in range(reps):
            vb.a_flat.copy_(a0)
            vb.b_flat.copy_(b0)
            batched.posv_vbatched_(vb, h)
    elif name == "adjoint":  # GP gradient, n = 4096
        k, y = workloads.gp_kernel_torch(4096, seed=0, device=dev)
        for _ in range(reps):
            k1 = k.detach().requires_grad_(True)
            mad.gp_log_marginal(k1, y).backward()
    elif name == "adjoint_batch":  # GP gradients, 256 systems of n = 512
        ks, ys = zip(*[workloads.gp_kernel_torch(512, seed=100 + i, device=dev) for i in range(256)])
        kb, yb = torch.stack(ks), torch.stack(ys)
        for _ in range(reps):
            k1 = kb.detach().requires_grad_(True)
            mad.gp_log_marginal_batched(k1, yb).sum().backward()
    else:
        raise SystemExit(f"unknown counter workload {name!r}")
    torch.cuda.synchronize()
    if h.status() != 0:
        raise SystemExit("device status reported a failure")
    log(f"[bench] counter workload {name}: {reps} calls done")


COUNTER_WORKLOADS = ("factor_fp64", "solve_fp64", "mixed_bf16x3", "mixed_bf16x6", "mixed_fp32", "vbatched_fp64",
                     "vbatched_fp32", "adjoint", "adjoint_batch")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-legs", action="store_true", help="skip the mixed / batched / adjoint legs")
    ap.add_argument("--no-host-api", action="store_true", help="skip the PCIe-inclusive LAPACKE drop-in timing")
    ap.add_argument("--counter-workload", choices=COUNTER_WORKLOADS,
                    help="run a short fixed workload for rocprofv3 counter runs instead of the bench")
    ap.add_argument("--reps", type=int, default=2, help="calls of the counter workload")
    args = ap.parse_args()
    if args.counter_workload:
        counter_workload(args.counter_workload, args.reps)
        return
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
    # HBM bytes per launch: rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload, scaled by a
    # copy kernel of known size (committed summary); none for other sizes
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "factor_step_fp64_n8192_traffic.json")
    if n == 8192 and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["hbm_bytes_per_dispatch"]
    roofline = {"bound": "mfma", "kernel": "factor_step_kernel (panel POTRF/TRSM + trailing SYRK, FP64 MFMA)",
                "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                "per_launch": {"avg_ms": st_ms / st_n, "avg_algorithmic_flops": st_fl / st_n,
                               "launches": st_n}}

    cpu = None
    host_api = None
    if not args.no_cpu or not args.no_host_api:
        a_host, b_host = np.asfortranarray(a.cpu().numpy()), np.asfortranarray(b.cpu().numpy())
        if not args.no_host_api:
            log("[bench] host-API (LAPACKE drop-in, PCIe-inclusive) timing")
            host_api = host_api_rate(a_host, b_host, n, r)
        if not args.no_cpu:
            log("[bench] cpu baseline (Eigen LLT on the host)")
            cpu = cpu_baseline(a_host, b_host, n, r)
        del a_host, b_host

    legs = None
    if not args.no_legs:
        del da, db, a, b
        torch.cuda.empty_cache()
        legs = {"mixed": leg_mixed(), "batched": leg_batched(), "adjoint": leg_adjoint()}

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
        "host_api": host_api,
        "wall_s_timed": round(t_wall, 4),
        "legs": legs,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
