"""Batched uneven SPD factor + solve (``mi_dposv_vbatched`` / ``mi_sposv_vbatched``) and its
autotuner.

Counterpart in the reference: a loop of ``LLT<MatrixXd>(A_i).solve(B_i)`` over independent
systems (Eigen/src/Cholesky/LLT.h:104-110, 430-460, 483-498).

:class:`VBatch` keeps every system in two flat device buffers (A_i column-major n_i x n_i with
lda = n_i, B_i n_i x nrhs_i with ldb = n_i) plus the device arrays the C ABI takes.  The
launch allocates nothing once the handle's batched workspace is large enough, so repeated
calls can be captured in a CUDA/HIP graph (``torch.cuda.graph``).

:func:`autotune` times the schedule parameters of the launch (workgroups of 4 worker waves
per CU, tile rows per register block, longest-first order) on the caller's batch and keeps
the fastest on the handle; results are cached per batch signature in a JSON file.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import time

import numpy as np
import torch

from . import Handle, _VP, default_handle, lib

MAX_N = 256


class VBatch:
    """Independent systems of orders ``ns`` with ``nrhs`` right-hand sides each."""

    def __init__(self, ns, nrhs, dtype=torch.float64, device="cuda"):
        ns = np.asarray(ns, dtype=np.int64)
        rs = np.asarray(nrhs, dtype=np.int64)
        if ns.shape != rs.shape or ns.ndim != 1:
            raise ValueError("ns and nrhs must be 1-D of the same length")
        self.count = int(ns.size)
        self.dtype = dtype
        self.device = torch.device(device)
        self.ns, self.rs = ns, rs
        asz = ns * ns
        bsz = ns * rs
        self.a_off = np.concatenate([[0], np.cumsum(asz)[:-1]]).astype(np.int64) if self.count else asz
        self.b_off = np.concatenate([[0], np.cumsum(bsz)[:-1]]).astype(np.int64) if self.count else bsz
        self.a_flat = torch.zeros(int(asz.sum()) + 1, dtype=dtype, device=self.device)
        self.b_flat = torch.zeros(int(bsz.sum()) + 1, dtype=dtype, device=self.device)
        es = self.a_flat.element_size()
        a_ptr = self.a_flat.data_ptr() + self.a_off * es
        b_ptr = self.b_flat.data_ptr() + self.b_off * es
        self.d_A = torch.from_numpy(a_ptr.astype(np.int64)).to(self.device)
        self.d_B = torch.from_numpy(b_ptr.astype(np.int64)).to(self.device)
        ld = np.maximum(ns, 1).astype(np.int32)
        self.d_n = torch.from_numpy(ns.astype(np.int32)).to(self.device)
        self.d_lda = torch.from_numpy(ld).to(self.device)
        self.d_nrhs = torch.from_numpy(rs.astype(np.int32)).to(self.device)
        self.d_ldb = torch.from_numpy(ld.copy()).to(self.device)
        self.d_info = torch.zeros(self.count, dtype=torch.int32, device=self.device)

    def A(self, i: int) -> torch.Tensor:
        """Column-major view of A_i (n_i x n_i)."""
        n, o = int(self.ns[i]), int(self.a_off[i])
        return self.a_flat[o:o + n * n].view(n, n).t() if n else self.a_flat[o:o].view(0, 0)

    def B(self, i: int) -> torch.Tensor:
        """Column-major view of B_i (n_i x nrhs_i)."""
        n, r, o = int(self.ns[i]), int(self.rs[i]), int(self.b_off[i])
        return self.a_flat.new_empty(0) if n * r == 0 else self.b_flat[o:o + n * r].view(r, n).t()

    def flops(self) -> float:
        """Algorithmic flops: n(n+1)(2n+1)/6 + 2 n^2 r per system (LAPACK counts)."""
        n, r = self.ns.astype(np.float64), self.rs.astype(np.float64)
        return float((n * (n + 1) * (2 * n + 1) / 6 + 2 * n * n * r).sum())

    def bytes(self) -> float:
        """Algorithmic HBM bytes: read the lower triangle of A and write L, read B and write X."""
        n, r = self.ns.astype(np.float64), self.rs.astype(np.float64)
        return float(((n * (n + 1) / 2) * 2 + 2 * n * r).sum() * self.a_flat.element_size())

    def signature(self) -> str:
        h = hashlib.sha1()
        h.update(np.bincount((self.ns + 15) // 16, minlength=17).astype(np.int64).tobytes())
        h.update(str((self.count, str(self.dtype))).encode())
        return h.hexdigest()[:16]


def posv_vbatched_(vb: VBatch, handle: Handle | None = None) -> torch.Tensor:
    """Factor every A_i in place (lower) and overwrite B_i with A_i^{-1} B_i; asynchronous.
    Returns ``vb.d_info`` (LAPACK info per system; negative: argument error)."""
    h = handle.bind_stream() if handle else default_handle()
    L = lib()
    f = L.mi_dposv_vbatched if vb.dtype == torch.float64 else L.mi_sposv_vbatched
    rc = f(h.ptr, b"L", vb.count, _VP(vb.d_n.data_ptr()), _VP(vb.d_A.data_ptr()), _VP(vb.d_lda.data_ptr()),
           _VP(vb.d_nrhs.data_ptr()), _VP(vb.d_B.data_ptr()), _VP(vb.d_ldb.data_ptr()), _VP(vb.d_info.data_ptr()))
    h.check(rc, "mi_?posv_vbatched")
    return vb.d_info


def set_config(handle: Handle, wg_per_cu: int, group: int, longest_first: bool):
    handle.check(lib().mi_llt_set_vbatched_config(handle.ptr, int(wg_per_cu), int(group), int(bool(longest_first))),
                 "mi_llt_set_vbatched_config")


def get_config(handle: Handle) -> tuple:
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    handle.check(lib().mi_llt_get_vbatched_config(handle.ptr, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                 "mi_llt_get_vbatched_config")
    return int(a.value), int(b.value), bool(c.value)


CANDIDATES = [(w, g, o) for w in (1, 2, 3, 4) for g in (2, 4) for o in (True, False)]
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "mi_llt_vbatched.json")


def autotune(vb: VBatch, handle: Handle | None = None, candidates=None, reps: int = 3, cache: str | None = CACHE,
             restore=None) -> dict:
    """Pick the fastest schedule for this batch and set it on the handle.

    Each candidate (wg_per_cu, group, longest_first) runs ``reps`` times on a scratch copy
    of the batch (``restore`` re-creates the inputs; by default the flat buffers are copied
    back before each run) and is timed with events; the best median wins.  Returns
    ``{"config": best, "times_ms": {candidate: median}}``; cached per batch signature."""
    h = handle.bind_stream() if handle else default_handle()
    key = vb.signature()
    if cache and os.path.exists(cache):
        try:
            table = json.load(open(cache))
            if key in table:
                best = tuple(table[key]["config"])
                set_config(h, *best)
                return table[key]
        except (OSError, ValueError):
            pass
    a0, b0 = vb.a_flat.clone(), vb.b_flat.clone()
    cands = candidates or CANDIDATES
    times = {}
    for c in cands:
        set_config(h, *c)
        ts = []
        for _ in range(reps + 1):
            vb.a_flat.copy_(a0)
            vb.b_flat.copy_(b0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            posv_vbatched_(vb, h)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        times[c] = float(np.median(ts[1:]))
    best = min(times, key=times.get)
    set_config(h, *best)
    vb.a_flat.copy_(a0)
    vb.b_flat.copy_(b0)
    out = {"config": list(best), "times_ms": {str(k): v for k, v in times.items()}, "when": time.time()}
    if cache:
        try:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
            table = json.load(open(cache)) if os.path.exists(cache) else {}
            table[key] = out
            json.dump(table, open(cache, "w"))
        except (OSError, ValueError):
            pass
    return out


def gram_batch(count: int, nmin: int = 8, nmax: int = 256, rmin: int = 1, rmax: int = 8, seed: int = 0,
               dtype=torch.float64, device="cuda", shift: float = 0.1) -> VBatch:
    """The batched workload: n_i ~ U{nmin..nmax}, nrhs_i ~ U{rmin..rmax} (numpy PCG64 ``seed``),
    A_i = B_i B_i^T / n_i + shift I with B_i iid N(0,1), right-hand sides iid N(0,1).  The
    Gaussians come from a torch generator seeded with ``seed``, drawn per size class in
    increasing n (setup only; FP64 products, rounded to ``dtype``)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=count)
    rs = rng.integers(rmin, rmax + 1, size=count)
    vb = VBatch(ns, rs, dtype=dtype, device=device)
    gen = torch.Generator(device=vb.device)
    gen.manual_seed(seed)
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        n = int(n)
        if n == 0:
            continue
        g = torch.randn(len(idx), n, n, dtype=torch.float64, device=vb.device, generator=gen)
        a = (torch.matmul(g, g.transpose(1, 2).contiguous()) / n).to(dtype)
        a.diagonal(dim1=1, dim2=2).add_(shift)
        bv = torch.randn(len(idx), rmax, n, dtype=torch.float64, device=vb.device, generator=gen).to(dtype)
        # one scatter per size class (A_i symmetric: row- and column-major agree; B_i column-major
        # n x r_i, the first r_i rows of bv[i])
        ia = (vb.a_off[idx][:, None] + np.arange(n * n)[None, :]).reshape(-1)
        vb.a_flat[torch.from_numpy(ia).to(vb.device)] = a.reshape(-1)
        r_of = rs[idx]
        sel = np.arange(rmax)[None, :] < r_of[:, None]                      # (systems, rmax)
        ib = vb.b_off[idx][:, None, None] + (np.arange(rmax)[None, :, None] * n + np.arange(n)[None, None, :])
        ib = ib[sel]                                                          # (sum r_i, n)
        if ib.size:
            vals = bv[torch.from_numpy(sel).to(vb.device)]                    # same order
            vb.b_flat[torch.from_numpy(ib.reshape(-1)).to(vb.device)] = vals.reshape(-1)
    return vb
