"""Calls on several streams at once: each stream gets its own library handle (workspace,
tickets and flags of one handle serve one stream's calls in order), so independent
factorizations, solves and adjoints run concurrently and give the same bits as one after the
other.

A fresh handle's synchronisation words used to be zeroed by a memset on the null stream,
which is not ordered with torch's non-blocking streams: a factorization queued right after
on such a stream could draw its workgroup roles from an unzeroed ticket counter.  They are
now zeroed by a kernel on the handle's own stream, the counters are reset at the start of
every factorization, and a role outside its launch stops with a status bit instead of
becoming an address."""
import numpy as np
import pytest
import torch

import mi_llt
from mi_llt import autograd as mad
from mi_llt import workloads

pytestmark = pytest.mark.gpu


def test_concurrent_streams_match_sequential(dev):
    print("[gpu] 8 systems on 4 streams vs sequential", flush=True)
    sizes = [300, 512, 700, 129, 1000, 64, 512, 385]
    mats = [torch.from_numpy(workloads.gram_spd(n, seed=50 + i)).to(dev) for i, n in enumerate(sizes)]
    rhs = [torch.from_numpy(workloads.rhs(n, 3, seed=80 + i)).to(dev) for i, n in enumerate(sizes)]

    def run_one(a, b):
        l = mi_llt.colmajor(a.clone())
        assert mi_llt.potrf_(l) == 0
        x = mi_llt.colmajor(b.clone())
        mi_llt.potrs_(l, x)
        k1 = a.clone().requires_grad_(True)
        (mad.cholesky(k1) * 0.5).sum().backward()
        return torch.tril(l), x, k1.grad

    seq = [run_one(a, b) for a, b in zip(mats, rhs)]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in range(4)]
    main = torch.cuda.current_stream()
    for st in streams:
        st.wait_stream(main)
    par = [None] * len(sizes)
    for i, (a, b) in enumerate(zip(mats, rhs)):
        with torch.cuda.stream(streams[i % 4]):
            par[i] = run_one(a, b)
    for st in streams:
        main.wait_stream(st)
    torch.cuda.synchronize()
    for i, n in enumerate(sizes):
        for got, want in zip(par[i], seq[i]):
            assert torch.equal(got, want), f"n={n}: stream result differs from the sequential one"
    for st in streams:
        with torch.cuda.stream(st):
            assert mi_llt.default_handle().status() == 0
