"""Robustness of the factorization's device counters and of the failure reporting.

* Graph capture: the factorization and the factor + solve draw workgroup roles from a ticket
  counter and wait on a panel flag; both are reset on the stream at the start of every call,
  so a captured call replays with the same accounting as an eager one.  Replays (with eager
  work of the program in between) and an eager call after the capture must give the eager
  result bit for bit (the kernels' sums run in a fixed order).
* Failure path: a device status bit set from the host (as a timeout would set it) makes the
  queued factorization skip its work, and the synchronous entry point report
  MI_LLT_E_DEVICE instead of info = 0; the status is then clear and the next call is normal.
"""
import numpy as np
import pytest
import torch

import mi_llt
from mi_llt import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,r", [(700, 5), (1537, 16)])
def test_factor_and_solve_graph_replay(dev, n, r):
    print(f"[gpu] graph replay posv n={n} r={r}", flush=True)
    a0 = mi_llt.colmajor(torch.from_numpy(workloads.gram_spd(n, seed=n)).to(dev))
    b0 = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, r, seed=n + 1)).to(dev))
    a, b = a0.clone(), b0.clone()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    h = mi_llt.Handle()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # eager on the capture stream: sizes the workspace
        mi_llt.posv_async_(a, b, info, handle=h)
    torch.cuda.synchronize()
    ref_a, ref_b = a.clone(), b.clone()
    assert int(info.item()) == 0
    g = torch.cuda.CUDAGraph()
    a.copy_(a0)
    b.copy_(b0)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        mi_llt.posv_async_(a, b, info, handle=h)
    for rep in range(3):
        a.copy_(a0)
        b.copy_(b0)
        info.fill_(-7)
        _ = (a * 2.0).sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert int(info.item()) == 0, f"replay {rep}"
        assert torch.equal(a, ref_a) and torch.equal(b, ref_b), f"replay {rep}"
    # an eager call on the same handle after the capture
    a.copy_(a0)
    b.copy_(b0)
    with torch.cuda.stream(s):
        mi_llt.posv_async_(a, b, info, handle=h)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    assert torch.equal(a, ref_a) and torch.equal(b, ref_b)
    assert h.bind_stream(s).status() == 0


def test_factor_only_graph_replay(dev):
    n = 1000
    print(f"[gpu] graph replay potrf n={n}", flush=True)
    a0 = mi_llt.colmajor(torch.from_numpy(workloads.gram_spd(n, seed=5)).to(dev))
    a = a0.clone()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    h = mi_llt.Handle()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        mi_llt.potrf_async_(a, info, handle=h)
    torch.cuda.synchronize()
    ref = a.clone()
    g = torch.cuda.CUDAGraph()
    a.copy_(a0)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        mi_llt.potrf_async_(a, info, handle=h)
    for rep in range(3):
        a.copy_(a0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(a, ref), f"replay {rep}"
    # eager on the default stream after the capture (the handle switches streams)
    a.copy_(a0)
    assert mi_llt.potrf_(a, handle=h) == 0
    assert torch.equal(a, ref)


def test_device_status_is_reported_and_cleared(dev):
    print("[gpu] injected device status", flush=True)
    n = 400
    a0 = mi_llt.colmajor(torch.from_numpy(workloads.gram_spd(n, seed=3)).to(dev))
    h = mi_llt.Handle()
    a = a0.clone()
    assert mi_llt.potrf_(a, handle=h) == 0
    ref = a.clone()
    # a status bit set on the stream: the factorization kernels skip their work and the
    # synchronous call reports the device failure instead of a success
    h.bind_stream().inject_status(mi_llt.STATUS_PANEL_TIMEOUT)
    a = a0.clone()
    with pytest.raises(mi_llt.MiLLTError) as ei:
        mi_llt.potrf_(a, handle=h)
    assert ei.value.code == mi_llt.E_DEVICE
    assert "panel solve timed out" in str(ei.value)
    assert torch.equal(a, a0)  # nothing was written
    # the status was read and cleared: the next call works
    a = a0.clone()
    assert mi_llt.potrf_(a, handle=h) == 0
    assert torch.equal(a, ref)
    # the asynchronous solve also skips its work while a bit is set; mi_llt_status returns it
    b0 = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, 3, seed=4)).to(dev))
    b = b0.clone()
    h.bind_stream().inject_status(mi_llt.STATUS_SOLVE_TIMEOUT)
    mi_llt.potrs_(ref, b, handle=h)
    assert h.status() == mi_llt.STATUS_SOLVE_TIMEOUT
    assert torch.equal(b, b0)
    mi_llt.potrs_(ref, b, handle=h)
    assert h.status() == 0
    x = np.linalg.solve(workloads.gram_spd(n, seed=3), b0.cpu().numpy())
    assert np.abs(b.cpu().numpy() - x).max() <= 1e-10 * np.abs(x).max()
    # the mixed-precision driver reads the status at each synchronisation
    h.bind_stream().inject_status(mi_llt.STATUS_FACTOR_TICKET)
    with pytest.raises(mi_llt.MiLLTError) as ei:
        mi_llt.posv_mixed(a0, b0, prec="fp32", handle=h)
    assert ei.value.code == mi_llt.E_DEVICE
    _, iters, berr, info = mi_llt.posv_mixed(a0, b0, prec="fp32", handle=h)
    assert info == 0 and berr <= 1e-14
