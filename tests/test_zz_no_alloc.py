"""mi_dposv_noalloc (factor + solve with the scratch in a fixed device array) against
mi_dposv_async (the handle entry point running the same kernels): identical results on two
calls in a row, the second smaller, then argument and size errors."""
import ctypes

import numpy as np
import pytest
import torch

import mi_llt
from mi_llt import workloads

E_UNSUPPORTED, E_TOO_LARGE = -1004, -1006


def _noalloc():
    L = mi_llt.lib()
    f = L.mi_dposv_noalloc
    f.argtypes = [ctypes.c_char, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                  ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def _inputs(n, r, seed, dev):
    a = mi_llt.colmajor(torch.from_numpy(workloads.gram_spd(n, seed=seed)).to(dev))
    b = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, r, seed=seed + 1)).to(dev))
    return a, b


def _noalloc_call(f, a, b, dev):
    n, r = int(a.shape[0]), int(b.shape[1])
    info = torch.full((1,), 77, dtype=torch.int32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = f(b"L", n, r, ctypes.c_void_p(a.data_ptr()), a.stride(1), ctypes.c_void_p(b.data_ptr()), b.stride(1),
           ctypes.c_void_p(info.data_ptr()), stream)
    assert rc == 0, rc
    return info


@pytest.mark.gpu
def test_noalloc_matches_handle_entry_point_twice(dev):
    f = _noalloc()
    h = mi_llt.Handle().bind_stream()
    for n, r, seed in ((1000, 16, 3), (300, 5, 11)):  # the second call smaller than the first
        print(f"noalloc vs dposv_async: n={n} nrhs={r} seed={seed}", flush=True)
        a, b = _inputs(n, r, seed, dev)
        a2, b2 = a.clone(), b.clone()
        info = _noalloc_call(f, a, b, dev)
        info2 = torch.zeros(1, dtype=torch.int32, device=dev)
        mi_llt.posv_async_(a2, b2, info2, handle=h)
        torch.cuda.synchronize()
        assert int(info.item()) == 0 and int(info2.item()) == 0
        assert torch.equal(torch.tril(a), torch.tril(a2)), "factor differs from mi_dposv_async"
        assert torch.equal(b, b2), "solution differs from mi_dposv_async"
        assert torch.equal(torch.triu(a, 1), torch.triu(a2, 1)), "upper triangle touched"
        an = a2.new_tensor(workloads.gram_spd(n, seed=seed))
        x = b.cpu().numpy()
        bn = workloads.rhs(n, r, seed=seed + 1)
        ah = an.cpu().numpy()
        berr = np.abs(ah @ x - bn).max() / (np.abs(ah).sum(1).max() * np.abs(x).max())
        assert berr < 1e-14, berr


@pytest.mark.gpu
def test_noalloc_failing_pivot_and_errors(dev):
    f = _noalloc()
    n, r = 257, 3
    a, b = _inputs(n, r, 5, dev)
    a[200, 200] = -1.0  # column 200 fails
    b0 = b.clone()
    info = _noalloc_call(f, a, b, dev)
    torch.cuda.synchronize()
    assert int(info.item()) == 201
    assert torch.equal(b, b0), "B must be unchanged after a failing factor"
    vp = ctypes.c_void_p
    d = torch.zeros(1, dtype=torch.int32, device=dev)
    s = vp(torch.cuda.current_stream().cuda_stream)
    p = vp(a.data_ptr())
    assert f(b"U", n, r, p, n, vp(b.data_ptr()), n, vp(d.data_ptr()), s) == E_UNSUPPORTED
    assert f(b"X", n, r, p, n, vp(b.data_ptr()), n, vp(d.data_ptr()), s) == -1
    assert f(b"L", n, r, p, n - 1, vp(b.data_ptr()), n, vp(d.data_ptr()), s) == -5
    assert f(b"L", n, r, p, n, vp(b.data_ptr()), n, vp(None), s) == -8
    # beyond the fixed scratch: rejected before anything is queued (the pointers are never used)
    assert f(b"L", 8192, 113, p, 8192, vp(b.data_ptr()), 8192, vp(d.data_ptr()), s) == E_TOO_LARGE
    assert f(b"L", 145664, 0, p, 145664, vp(b.data_ptr()), 145664, vp(d.data_ptr()), s) == E_TOO_LARGE
    torch.cuda.synchronize()
