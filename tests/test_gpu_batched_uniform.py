"""Batched systems of one order (mi_dpotrf_batched_async, mi_dpotrs_batched and the batched
adjoints): every system's result must be bit-for-bit the single-system entry point's (the same
kernels, each launch covering the whole batch), failing systems report their own info, and the
batched Gaussian-process gradient matches the closed form 1/2 (a a^T - K^{-1})."""
import ctypes

import numpy as np
import pytest
import torch

import mi_llt
import oracle
from mi_llt import autograd as mad
from mi_llt import workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def stack(n, batch, seed, ld=None, pad=0):
    """(batch, n, n) SPD stack, column-major per matrix with leading dimension ld and pad extra
    elements between systems (a non-trivial stride)."""
    ld = ld or n
    mats = [workloads.gram_spd(n, seed=seed + z) for z in range(batch)]
    stride = ld * n + pad
    buf = torch.zeros(batch * stride + 1, dtype=torch.float64, device="cuda")
    a = torch.as_strided(buf, (batch, n, n), (stride, 1, ld))
    a.copy_(torch.from_numpy(np.stack(mats)).cuda())
    return a, mats


@pytest.mark.parametrize("n,batch,nrhs", [(1, 3, 1), (100, 5, 2), (129, 4, 3), (300, 7, 16), (512, 6, 1),
                                          (640, 3, 17), (1000, 2, 5), (2049, 2, 3)])
def test_batched_factor_and_solve_match_single(dev, n, batch, nrhs):
    say(f"[gpu] batched potrf/potrs n={n} batch={batch} nrhs={nrhs}")
    a, mats = stack(n, batch, seed=10 * n)
    b = mi_llt.colmajor_batched(torch.from_numpy(np.stack([workloads.rhs(n, nrhs, seed=7 + z)
                                                           for z in range(batch)])).to(dev))
    info = torch.full((batch,), 99, dtype=torch.int32, device=dev)
    h = mi_llt.Handle().bind_stream()
    mi_llt.potrf_batched_async_(a, info, handle=h)
    x = b.clone()
    mi_llt.potrs_batched_(a, x, handle=h)
    torch.cuda.synchronize()
    assert h.status() == 0
    assert info.tolist() == [0] * batch
    h1 = mi_llt.Handle().bind_stream()
    for z in range(batch):
        l1 = mi_llt.colmajor(torch.from_numpy(mats[z]).to(dev))
        assert mi_llt.potrf_(l1, handle=h1) == 0
        x1 = mi_llt.colmajor(b[z].clone())
        mi_llt.potrs_(l1, x1, handle=h1)
        torch.cuda.synchronize()
        assert torch.equal(torch.tril(a[z]), torch.tril(l1)), f"factor of system {z} differs"
        assert torch.equal(torch.triu(a[z], 1), torch.triu(torch.from_numpy(mats[z]).to(dev), 1)), "upper touched"
        assert torch.equal(x[z], x1), f"solution of system {z} differs"
    # and against the oracle for one system
    lo, _ = oracle.llt_blocked(mats[batch - 1])
    np.testing.assert_allclose(np.tril(a[batch - 1].cpu().numpy()), np.tril(lo), rtol=0, atol=1e-12)


def test_batched_leading_dimension_stride_and_failures(dev):
    n, batch = 257, 5
    say(f"[gpu] batched potrf n={n} batch={batch} lda={n + 3} padded stride, failing systems 1 and 3")
    a, mats = stack(n, batch, seed=3, ld=n + 3, pad=11)
    a[1, 200, 200] = -1.0    # system 1 fails at column 200
    a[3, 0, 0] = 0.0         # system 3 fails at column 0
    info = torch.zeros(batch, dtype=torch.int32, device=dev)
    mi_llt.potrf_batched_async_(a, info)
    torch.cuda.synchronize()
    assert info.tolist() == [0, 201, 0, 1, 0]
    for z in (0, 2, 4):
        lo, _ = oracle.llt_blocked(mats[z])
        np.testing.assert_allclose(np.tril(a[z].cpu().numpy()), np.tril(lo), rtol=0, atol=1e-12)


def test_batched_adjoints_match_single(dev):
    n, batch, r = 384, 4, 2
    say(f"[gpu] batched adjoints n={n} batch={batch} r={r}")
    a, mats = stack(n, batch, seed=21)
    info = torch.zeros(batch, dtype=torch.int32, device=dev)
    mi_llt.potrf_batched_async_(a, info)
    l = mi_llt.colmajor_batched(torch.tril(a))
    g = torch.Generator(device="cpu").manual_seed(5)
    lbar = mi_llt.colmajor_batched(torch.tril(torch.randn(batch, n, n, generator=g, dtype=torch.float64)).to(dev))
    xbar = mi_llt.colmajor_batched(torch.randn(batch, n, r, generator=g, dtype=torch.float64).to(dev))
    x = mi_llt.colmajor_batched(torch.randn(batch, n, r, generator=g, dtype=torch.float64).to(dev))
    h = mi_llt.default_handle()
    L = mi_llt.lib()
    vp = ctypes.c_void_p
    gb = lbar.clone()
    rc = L.mi_dpotrf_adjoint_batched(h.ptr, b"L", n, batch, vp(l.data_ptr()), n, n * n, vp(gb.data_ptr()), n, n * n)
    assert rc == 0
    xb = xbar.clone()
    lb = mi_llt.colmajor_batched(torch.zeros(batch, n, n, dtype=torch.float64, device=dev))
    rc = L.mi_dpotrs_adjoint_batched(h.ptr, b"L", n, r, batch, vp(l.data_ptr()), n, n * n, vp(x.data_ptr()), n, n * r,
                                     vp(xb.data_ptr()), n, n * r, vp(lb.data_ptr()), n, n * n)
    assert rc == 0
    torch.cuda.synchronize()
    for z in range(batch):
        g1 = mi_llt.colmajor(lbar[z].clone())
        assert L.mi_dpotrf_adjoint(h.ptr, b"L", n, vp(l[z].data_ptr()), n, vp(g1.data_ptr()), n) == 0
        xb1 = mi_llt.colmajor(xbar[z].clone())
        lb1 = mi_llt.colmajor(torch.zeros(n, n, dtype=torch.float64, device=dev))
        assert L.mi_dpotrs_adjoint(h.ptr, b"L", n, r, vp(l[z].data_ptr()), n, vp(x[z].data_ptr()), n,
                                   vp(xb1.data_ptr()), n, vp(lb1.data_ptr()), n) == 0
        torch.cuda.synchronize()
        assert torch.equal(gb[z], g1), f"factor adjoint of system {z} differs"
        assert torch.equal(xb[z], xb1) and torch.equal(torch.tril(lb[z]), torch.tril(lb1)), f"solve adjoint {z}"
    assert h.status() == 0


@pytest.mark.parametrize("n,batch,nrhs", [(96, 3, 1), (512, 5, 1), (700, 2, 4)])
def test_batched_posv_and_its_adjoint_match_single(dev, n, batch, nrhs):
    say(f"[gpu] batched posv + posv adjoint n={n} batch={batch} nrhs={nrhs}")
    a, mats = stack(n, batch, seed=40 + n)
    b = mi_llt.colmajor_batched(torch.from_numpy(np.stack([workloads.rhs(n, nrhs, seed=70 + z)
                                                           for z in range(batch)])).to(dev))
    info = torch.full((batch,), 5, dtype=torch.int32, device=dev)
    x = b.clone()
    mi_llt.posv_batched_async_(a, x, info)
    g = torch.Generator(device="cpu").manual_seed(9)
    xbar = mi_llt.colmajor_batched(torch.randn(batch, n, nrhs, generator=g, dtype=torch.float64).to(dev))
    lbar0 = mi_llt.colmajor_batched(torch.tril(torch.randn(batch, n, n, generator=g, dtype=torch.float64)).to(dev))
    xb, lb = xbar.clone(), lbar0.clone()
    h = mi_llt.default_handle()
    L = mi_llt.lib()
    vp = ctypes.c_void_p
    assert L.mi_dposv_adjoint_batched(h.ptr, b"L", n, nrhs, batch, vp(a.data_ptr()), n, n * n, vp(x.data_ptr()), n,
                                      n * nrhs, vp(xb.data_ptr()), n, n * nrhs, vp(lb.data_ptr()), n, n * n) == 0
    torch.cuda.synchronize()
    assert info.tolist() == [0] * batch and h.status() == 0
    for z in range(batch):
        a1 = mi_llt.colmajor(torch.from_numpy(mats[z]).to(dev))
        x1 = mi_llt.colmajor(b[z].clone())
        i1 = torch.zeros(1, dtype=torch.int32, device=dev)
        mi_llt.posv_async_(a1, x1, i1)
        torch.cuda.synchronize()
        assert torch.equal(torch.tril(a[z]), torch.tril(a1)) and torch.equal(x[z], x1), f"posv of system {z}"
        # single-system reverse mode: the solve's adjoint, then the factor's on the accumulated Lbar
        xb1 = mi_llt.colmajor(xbar[z].clone())
        lb1 = mi_llt.colmajor(lbar0[z].clone())
        assert L.mi_dpotrs_adjoint(h.ptr, b"L", n, nrhs, vp(a1.data_ptr()), n, vp(x1.data_ptr()), n,
                                   vp(xb1.data_ptr()), n, vp(lb1.data_ptr()), n) == 0
        assert L.mi_dpotrf_adjoint(h.ptr, b"L", n, vp(a1.data_ptr()), n, vp(lb1.data_ptr()), n) == 0
        torch.cuda.synchronize()
        assert torch.equal(xb[z], xb1), f"Bbar of system {z}"
        assert torch.equal(lb[z], lb1), f"Abar of system {z}"


def test_batched_gp_gradient_closed_form_and_single(dev):
    n, batch = 200, 6
    say(f"[gpu] batched GP gradient n={n} batch={batch}")
    ks, ys = zip(*[workloads.gp_kernel_torch(n, seed=300 + z, device=dev) for z in range(batch)])
    k = torch.stack(ks).requires_grad_(True)
    y = torch.stack(ys)
    ll = mad.gp_log_marginal_batched(k, y)
    ll.sum().backward()
    for z in range(batch):
        kz = ks[z]
        kinv = torch.linalg.inv(kz)
        alpha = kinv @ ys[z]
        ref = 0.5 * (torch.outer(alpha, alpha) - kinv)
        err = float((k.grad[z] - ref).abs().max() / ref.abs().max())
        assert err <= 1e-9, (z, err)
        k1 = kz.clone().requires_grad_(True)
        l1 = mad.gp_log_marginal(k1, ys[z])
        l1.backward()
        assert abs(float(ll[z]) - float(l1)) <= 1e-12 * abs(float(l1))
        dif = float((k.grad[z] - k1.grad).abs().max() / k1.grad.abs().max())
        assert dif <= 1e-13, f"system {z}: batched gradient differs from the single path by {dif}"


def test_batched_argument_errors(dev):
    h = mi_llt.default_handle()
    L = mi_llt.lib()
    vp = ctypes.c_void_p
    a = torch.zeros(2, 8, 8, dtype=torch.float64, device=dev)
    d = torch.zeros(2, dtype=torch.int32, device=dev)
    p, q = vp(a.data_ptr()), vp(d.data_ptr())
    assert L.mi_dpotrf_batched_async(h.ptr, b"U", 8, 2, p, 8, 64, q) == -1004
    assert L.mi_dpotrf_batched_async(h.ptr, b"L", 8, -1, p, 8, 64, q) == -4
    assert L.mi_dpotrf_batched_async(h.ptr, b"L", 8, 2, p, 7, 64, q) == -6
    assert L.mi_dpotrf_batched_async(h.ptr, b"L", 8, 2, p, 8, 63, q) == -7  # overlapping systems
    assert L.mi_dpotrf_batched_async(h.ptr, b"L", 8, 2, p, 8, 64, vp(None)) == -8
    assert L.mi_dpotrf_batched_async(h.ptr, b"L", 8, 0, p, 8, 64, q) == 0
    torch.cuda.synchronize()
