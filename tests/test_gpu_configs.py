"""The five stated workloads, each run through the library at its stated size and checked:

1. n = 1024, 16 right-hand sides, A = B B^T / n + 0.1 I from seed 0: L and X element-wise
   against the oracle and against the genuine Eigen LLT (oracle/_ref), backward errors <= 1e-14.
2. n = 8192, 64 right-hand sides (the bench workload): in test_gpu_llt.py
   (test_full_size_backward_error).
3. n = 16384, A = Q diag(lambda) Q^T, kappa in {1e2, 1e4, 1e6}, 16 right-hand sides, the
   mixed-precision solve with FP64 refinement for FP32, split-BF16 and 3-part split-BF16
   trailing updates: FP64 backward error <= 1e-14 recomputed independently (torch FP64 products),
   and the FP64 factor's logdet = sum log lambda_i (the planted spectrum).
4. 20,000 systems, n ~ U{8..256}, 1..8 right-hand sides, FP32: every system's factor and a
   sample of every size class's solves within FP32 bounds of the FP64 arithmetic (the FP64
   variant is test_vbatched_full_workload_backward_error in test_gpu_batched.py).
5. The Gaussian-process gradient dL/dK for a batch of 256 kernels of n = 512, computed in one
   batched forward + backward (every kernel launch covers all 256 systems): every gradient
   against the closed form 1/2 (a a^T - K^{-1}) from torch's FP64 inverse (an independent
   library routine, used as a checker only), and a sample against the same gradient computed
   by Eigen's LLT on the CPU (oracle/_ref).  (The single n = 4096 case: test_gpu_adjoint.py.)
"""
import numpy as np
import pytest
import torch

import mi_llt
import oracle
from mi_llt import autograd as mad
from mi_llt import batched, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def sym(a):
    return torch.tril(a) + torch.tril(a, -1).t()


def berr_fp64(a, x, b):
    asym = sym(a)
    r = asym @ x - b
    return float((r.norm(dim=0) / (torch.linalg.matrix_norm(asym) * x.norm(dim=0))).max())


def test_config1_n1024_seed0(dev):
    say("[gpu] config 1: n=1024, 16 RHS, seed 0")
    n, r = 1024, 16
    a = workloads.gram_spd(n, seed=0)
    b = workloads.rhs(n, r, seed=1)
    da = mi_llt.colmajor(torch.from_numpy(a).to(dev))
    db = mi_llt.colmajor(torch.from_numpy(b).to(dev))
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    mi_llt.posv_async_(da, db, info)
    torch.cuda.synchronize()
    assert int(info.item()) == 0 and mi_llt.default_handle().status() == 0
    m, x = da.cpu().numpy(), db.cpu().numpy()
    lo, idx = oracle.llt_blocked(a)
    assert idx == -1
    np.testing.assert_allclose(np.tril(m), np.tril(lo), rtol=0, atol=1e-12)
    np.testing.assert_array_equal(np.triu(m, 1), np.triu(a, 1))
    xo = oracle.llt_solve(np.tril(lo), b)
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-10 * np.abs(xo).max())
    assert oracle.factor_berr(a, np.tril(m)) <= 1e-14
    assert oracle.solve_berr(a, x, b) <= 1e-14
    if oracle.have_eigen():
        e = oracle.eigen_llt(a, b)
        assert e["info"] == 0
        np.testing.assert_allclose(np.tril(m), np.tril(e["L"]), rtol=0, atol=1e-12)
        np.testing.assert_allclose(x, e["X"], rtol=0, atol=1e-10 * np.abs(e["X"]).max())


@pytest.fixture(scope="module")
def config3(dev):
    """The n = 16384 right-hand sides (seed 1) and one matrix per kappa (built lazily)."""
    n, r = 16384, 16
    b = mi_llt.colmajor(torch.from_numpy(workloads.rhs(n, r, seed=1)).to(dev))
    cache = {}

    def get(kappa):
        if kappa not in cache:
            cache.clear()
            torch.cuda.empty_cache()
            cache[kappa] = workloads.spectrum_spd_torch(n, kappa, seed=0, device=dev)
        return cache[kappa]

    return n, b, get


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16x6"])
def test_config3_mixed_n16384(dev, config3, kappa, prec):
    n, b, get = config3
    say(f"[gpu] config 3: n={n} kappa={kappa:g} {prec}")
    a = get(kappa)
    x, iters, berr, info = mi_llt.posv_mixed(a, b, prec=prec, tol=1e-14, maxit=30)
    assert info == 0
    assert berr <= 1e-14
    assert berr_fp64(a, x, b) <= 2e-14  # recomputed here, with torch's FP64 products
    if iters >= 0:  # refinement converged without the FP64 fallback
        assert 1 <= iters <= 30


@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_config3_fp64_logdet(dev, config3, kappa):
    n, _, get = config3
    say(f"[gpu] config 3: FP64 logdet n={n} kappa={kappa:g}")
    l = get(kappa).clone()
    assert mi_llt.potrf_(l) == 0
    logdet = 2.0 * float(torch.log(torch.diagonal(l)).sum())
    ref = float(np.log(workloads.spectrum(n, kappa, seed=0)).sum())
    assert abs(logdet - ref) <= 1e-10 * abs(ref)


def test_config4_fp32_20000_systems(dev):
    say("[gpu] config 4: 20,000 systems FP32")
    vb = batched.gram_batch(20000, seed=0, dtype=torch.float32, device=dev)
    a0, b0 = vb.a_flat.clone(), vb.b_flat.clone()
    info = batched.posv_vbatched_(vb)
    assert int((info != 0).sum()) == 0
    worst_f, worst_s, checked = 0.0, 0.0, 0
    for n in np.unique(vb.ns):
        idx = np.nonzero(vb.ns == n)[0]
        n = int(n)
        ao = torch.from_numpy(vb.a_off[idx]).to(dev)
        pos = ao[:, None] + torch.arange(n * n, device=dev)[None, :]
        A = a0[pos].view(-1, n, n).transpose(1, 2).double()
        A = torch.tril(A) + torch.tril(A, -1).transpose(1, 2)
        Lf = torch.tril(vb.a_flat[pos].view(-1, n, n).transpose(1, 2)).double()
        rf = torch.linalg.matrix_norm(A - Lf @ Lf.transpose(1, 2).contiguous()) / torch.linalg.matrix_norm(A)
        worst_f = max(worst_f, float(rf.max()))
        checked += len(idx)
        for k in idx[:8]:
            r = int(vb.rs[k])
            o = int(vb.b_off[k])
            Ak = a0[int(vb.a_off[k]):int(vb.a_off[k]) + n * n].view(n, n).t().double()
            Ak = torch.tril(Ak) + torch.tril(Ak, -1).t()
            Bk = b0[o:o + n * r].view(r, n).t().double()
            Xk = vb.b_flat[o:o + n * r].view(r, n).t().double()
            be = float(((Ak @ Xk - Bk).norm(dim=0) / (torch.linalg.matrix_norm(Ak) * Xk.norm(dim=0))).max())
            worst_s = max(worst_s, be)
    assert checked == 20000
    # FP32 storage and arithmetic: backward errors a small multiple of u = 6e-8
    assert worst_f <= 2e-6, worst_f
    assert worst_s <= 2e-6, worst_s


def test_config5_gp_gradient_batch_256_n512(dev):
    say("[gpu] config 5: 256 GP gradients, n=512, one batched forward + backward")
    n, batch = 512, 256
    ks, ys = zip(*[workloads.gp_kernel_torch(n, seed=100 + i, device=dev) for i in range(batch)])
    k = torch.stack(ks).requires_grad_(True)
    y = torch.stack(ys)
    mad.gp_log_marginal_batched(k, y).sum().backward()
    worst = 0.0
    grads = []
    for i in range(batch):
        kinv = torch.linalg.inv(sym(ks[i]))
        alpha = kinv @ ys[i]
        ref = 0.5 * (torch.outer(alpha, alpha) - kinv)
        worst = max(worst, float((k.grad[i] - ref).abs().max() / ref.abs().max()))
        if i in (0, 85, 170, 255):
            grads.append((i, ks[i].cpu().numpy(), ys[i].cpu().numpy(), k.grad[i].cpu().numpy()))
    assert worst <= 1e-9, worst
    if oracle.have_eigen():
        for i, kk, yy, g in grads:  # the gradient computed by Eigen's LLT on the CPU
            ge = oracle.eigen_gp_grad(kk, yy)
            assert np.abs(g - ge).max() <= 1e-9 * np.abs(ge).max(), i
