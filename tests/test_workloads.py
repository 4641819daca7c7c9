"""Synthetic inputs: the two constructions of the planted-spectrum matrix agree, and the
planted spectrum is what the matrix has (CPU)."""
import numpy as np
import torch

from mi_llt import workloads


def test_spectrum_spd_two_sided_equals_dense_product():
    n, kappa = 96, 1e6
    a_dense = workloads.spectrum_spd(n, kappa, seed=11, n_reflectors=16)
    a_two = workloads.spectrum_spd_torch(n, kappa, seed=11, n_reflectors=16, device="cpu").numpy()
    assert np.abs(a_dense - a_two).max() <= 1e-12 * kappa


def test_spectrum_spd_has_the_planted_eigenvalues():
    n, kappa = 80, 1e4
    a = workloads.spectrum_spd_torch(n, kappa, seed=2, n_reflectors=64, device="cpu").numpy()
    lam = np.sort(workloads.spectrum(n, kappa, seed=2))
    ev = np.linalg.eigvalsh(a)
    assert np.allclose(ev, lam, rtol=1e-10, atol=1e-10)
    assert abs(ev[0] - 1.0) < 1e-10 and abs(ev[-1] - kappa) < 1e-8 * kappa


def test_gram_batch_scatter_equals_per_system_construction():
    """The batched workload's inputs: one scatter per size class gives exactly the values of
    building every system one by one from the same generator draws."""
    import torch

    from mi_llt import batched

    vb = batched.gram_batch(40, seed=5, device="cpu")
    rng = np.random.default_rng(5)
    ns = rng.integers(8, 257, size=40)
    rs = rng.integers(1, 9, size=40)
    ref = batched.VBatch(ns, rs, device="cpu")
    gen = torch.Generator(device="cpu")
    gen.manual_seed(5)
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        n = int(n)
        g = torch.randn(len(idx), n, n, dtype=torch.float64, generator=gen)
        a = torch.matmul(g, g.transpose(1, 2).contiguous()) / n
        a.diagonal(dim1=1, dim2=2).add_(0.1)
        bv = torch.randn(len(idx), 8, n, dtype=torch.float64, generator=gen)
        for i, k in enumerate(idx):
            o = int(ref.a_off[k])
            ref.a_flat[o:o + n * n] = a[i].reshape(-1)
            o, r = int(ref.b_off[k]), int(rs[k])
            ref.b_flat[o:o + n * r] = bv[i, :r].reshape(-1)
    assert torch.equal(vb.a_flat, ref.a_flat) and torch.equal(vb.b_flat, ref.b_flat)
