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
