"""The seeded input-field generator (workload tool, not the hot path).

Pins what the construction fixes exactly or statistically: the r2c layout
and Hermitian kz=0 plane of a real field, the zero mean (the quadratic term
subtracts <dG^2>), the Kaiser factor mode by mode, the cut at k >= NN/3 of
the Gaussian field, and the shell-averaged power of the Gaussian field
against P(k) = A k^-1.5.
"""
import numpy as np
import pytest

import barbie_amd as B


def modes(f):
    return f.view(np.complex128)


def kgrid(NN):
    k = np.fft.fftfreq(NN, d=1.0 / NN)
    kz = np.arange(NN // 2 + 1, dtype=np.float64)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    return np.sqrt(k2), kz


def test_layout_seed_and_hermitian_plane():
    NN = 24
    f = B.make_field(NN, NRi=3)
    assert f.shape == (NN, NN, NN + 2) and f.dtype == np.float64
    assert np.array_equal(f, B.make_field(NN, NRi=3))          # seeded
    assert not np.array_equal(f, B.make_field(NN, NRi=4))
    D = modes(f)
    # the kz = 0 and kz = NN/2 planes of a real field: D(-kx,-ky) = conj D(kx,ky)
    idx = (-np.arange(NN)) % NN
    for kz in (0, NN // 2):
        P = D[:, :, kz]
        assert np.allclose(P[np.ix_(idx, idx)], np.conj(P), atol=1e-9 * np.abs(P).max())
    assert abs(D[0, 0, 0]) < 1e-9 * np.abs(D).max()             # zero mean


def test_kaiser_factor_mode_by_mode():
    NN = 24
    a = modes(B.make_field(NN, NRi=2, beta=0.0))
    b = modes(B.make_field(NN, NRi=2, beta=1.0))
    kk, kz = kgrid(NN)
    with np.errstate(divide="ignore", invalid="ignore"):
        mu2 = np.where(kk > 0, kz[None, None, :] ** 2 / kk ** 2, 0.0)
    assert np.allclose(b, a * (1.0 + mu2), rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_gaussian_field_power_and_cut():
    NN = 48
    D = modes(B.make_field(NN, NRi=1, fnl=0.0, beta=0.0))
    kk, _ = kgrid(NN)
    # modes at or above the cut, and k = 0, carry no power
    assert np.abs(D[(kk >= NN / 3) | (kk == 0)]).max() < 1e-9 * np.abs(D).max()
    # shell averages of |D|^2 / (V P(k)) for k in [3, NN/3): mean 1 within the
    # sampling error of the shell (relative std 1 / sqrt(modes))
    V = float(NN) ** 3
    ratios = []
    for lo in range(3, NN // 3 - 1):
        sel = (kk >= lo) & (kk < lo + 1)
        # each mode with 0 < kz < NN/2 stands for two (its conjugate)
        r = np.abs(D[sel]) ** 2 / (V * 1e3 * kk[sel] ** -1.5)
        ratios.append((r.mean(), r.size))
    for m, n in ratios:
        assert abs(m - 1.0) < 5.0 / np.sqrt(n) + 0.02, (m, n)
    # and over all shells together, tightly
    tot = sum(m * n for m, n in ratios) / sum(n for _, n in ratios)
    assert abs(tot - 1.0) < 0.05


def test_quadratic_term_moves_power_but_keeps_zero_mean():
    NN = 24
    g = modes(B.make_field(NN, NRi=5, fnl=0.0, beta=0.0))
    n = modes(B.make_field(NN, NRi=5, fnl=1.0, beta=0.0))
    kk, _ = kgrid(NN)
    assert abs(n[0, 0, 0]) < 1e-9 * np.abs(n).max()
    # the square adds power beyond the Gaussian cut
    assert np.abs(n[kk >= NN / 3]).max() > 1e-3 * np.abs(g).max()


@pytest.mark.gpu
def test_device_generator_statistics():
    import torch
    NN = 48
    D = torch.view_as_complex(B.make_field_device(NN, NRi=1, fnl=0.0, beta=0.0)
                              .reshape(NN, NN, NN // 2 + 1, 2)).cpu().numpy()
    kk, _ = kgrid(NN)
    assert np.abs(D[(kk >= NN / 3) | (kk == 0)]).max() < 1e-9 * np.abs(D).max()
    V = float(NN) ** 3
    sel = (kk >= 3) & (kk < NN / 3)
    r = np.abs(D[sel]) ** 2 / (V * 1e3 * kk[sel] ** -1.5)
    assert abs(r.mean() - 1.0) < 0.05
