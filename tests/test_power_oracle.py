"""CPU oracle of the power-spectrum multipoles (calpow_mom, funcs.c:860-1097).

Pins from the mathematics, not from the loops: the visited modes are one of
each Hermitian pair inside the Nyquist sphere (their count is half the
lattice count); the estimator is a sum over modes, so it is additive over
fields with disjoint mode supports and quadratic in the amplitude; for a
Gaussian field P0 follows the input P(k), and for a Kaiser field the P2/P0
and P4/P0 ratios follow the Legendre expansion of (1 + beta mu^2)^2.
"""
import numpy as np

import oracle_lib as O
import barbie_amd as B


def lattice_pairs(NN):
    h = NN // 2
    k = np.arange(-h + 1, h + 1)
    n = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    return int(np.count_nonzero((n > 0) & (n < h * h))) // 2


def test_visited_modes_are_half_the_nyquist_sphere():
    for NN in (16, 24, 48):
        ones = np.zeros((NN, NN, NN + 2))
        ones[..., 0::2] = 1.0
        o = O.calpow_mom(ones, NN)
        assert int(o[:, 5].sum()) == lattice_pairs(NN)
        # |Delta|^2 = 1 everywhere: P0 = 1 / (N^3 LL^3) in every populated bin
        pop = o[:, 5] > 0
        assert np.allclose(o[pop, 1], 1.0 / (NN ** 3 * 0.56 ** 3), rtol=1e-12)


def test_additive_over_disjoint_supports_and_quadratic():
    NN = 24
    f = B.make_field(NN, NRi=1)
    H = NN // 2 + 1
    c = f.view(np.complex128).reshape(NN, NN, H)
    k = np.arange(H)
    lo = np.zeros_like(f)
    hi = np.zeros_like(f)
    lo.view(np.complex128).reshape(NN, NN, H)[...] = np.where(k[None, None, :] < 4, c, 0)
    hi.view(np.complex128).reshape(NN, NN, H)[...] = np.where(k[None, None, :] >= 4, c, 0)
    a, b, ab = O.calpow_mom(lo, NN), O.calpow_mom(hi, NN), O.calpow_mom(f, NN)
    # the mode counts and mean |k| do not depend on the field; the weighted
    # sums add over the two supports
    assert np.array_equal(ab[:, 5], a[:, 5]) and np.array_equal(ab[:, 5], b[:, 5])
    num = ab[:, 5]
    pop = num > 0
    assert np.allclose(ab[pop, 1], a[pop, 1] + b[pop, 1], rtol=1e-12)
    assert np.allclose(ab[pop, 2:5], a[pop, 2:5] + b[pop, 2:5], rtol=1e-10,
                       atol=1e-12 * np.abs(ab[pop, 1:5]).max())
    s = O.calpow_mom(3.0 * f, NN)
    assert np.allclose(s[:, 1:5], 9.0 * ab[:, 1:5], rtol=1e-12, atol=1e-12 * np.abs(ab).max())
    assert np.array_equal(s[:, 5], ab[:, 5]) and np.array_equal(s[:, 0], ab[:, 0])


def test_gaussian_monopole_follows_the_input_spectrum():
    NN = 64
    f = B.make_field(NN, NRi=3, fnl=0.0, beta=0.0)
    o = O.calpow_mom(f, NN)
    LL = 0.56
    # the bins' mean of P(k) = 1e3 k^-1.5 over the lattice modes they hold
    h = NN // 2
    k = np.arange(-h + 1, h + 1)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2).ravel()
    kk = kk[(kk > 0) & (kk < h)]
    scale = np.log10(0.5 * NN) / 10
    d = np.floor(np.log10(kk) / scale).astype(int)
    checked = 0
    for b, (kmean, P0, P2, P4, P6, num) in enumerate(o):
        kb = kk[d == b]
        if num < 200 or kb.max() > NN / 3 - 1:
            continue
        expect = np.mean(1e3 * kb ** -1.5) / LL ** 3
        assert abs(P0 / expect - 1.0) < 5.0 / np.sqrt(num) + 0.02, (b, P0, expect)
        checked += 1
    assert checked >= 2


def test_kaiser_multipole_ratios():
    NN = 64
    beta = 1.0
    f = B.make_field(NN, NRi=4, fnl=0.0, beta=beta)
    o = O.calpow_mom(f, NN)
    a0 = 1 + 2 * beta / 3 + beta ** 2 / 5
    a2 = 4 * beta / 3 + 4 * beta ** 2 / 7
    a4 = 8 * beta ** 2 / 35
    rows = [r for r in o if r[5] > 2000 and r[0] * 0.56 * NN / (2 * np.pi) < NN / 3 - 2]
    assert rows
    r2 = np.mean([r[2] / r[1] for r in rows])
    r4 = np.mean([r[3] / r[1] for r in rows])
    assert abs(r2 - a2 / a0) < 0.1, r2
    assert abs(r4 - a4 / a0) < 0.05, r4
