"""CPU pins of the generator's restatement (oracle/oracle_field.c), which the
GPU generator (bb_generate_field) must reproduce:

* Philox4x64-10 equals numpy.random.Philox's raw stream bit for bit
  (numpy's bit generator increments its 256-bit counter before each block,
  so block c is numpy's first block from counter c - 1);
* the own logarithm and sine/cosine stay within a few ulp of the C
  library's over their whole argument ranges;
* the Box-Muller pairs are standard normal (moments and a two-sided
  Kolmogorov-Smirnov test);
* the field: the Gaussian modes are Hermitian on the kz = 0 and NN/2 planes,
  their power follows V P(k) / 2 per component (chi-square bound), the
  field without non-Gaussianity and RSD returns its modes through the
  c2r / r2c round trip, and the RSD factor matches (1 + beta mu^2).
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle_lib as O


def numpy_block(ctr, key):
    """Philox4x64-10 of counter ctr (4 words) and key (2 words) via numpy."""
    c = sum(int(w) << (64 * i) for i, w in enumerate(ctr))
    c = (c - 1) % (1 << 256)
    k = sum(int(w) << (64 * i) for i, w in enumerate(key))
    bg = np.random.Philox(counter=c, key=k)
    return [int(x) for x in bg.random_raw(4)]


@pytest.mark.parametrize("seed", [0, 1, 100020 + 146 * 7, (1 << 64) - 1])
def test_philox_equals_numpy(seed):
    rng = np.random.default_rng(3)
    ctrs = [[0, 0, 0, 0], [1, 0, 0, 0], [123456789, 0, 0, 0], [(1 << 64) - 1, 0, 0, 0]]
    ctrs += [[int(x) for x in rng.integers(0, 1 << 63, 4)] for _ in range(8)]
    for ctr in ctrs:
        key = [seed, 0]
        assert O.philox4x64(ctr, key) == numpy_block(ctr, key), (ctr, seed)
    key = [int(x) for x in rng.integers(0, 1 << 63, 2)]
    assert O.philox4x64(ctrs[5], key) == numpy_block(ctrs[5], key)


def ulp_diff(a, b):
    return abs(a - b) / math.ulp(max(abs(b), 1e-300))


def test_own_log_within_ulps():
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(0, 1, 20000), 10.0 ** rng.uniform(-16, 0, 5000),
                         [1.0, 0.5, 2.0 ** -53, 0.7071067811865476, 0.70710678118654746]])
    worst = max(ulp_diff(O.gen_log(x), math.log(x)) for x in xs if x > 0)
    assert worst <= 4, worst
    assert O.gen_log(1.0) == 0.0


def test_own_sincos_within_ulps():
    rng = np.random.default_rng(6)
    us = np.concatenate([rng.uniform(0, 1, 20000), np.arange(0, 1, 1 / 64)])
    worst = 0.0
    two_pi = 2 * np.arccos(np.longdouble(-1.0))   # extended precision reference
    for u in us:
        s, c = O.gen_sincos_2pi(u)
        a = two_pi * np.longdouble(u)
        worst = max(worst, float(abs(s - np.sin(a))), float(abs(c - np.cos(a))))
        assert s * s + c * c == pytest.approx(1.0, abs=4e-16)
    assert worst <= 3e-16, worst
    assert O.gen_sincos_2pi(0.0) == (0.0, 1.0)
    assert O.gen_sincos_2pi(0.25) == (1.0, 0.0) or O.gen_sincos_2pi(0.25)[0] == 1.0


def test_normal_pairs_are_standard_normal():
    g = np.array([O.gen_normal_pair(c, 1234) for c in range(40000)]).ravel()
    assert abs(g.mean()) < 4 / math.sqrt(g.size)
    assert abs(g.var() - 1) < 4 * math.sqrt(2 / g.size)
    assert abs(stats.skew(g)) < 4 * math.sqrt(6 / g.size)
    assert abs(stats.kurtosis(g)) < 4 * math.sqrt(24 / g.size)
    assert stats.kstest(g, "norm").pvalue > 1e-3
    # the two members of a pair are uncorrelated
    p = g.reshape(-1, 2)
    assert abs(np.corrcoef(p[:, 0], p[:, 1])[0, 1]) < 4 / math.sqrt(len(p))


@pytest.mark.parametrize("NN", [16, 24])
def test_modes_are_hermitian_and_follow_the_power(NN):
    m = O.generate_field(NN, 77, stage=1).view(np.complex128)   # [NN][NN][H]
    H = NN // 2 + 1
    for k in (0, NN // 2):
        pl = m[:, :, k]
        mir = pl[(-np.arange(NN)) % NN][:, (-np.arange(NN)) % NN]
        assert np.array_equal(pl, np.conj(mir))
    assert m[0, 0, 0] == 0
    # per-mode power against V P(k) (V = NN^3): chi-square of the complex modes
    i = np.fft.fftfreq(NN, 1.0 / NN)
    kz = np.arange(H)
    kk = np.sqrt(i[:, None, None] ** 2 + i[None, :, None] ** 2 + kz[None, None, :] ** 2)
    inside = (kk > 0) & (kk < NN / 3)
    pk = np.where(inside, 1e3 * np.where(kk > 0, kk, 1) ** -1.5, 0)
    assert np.all(m[~inside] == 0)
    cplx = inside & (kz[None, None, :] > 0) & (kz[None, None, :] < NN // 2)
    ratio = np.abs(m[cplx]) ** 2 / (NN ** 3 * pk[cplx])    # each ~ Exp(1)
    n = ratio.size
    assert abs(ratio.mean() - 1) < 5 / math.sqrt(n)


def test_round_trip_and_rsd():
    NN = 24
    g = O.generate_field(NN, 9, stage=1, fnl=0.0, beta=0.0).view(np.complex128)
    f = O.generate_field(NN, 9, fnl=0.0, beta=0.0).view(np.complex128)
    # without non-Gaussianity and RSD the field returns its modes (c2r / V, r2c)
    assert np.max(np.abs(f - g)) <= 1e-12 * np.max(np.abs(g))
    r = O.generate_field(NN, 9, fnl=0.0, beta=0.5).view(np.complex128)
    H = NN // 2 + 1
    i = np.fft.fftfreq(NN, 1.0 / NN)
    kz = np.arange(H)
    kk = np.sqrt(i[:, None, None] ** 2 + i[None, :, None] ** 2 + kz[None, None, :] ** 2)
    mu2 = np.where(kk > 0, kz[None, None, :] ** 2 / np.where(kk > 0, kk, 1) ** 2, 0)
    assert np.max(np.abs(r - f * (1 + 0.5 * mu2))) <= 1e-12 * np.max(np.abs(f))
    # the non-Gaussian term: delta + fnl (delta^2 - <delta^2>) in real space
    n = O.generate_field(NN, 9, fnl=0.7, beta=0.0).view(np.complex128)
    d = np.fft.irfftn(g, s=(NN, NN, NN), axes=(0, 1, 2))   # numpy's irfftn divides by NN^3 = V
    dn = d + 0.7 * (d * d - np.mean(d * d))
    ref = np.fft.rfftn(dn)
    assert np.max(np.abs(n - ref)) <= 1e-11 * np.max(np.abs(ref))
