"""Properties of the estimator at full size, where the CPU oracle would take
an hour or more per field: the reference configuration (NN=384, Nmax=100,
Nr=40, UL_k1=51; barbie.c:61-84) and the largest served one (NN=512,
Nmax=134, Nr=54 uniform rings, UL_k1=73).  Each follows from the
estimator's definition
(S = sum_x Re D1 D2 D3, barbie.c:326-337, over closed triangles of
isotropic rings) and holds at any size:

* cubic in the field: scaling the input by 2 scales every floating-point
  step by a power of two, so bin_BS scales by exactly 8 and the
  normalisations and the analytic column do not change (FP64 bit for bit;
  BF16x3 within 1e-12, where subnormal products could round differently);
  negating the input negates bin_BS (FP64 bit for bit; BF16x3 within 1e-5
  per multipole: every step of this build is symmetric under negation, but
  the bf16 matrix instruction's FP32 accumulation is not, measured here at
  about 2e-6);
* translation invariant: moving the real-space field by one cell along x
  multiplies mode (i, j, k) by exp(-2 pi i i / NN); closed triangles carry
  no net phase, so the bins are unchanged (FP64 within 1e-10 of the largest
  bin, the parity bar; BF16x3 within its 1e-5 per multipole of FP64);
* symmetric under exchanging x and y: the rings depend on ia^2 + ja^2 only,
  and the line of sight is z (FP64 within 1e-10).

The power-spectrum multipoles (calpow_mom) scale by exactly 4 and are
unchanged by the exchange of x and y (within 1e-12: the summation order
changes).
"""
import math

import numpy as np
import pytest

import barbie_amd as B

pytestmark = pytest.mark.gpu

GEOMS = {
    "reference384": dict(NN=384, UL_k1=51, LL_k1=1, BW_k1=1),
    "uniform512": dict(NN=512, UL_k1=73, LL_k1=1, BW_k1=1, Nmax=134, Nr=54, n_mtbin=5,
                       ring_rule=B.RINGS_UNIFORM, ring_width=134 / 54, last_ring_to_nmax=1),
}
_EST = {}


def estimator(geom, prec):
    if (geom, prec) not in _EST:
        g = B.Geometry(**GEOMS[geom])
        # the tuned BF16x3 layout of the reference geometry; defaults otherwise
        opts = (dict(grid_levels=4, waves=4, occupancy=3)
                if prec == "bf16x3" and geom == "reference384" else {})
        _EST[(geom, prec)] = (g, B.Estimator(B.Plan(g), prec, **opts))
    return _EST[(geom, prec)]


def bins(case, prec, d):
    """The three bin arrays of the field tensor d (its producer is waited
    for: the estimator reads it on its own stream)."""
    import torch
    g, est = estimator(case[0], prec)
    nb = g.nbins * 16
    out = torch.zeros(3 * nb, dtype=torch.float64, device="cuda")
    p = out.data_ptr()
    torch.cuda.synchronize()
    est.estimate_device(d.data_ptr(), p, p + 8 * nb, p + 16 * nb)
    o = out.cpu().numpy()
    return o[:nb].reshape(-1, 16), o[nb:2 * nb].reshape(-1, 16), o[2 * nb:].reshape(-1, 16)


@pytest.fixture(scope="module", params=list(GEOMS))
def case(request):
    """(geometry name, device field); the previous geometry's estimators are
    released first."""
    for k in [k for k in _EST if k[0] != request.param]:
        _EST.pop(k)[1].close()
    NN = GEOMS[request.param]["NN"]
    return request.param, B.make_field_device(NN, NRi=3, fnl=0.3)


def shifted_x(d):
    """The r2c buffer of the real-space field moved by one cell along x."""
    import torch
    NN = d.shape[0]
    i = torch.arange(NN, dtype=torch.float64, device=d.device)
    ph = -2.0 * math.pi * i / NN
    c, s = torch.cos(ph)[:, None, None], torch.sin(ph)[:, None, None]
    re, im = d[..., 0::2], d[..., 1::2]
    out = torch.empty_like(d)
    out[..., 0::2] = re * c - im * s
    out[..., 1::2] = re * s + im * c
    return out


def swapped_xy(d):
    """The r2c buffer of the field with the x and y axes exchanged."""
    return d.transpose(0, 1).contiguous()


def per_multipole(a, b):
    return [float(np.linalg.norm(a[:, i] - b[:, i]) / np.linalg.norm(b[:, i])) for i in range(16)]


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_bins_are_cubic_in_the_field(case, prec):
    field = case[1]
    a = bins(case, prec, field)
    b = bins(case, prec, 2.0 * field)
    n = bins(case, prec, -field)
    assert np.abs(a[0]).max() > 0
    if prec == "fp64":
        assert np.array_equal(b[0], 8.0 * a[0])
        assert np.array_equal(n[0], -a[0])
    else:
        assert np.abs(b[0] - 8.0 * a[0]).max() <= 1e-12 * np.abs(8.0 * a[0]).max()
        err = per_multipole(n[0], -a[0])
        print("bf16x3 negation, per-multipole rel diff", ["%.1e" % e for e in err])
        assert max(err) <= 1e-5
    for k in (1, 2):
        assert np.array_equal(b[k], a[k]) and np.array_equal(n[k], a[k])


def test_bins_are_translation_invariant(case):
    field = case[1]
    a = bins(case, "fp64", field)
    t = bins(case, "fp64", shifted_x(field))
    for k in range(3):
        scale = np.abs(a[k]).max()
        assert np.abs(t[k] - a[k]).max() <= 1e-10 * scale, k
    tb = bins(case, "bf16x3", shifted_x(field))
    assert max(per_multipole(tb[0], a[0])) <= 1e-5
    assert np.abs(tb[1] - a[1]).max() <= 1e-12 * np.abs(a[1]).max()


def test_bins_are_symmetric_in_x_and_y(case):
    field = case[1]
    a = bins(case, "fp64", field)
    s = bins(case, "fp64", swapped_xy(field))
    for k in range(3):
        scale = np.abs(a[k]).max()
        assert np.abs(s[k] - a[k]).max() <= 1e-10 * scale, k


def test_power_multipoles_scale_and_symmetry(case):
    import torch
    field = case[1]
    NN = field.shape[0]
    doubled, swapped = 2.0 * field, swapped_xy(field)
    torch.cuda.synchronize()   # the library reads device fields on its own stream
    a = B.power_multipoles(field.data_ptr(), NN)
    b = B.power_multipoles(doubled.data_ptr(), NN)
    s = B.power_multipoles(swapped.data_ptr(), NN)
    assert np.array_equal(b[:, [0, 5]], a[:, [0, 5]])
    assert np.array_equal(b[:, 1:5], 4.0 * a[:, 1:5])
    assert np.array_equal(s[:, 5], a[:, 5])
    pop = a[:, 5] > 0
    assert np.all(np.abs(s[pop] - a[pop]) <= 1e-12 * np.abs(a[pop]) + 1e-12 * np.abs(a[pop, 1:2]))
