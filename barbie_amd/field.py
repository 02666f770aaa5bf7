"""Seeded synthetic input fields in the reference's input contract.

Produces the FFTW in-place r2c buffer [NN][NN][NN+2] float64 that `barbie`
reads from field_N<NN>_R<NRi>.txt.  The construction follows the order of the
reference generator (generate_field.c:44-107: Fill_Modes -> Make_Image ->
Non_Gauss -> Get_Modes -> RSD, ngrf_2d_funcs.c:341-577, 734-765) with the
workload's own parameters: P(k) = A k^-1.5 in grid units (A = 1e3), modes
with k >= NN/3 and k = 0 set to zero, delta = dG + f (dG^2 - <dG^2>) with
f = 0.3, then Kaiser (1 + beta mu^2), beta = 1, mu = kz/k, cell size 1.
The reference draws its Gaussian modes with GSL's mt19937 + ziggurat (absent
from this image).  Two generators here, both seeded with 100020 + 146*NRi
(generate_field.c:47):
  * make_field (host, numpy): PCG64 draws a real white-noise cube whose FFT
    gives Hermitian Gaussian modes with <|Delta|^2> = V P(k);
  * make_field_device (GPU): the library's bb_generate_field, which draws
    every Fourier mode from a Philox4x64-10 counter-based stream and runs
    the same pipeline in HIP.

These are input tools (the step before the estimator), not part of the
estimator's hot path.
"""
from __future__ import annotations

import numpy as np


def seed_for(NRi: int) -> int:
    return 100020 + 146 * NRi


def make_field(NN: int, NRi: int = 1, fnl: float = 0.3, amp: float = 1e3, index: float = -1.5,
               beta: float = 1.0, kcut_frac: float = 1.0 / 3.0) -> np.ndarray:
    """Returns float64 array (NN, NN, NN+2): interleaved complex r2c modes."""
    rng = np.random.default_rng(seed_for(NRi))
    white = rng.standard_normal((NN, NN, NN))
    W = np.fft.rfftn(white)                       # <|W|^2> = NN^3 = V
    del white
    k = np.fft.fftfreq(NN, d=1.0 / NN)
    kz = np.arange(NN // 2 + 1, dtype=np.float64)
    k2 = (k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    kk = np.sqrt(k2)
    with np.errstate(divide="ignore"):
        pk = np.where((kk > 0) & (kk < NN * kcut_frac), amp * kk ** index, 0.0)
    W *= np.sqrt(pk)
    del pk
    dG = np.fft.irfftn(W, s=(NN, NN, NN), axes=(0, 1, 2))      # c2r with 1/V (Make_Image)
    del W
    var = np.mean(dG * dG)
    dG += fnl * (dG * dG - var)                   # Non_Gauss
    D = np.fft.rfftn(dG)                          # Get_Modes (LL^3 = 1)
    del dG
    with np.errstate(divide="ignore", invalid="ignore"):
        mu2 = np.where(kk > 0, kz[None, None, :] ** 2 / k2, 0.0)
    D *= (1.0 + beta * mu2)                       # RSD (k = 0 untouched)
    out = np.empty((NN, NN, NN + 2), dtype=np.float64)
    out.view(np.complex128)[...] = D
    return out


def write_field(path: str, field: np.ndarray):
    """Raw native-endian doubles, as barbie.c:176-185 reads them."""
    np.ascontiguousarray(field, dtype=np.float64).tofile(path)


def read_field(path: str, NN: int) -> np.ndarray:
    return np.fromfile(path, dtype=np.float64).reshape(NN, NN, NN + 2)


def make_field_device(NN: int, NRi: int = 1, fnl: float = 0.3, amp: float = 1e3,
                      index: float = -1.5, beta: float = 1.0, kcut_frac: float = 1.0 / 3.0,
                      device: str = "cuda"):
    """The same construction on the GPU by the library's own generator
    (bb_generate_field: a Philox4x64-10 stream with seed seed_for(NRi), one
    block per Fourier mode, then the c2r / non-Gaussian / r2c / Kaiser steps
    in HIP), so the realization differs from make_field's while the
    statistics are the same.  Returns a float64 tensor (NN, NN, NN+2) on
    `device`, complete when returned."""
    import torch
    from .native import generate_field
    out = torch.empty((NN, NN, NN + 2), dtype=torch.float64, device=device)
    torch.cuda.synchronize()
    generate_field(NN, seed_for(NRi), out=out.data_ptr(), amp=amp, index=index, kcut=NN * kcut_frac,
                   fnl=fnl, beta=beta)
    return out
