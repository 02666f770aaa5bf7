"""Writes tests/golden/ref_funcs.npz: input/output vectors of the
reference's own input_ps (funcs.c:73-81), sort (funcs.c:113-126) and
cal_coeff (funcs.c:708-745), evaluated by oracle/_ref/libref_funcs.so (built
from the reference's funcs.c by `make -C oracle ref`).  The vectors pin the
oracle's restatement of these routines where the reference is absent.

Run from the repository root: python tests/golden/make_ref_vectors.py
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_funcs.so")
OUT = os.path.join(ROOT, "tests", "golden", "ref_funcs.npz")
_dp = ctypes.POINTER(ctypes.c_double)


def ref_lib(path=LIB):
    L = ctypes.CDLL(path)
    L.input_ps.argtypes = [ctypes.c_double]
    L.input_ps.restype = ctypes.c_double
    L.sort.argtypes = [_dp]
    L.cal_coeff.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp]
    return L


def inputs(seed=20240514):
    rng = np.random.default_rng(seed)
    # (px, pz): the estimator's range (|pz| <= 1, px of either sign), special
    # points, and values beyond it (the polynomials are defined everywhere)
    pts = [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (0.0, -1.0), (1.0, 1.0), (-1.0, 0.5),
           (0.5, -0.5), (2.0, 0.3), (-0.7071067811865476, 0.7071067811865476)]
    pxpz = np.concatenate([np.array(pts), rng.uniform(-1.2, 1.2, (300, 2)),
                           rng.uniform(-1.0, 1.0, (200, 2))])
    # sort: {(|k|, k_z)} triples, with ties in |k| (the bubble sort keeps the
    # input order of equal keys) and integer-valued keys
    K = rng.uniform(0.1, 60.0, (300, 3, 2))
    ties = rng.integers(1, 6, (200, 3, 1)).astype(float)
    K = np.concatenate([K, np.concatenate([ties, rng.uniform(-5, 5, (200, 3, 1))], axis=2)])
    kk = np.concatenate([rng.uniform(0.01, 10.0, 100), 10.0 ** rng.uniform(-3, 3, 100),
                         [1.0, 2.0, 0.5, 3.0]])
    return pxpz, K, kk


def evaluate(L, pxpz, K, kk):
    bs = np.zeros((len(pxpz), 16))
    nm = np.zeros((len(pxpz), 16))
    for i, (px, pz) in enumerate(pxpz):
        L.cal_coeff(float(px), float(pz), bs[i].ctypes.data_as(_dp), nm[i].ctypes.data_as(_dp))
    Ks = np.ascontiguousarray(K.copy())
    for i in range(len(Ks)):
        L.sort(Ks[i].ctypes.data_as(_dp))
    ps = np.array([L.input_ps(float(x)) for x in kk])
    return bs, nm, Ks, ps


def main():
    pxpz, K, kk = inputs()
    bs, nm, Ks, ps = evaluate(ref_lib(), pxpz, K, kk)
    np.savez_compressed(OUT, pxpz=pxpz, bs=bs, norm=nm, K_in=K, K_sorted=Ks, kk=kk, input_ps=ps)
    print(f"wrote {OUT}: {len(pxpz)} cal_coeff points, {len(K)} sorts, {len(kk)} input_ps values")


if __name__ == "__main__":
    main()
