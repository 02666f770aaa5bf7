"""Pins of the oracle's restated routines against the reference's own code
and against explicit closed-form answers.

* input_ps (funcs.c:73-81), sort (funcs.c:113-126) and cal_coeff
  (funcs.c:708-745): bit for bit against the reference's functions, compiled
  from its funcs.c by `make -C oracle ref` (oracle/_ref/libref_funcs.so),
  through the vectors they produced (tests/golden/ref_funcs.npz, written by
  tests/golden/make_ref_vectors.py), and against the live library at fresh
  inputs when it is built.
* The triangle parameters (barbie.c:293-297) on explicit closed 3-D
  triangles: MU is minus the cosine between the two largest sides, pz the
  line-of-sight cosine of the largest, and px the line-of-sight component of
  the unit vector in the triangle's plane perpendicular to the largest side,
  on the side of the second (Gram-Schmidt).
* The output writer (barbie.c:383-397): the exact text of the 16 files for
  bins with hand-computed values.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref_funcs.npz")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_funcs.so")


def oracle_eval(pxpz, K, kk):
    bs = np.zeros((len(pxpz), 16))
    nm = np.zeros((len(pxpz), 16))
    for i, (px, pz) in enumerate(pxpz):
        bs[i], nm[i] = O.cal_coeff(px, pz)
    Ks = np.array([O.sort3(k) for k in K])
    ps = np.array([O.input_ps(x) for x in kk])
    return bs, nm, Ks, ps


def test_oracle_routines_equal_reference_vectors():
    g = np.load(GOLD)
    bs, nm, Ks, ps = oracle_eval(g["pxpz"], g["K_in"], g["kk"])
    assert np.array_equal(bs, g["bs"])
    assert np.array_equal(nm, g["norm"])
    assert np.array_equal(Ks, g["K_sorted"])
    assert np.array_equal(ps, g["input_ps"])
    # the vectors hold ties in |k|: the sort keeps equal keys in input order
    ties = [k for k in g["K_in"] if len(set(k[:, 0])) < 3]
    assert len(ties) > 50


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (needs the reference)")
def test_oracle_routines_equal_live_reference():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_ref_vectors as M
    rng = np.random.default_rng(7)
    pxpz = rng.uniform(-1.5, 1.5, (300, 2))
    K = np.concatenate([rng.uniform(0, 80, (200, 3, 2)),
                        np.concatenate([rng.integers(0, 4, (100, 3, 1)).astype(float),
                                        rng.normal(size=(100, 3, 1))], axis=2)])
    kk = 10.0 ** rng.uniform(-2, 2, 200)
    ref = M.evaluate(M.ref_lib(REF_LIB), pxpz, K, kk)
    got = oracle_eval(pxpz, K, kk)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def closed_triangles(n, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        k0 = rng.normal(size=3) * rng.uniform(1, 50)
        k1 = rng.normal(size=3) * rng.uniform(1, 50)
        ks = [k0, k1, -(k0 + k1)]
        out.append(ks)
    return out


def test_triangle_parameters_follow_the_3d_geometry():
    zhat = np.array([0.0, 0.0, 1.0])
    for ks in closed_triangles(400):
        K = np.array([[np.linalg.norm(k), k[2]] for k in ks])
        Ks = O.sort3(K)                                  # the reference's sort
        order = sorted(range(3), key=lambda i: -np.linalg.norm(ks[i]))
        v0, v1 = ks[order[0]], ks[order[1]]
        if len({round(float(np.linalg.norm(k)), 12) for k in ks}) < 3:
            continue                                     # ties: order fixed by the sort
        K1, T, MU, pz, px = O.triangle_params(Ks)
        u0, u1 = v0 / np.linalg.norm(v0), v1 / np.linalg.norm(v1)
        assert K1 == pytest.approx(np.linalg.norm(v0), rel=1e-14)
        assert T == pytest.approx(np.linalg.norm(v1) / np.linalg.norm(v0), rel=1e-14)
        assert MU == pytest.approx(-np.dot(u0, u1), abs=1e-12)
        assert pz == pytest.approx(np.dot(zhat, u0), abs=1e-14)
        e = u1 - np.dot(u1, u0) * u0                     # Gram-Schmidt: in plane, normal to u0
        e /= np.linalg.norm(e)
        assert px == pytest.approx(np.dot(zhat, e), abs=1e-9)
        # an invariant of the parameterisation: pz^2 + px^2 <= 1
        assert pz * pz + px * px <= 1 + 1e-9


def test_writer_text_is_exact(tmp_path):
    """Two size bins x 2 x 2 shape bins (n_mtbin=2), NN=4: every value of
    the files follows from the bins by hand."""
    params = dict(NN=4, Nmax=2, Nr=2, n_mtbin=2, UL_k1=3, LL_k1=1, BW_k1=1, LL=1.0, ff=0.1,
                  beta=1.0)
    nb = 2 * 2 * 2
    bs = np.zeros((nb, 16))
    nm = np.zeros((nb, 16))
    bi = np.zeros((nb, 16))
    # bin (0,0,0): norm 32 for all multipoles; bs = 64 * (ii + 1); biana = 16
    nm[0, :] = 32.0
    bs[0, :] = 64.0 * (np.arange(16) + 1)
    bi[0, :] = 16.0
    # bin (1,1,0): monopole norm 0.05 (below the .1 filter): not written
    nm[6, :] = 0.05
    bs[6, :] = 1.0
    # bin (1,0,1): monopole 2, multipole 2's own norm 0 (prints x/0 = -inf, inf)
    nm[5, :] = 2.0
    nm[5, 2] = 0.0
    bs[5, :] = -6.4
    bi[5, :] = 3.0
    O.write_outputs(params, bs.ravel(), nm.ravel(), bi.ravel(), str(tmp_path), 9)
    vol = 4.0 ** 3
    for ii in range(16):
        text = (tmp_path / f"BS_test_ii{ii}_N4_R9.txt").read_text()
        l0 = "0\t0\t0\t%e\t%e\t%f\n" % (64.0 * (ii + 1) / (vol * 32.0), 16.0 / 32.0, 32.0 / 16.0)
        if ii == 2:   # x / 0 with the multipole's own denominator 0
            l1 = "1\t0\t1\t-inf\tinf\t0.000000\n"
        else:
            l1 = "1\t0\t1\t%e\t%e\t%f\n" % (-6.4 / (vol * 2.0), 3.0 / 2.0, 2.0 / 16.0)
        assert text == l0 + l1, (ii, text)
    # literal spot checks of the formatting itself
    assert (tmp_path / "BS_test_ii0_N4_R9.txt").read_text().splitlines()[0] == \
        "0\t0\t0\t3.125000e-02\t5.000000e-01\t2.000000"
    assert (tmp_path / "BS_test_ii1_N4_R9.txt").read_text().splitlines()[1] == \
        "1\t0\t1\t-5.000000e-02\t1.500000e+00\t0.125000"
