"""Grid levels, host side (no GPU).

A contraction of rings (r1, r2, r3) sums exactly on a grid of NG points per
side when no sum of three of its modes wraps around (NG > K1 + K2 + K3, K_r
the largest |k_x| of ring r) and the ring transform holds every mode of each
ring (NG >= 2 K_r + 2).  These tests take K_r from the oracle's own ring
membership (the mask of orc_df, recovered with numpy's FFT), check both
bounds for every contraction, and check that the chosen grids minimise the
summed points against brute force over every admissible grid set.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as O
from geoms import tiny, small, medium, radix5, config1
from test_oracle import oracle_edges
import barbie_amd as B


def ring_kx(g):
    lo, hi, _ = oracle_edges(g)
    ones = np.zeros((g.NN, g.NN, g.NN + 2))
    ones[..., 0::2] = 1.0
    idx = np.arange(g.NN)
    signed = np.where(idx > g.NN // 2, idx - g.NN, idx)
    K = []
    for r in range(g.Nr):
        d = O.df(ones, g.NN, 1, lo[r], hi[r])
        m = np.abs(np.fft.fft2(d)) > 0.5 * g.NN ** 2     # fft2 of the field = NN^2 * mask
        rows = np.nonzero(m.any(axis=1))[0]
        K.append(int(np.max(np.abs(signed[rows]))) if rows.size else 0)
    return np.array(K)


def needs(g, plan):
    K = ring_kx(g)
    c = plan.contractions()
    k1, k2, k3 = K[c[:, 1]], K[c[:, 3]], K[c[:, 5]]
    return np.maximum(k1 + k2 + k3 + 1, 2 * np.maximum(np.maximum(k1, k2), k3) + 2), K


def smooth_even(n):
    if n % 2:
        return False
    for f in (2, 3, 5):
        while n % f == 0:
            n //= f
    return n == 1


@pytest.mark.parametrize("gf", [tiny, small, medium, radix5, config1])
def test_every_contraction_runs_on_an_admissible_grid(gf):
    g = gf()
    plan = B.Plan(g)
    need, K = needs(g, plan)
    grids, cl = plan.grid_levels("fp64")
    assert grids == sorted(set(grids)) and len(grids) <= 8
    assert all(smooth_even(n) and n <= g.NN for n in grids)
    G = np.array(grids)[cl]
    assert np.all(G >= need)
    # every level holds at least one contraction
    assert set(np.unique(cl)) == set(range(len(grids)))


@pytest.mark.parametrize("gf", [medium, config1])
@pytest.mark.parametrize("L", [1, 2, 3])
def test_chosen_grids_minimise_summed_points(gf, L):
    g = gf()
    plan = B.Plan(g)
    need, _ = needs(g, plan)
    grids, cl = plan.grid_levels("fp64", grid_levels=L)
    assert len(grids) <= L
    got = float(np.sum(np.array(grids, dtype=float)[cl] ** 2))
    cand = [n for n in range(8, g.NN + 1) if smooth_even(n)]
    top = min(n for n in cand if n >= need.max())
    cand = [n for n in cand if n <= top]
    best = np.inf
    for k in range(1, L + 1):
        for s in itertools.combinations(cand, k):
            if s[-1] != top:
                continue
            arr = np.array(s)
            G = arr[np.searchsorted(arr, need)]
            best = min(best, float(np.sum(G.astype(float) ** 2)))
    assert got == best
    if L == 1:
        assert grids == [top]


def test_full_grid_and_aliasing_grid_use_nn():
    g = medium()
    grids, cl = B.Plan(g).grid_levels("fp64", full_grid=True)
    assert grids == [g.NN] and np.all(cl == 0)
    # rings out to k = 10 on a 24 grid: 3 * 10 + 1 > 24, the NN-grid sums
    # alias and only the NN grid reproduces them
    ga = B.Geometry(NN=24, UL_k1=8, LL_k1=1, BW_k1=1, Nmax=6, Nr=10, ring_rule=B.RINGS_UNIFORM,
                    ring_width=1.0, last_ring_to_nmax=0)
    pa = B.Plan(ga)
    _, K = needs(ga, pa)
    assert 3 * K.max() + 1 > ga.NN
    grids, cl = pa.grid_levels("fp64")
    assert grids == [ga.NN] and np.all(cl == 0)


def test_level_count_option_is_validated():
    plan = B.Plan(medium())
    with pytest.raises(B.BarbieError):
        plan.grid_levels("fp64", grid_levels=17)


def test_waves_option_is_validated():
    plan = B.Plan(medium())
    o = B.native.BBOptions(B.PRECISIONS["fp64"], 0, 0, 0, 0, 5)
    out = (B.native.ctypes.c_longlong * len(B.native.WORKLIST_KEYS))()
    rc = B.lib().bb_plan_worklist(plan.handle, B.native.ctypes.byref(o), out)
    assert rc == -1
    w4 = plan.worklist("fp64")
    assert w4["waves"] == 8


def test_occupancy_option_is_validated():
    plan = B.Plan(medium())
    W = len(B.native.WORKLIST_KEYS)
    for prec, waves, occ in [("fp64", 4, 3), ("bf16x3", 8, 3), ("bf16x3", 4, 5)]:
        o = B.native.BBOptions(B.PRECISIONS[prec], 0, 0, 0, 0, waves, 0, 0, occ)
        out = (B.native.ctypes.c_longlong * W)()
        assert B.lib().bb_plan_worklist(plan.handle, B.native.ctypes.byref(o), out) == -1
    w3 = plan.worklist("bf16x3", waves=4, occupancy=3)
    w2 = plan.worklist("bf16x3", waves=4)
    assert w3["max_rings"] <= 96
    assert w3["tiles"] == w2["tiles"]


def test_pair_point_figures_are_consistent():
    """The work-list figures of the pair products: every contraction's pair
    row is formed once, so the pair points used lie between the points of
    the distinct pairs and those formed (16 rows per block); tile windows
    times points cover every contraction's points once per 16 columns."""
    g = medium()
    plan = B.Plan(g)
    C = plan.contractions()
    for prec, opts in (("bf16x3", dict(waves=4, occupancy=3)), ("fp64", {})):
        w = plan.worklist(prec, **opts)
        grids, cl = plan.grid_levels(prec)
        ng2 = np.array([grids[l] ** 2 for l in cl], dtype=np.int64)
        # distinct pair rows (level, p1, r1, signed p2, r2), each with its points
        pairs = {}
        for c in range(len(C)):
            pairs[(cl[c], C[c, 0], C[c, 1], C[c, 2], C[c, 3])] = ng2[c]
        distinct = sum(pairs.values())
        assert w["pair_points_used"] >= distinct
        assert w["pair_points"] >= w["pair_points_used"]
        assert w["pair_points"] % 16 == 0
        assert 256 * w["window_points"] >= w["points_summed"]
        assert w["points_summed"] == int(ng2.sum())
