"""CPU tests of the C ABI: the library loads and exports every declared
symbol, the host planner reproduces the oracle's enumeration and the
reference geometry's counts, and invalid input is refused with an error."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from geoms import tiny, small, medium, radix5, config1, oracle_params
from test_oracle import oracle_edges
import barbie_amd as B
from barbie_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "barbie.h")).read()
    declared = set(re.findall(r"\b(bb_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(native.EXPORTED)
    L = ctypes.CDLL(B.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert b"gfx950" in B.lib().bb_version()


def test_reference_geometry_counts():
    # barbie.c constants: Nmax=100, Nr=40, piecewise rings; UL=51 (config 2)
    p = B.Plan(B.Geometry(NN=384, UL_k1=51))
    s = p.stats
    assert s["rcount"] == 3620
    assert s["ncontr"] == 2756382
    assert s["npairs"] == 387848
    assert s["ngroups"] == 3826
    assert s["nfields_used"] == 1189
    assert s["n_k1bin"] == 50 and s["nbins"] == 50 * 25


def test_other_workload_counts():
    assert B.Plan(config1()).stats["ncontr"] == 83221
    g = config1()
    g.last_ring_to_nmax = 1
    assert B.Plan(g).stats["ncontr"] == 65840
    for NN, Nr, UL, n in [(128, 13, 18, 51371), (192, 19, 27, 269376), (256, 26, 36, 794278)]:
        st = B.Plan(B.Geometry(NN=NN, UL_k1=UL, Nmax=NN // 4, Nr=Nr)).stats
        assert st["ncontr"] == n


@pytest.mark.parametrize("gf", [tiny, small, medium, radix5])
def test_plan_matches_oracle_enumeration(gf):
    g = gf()
    plan = B.Plan(g)
    lo, hi, av = oracle_edges(g)
    plo, phi, pav = plan.ring_edges()
    assert np.array_equal(lo, plo) and np.array_equal(hi, phi) and np.array_equal(av, pav)
    rd = O.calcr(g.NN, lo, hi, av)
    assert np.array_equal(plan.triples(), rd[:, :3].astype(np.int32))
    field = make_small_field(g.NN)
    _, _, _, n, contr = O.estimate(oracle_params(g), lo, hi, av, rd, field, record=True)
    pc = plan.contractions()
    assert n == len(pc)
    assert np.array_equal(pc[:, :6], contr[:, :6].astype(np.int32))
    assert np.array_equal(pc[:, 7], contr[:, 7].astype(np.int32))


def make_small_field(NN):
    return B.make_field(NN, NRi=1)


@pytest.mark.parametrize("kw,msg", [
    (dict(NN=15), "NN"), (dict(NN=14), "NN"), (dict(BW_k1=0), "BW_k1"),
    (dict(UL_k1=1), "UL_k1"), (dict(UL_k1=8, BW_k1=3), "divisible"),
    (dict(Nmax=9), "Nmax"), (dict(Nr=0), "Nr"),
])
def test_invalid_geometry_is_refused(kw, msg):
    args = dict(NN=16, UL_k1=6, LL_k1=1, BW_k1=1, Nmax=4, Nr=4, ring_rule=B.RINGS_UNIFORM)
    args.update(kw)
    with pytest.raises(B.BarbieError) as ei:
        B.Plan(B.Geometry(**args))
    assert msg in str(ei.value)
    assert ei.value.code == -1


def test_ring_outside_grid_is_refused():
    with pytest.raises(B.BarbieError):
        B.Plan(B.Geometry(NN=16, UL_k1=6, Nmax=8, Nr=12, ring_rule=B.RINGS_UNIFORM,
                          last_ring_to_nmax=0))


def test_explicit_rings():
    g = tiny()
    lo, hi, av = oracle_edges(g)
    ge = B.Geometry(NN=16, UL_k1=6, Nmax=4, Nr=4, ring_rule=B.RINGS_EXPLICIT, ring_low=lo,
                    ring_high=hi)
    assert np.array_equal(B.Plan(ge).contractions(), B.Plan(g).contractions())


def test_cli_usage_and_validation(tmp_path):
    import subprocess
    r = subprocess.run([B.BIN_PATH], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "Usage:" in r.stdout
    r = subprocess.run([B.BIN_PATH, "384", "1", "51", "1", "0"], capture_output=True, text=True,
                       cwd=tmp_path)
    assert r.returncode == 2 and "BW_k1" in r.stderr
    r = subprocess.run([B.BIN_PATH, "16", "1", "6", "1", "1"], capture_output=True, text=True,
                       cwd=tmp_path, env=dict(os.environ, BARBIE_NMAX="4", BARBIE_NR="4",
                                              BARBIE_RING_RULE="uniform"))
    assert r.returncode == 3 and "field_N16_R1.txt" in r.stderr


def test_write_outputs_format_matches_oracle(tmp_path):
    g = tiny()
    plan = B.Plan(g)
    rng = np.random.default_rng(0)
    nb = g.nbins * 16
    bs, nm, bi = rng.standard_normal(nb), np.abs(rng.standard_normal(nb)), rng.standard_normal(nb)
    nm[::7] = 0.0
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    plan.write_outputs(bs, nm, bi, str(tmp_path / "a"), 5)
    O.write_outputs(oracle_params(g), bs, nm, bi, str(tmp_path / "b"), 5)
    for ii in range(16):
        name = f"BS_test_ii{ii}_N16_R5.txt"
        assert (tmp_path / "a" / name).read_text() == (tmp_path / "b" / name).read_text()


@pytest.mark.parametrize("gf", [medium, config1])
def test_bins_follow_the_law_of_cosines(gf):
    """Shape bins (barbie.c:287-298, 343-345) against an independent form:
    with K sorted descending (stable on ties, funcs.c:113-126), T = K1/K0 and
    MU is the cosine of the triangle's angle between its two longest sides,
    (K0^2 + K1^2 - K2^2) / (2 K0 K1), and K1_bin = floor((K0 - LL_k1)/BW_k1).
    Contractions whose MU or T lies within 1e-9 of a bin edge are skipped."""
    g = gf()
    plan = B.Plan(g)
    _, _, av = plan.ring_edges()
    c = plan.contractions().astype(np.int64)
    p = c[:, [0, 2, 4]].astype(float)
    r = c[:, [1, 3, 5]]
    K = np.sqrt(av[r] ** 2 + p ** 2)
    order = np.argsort(-K, axis=1, kind="stable")
    Ks = np.take_along_axis(K, order, axis=1)
    T = Ks[:, 1] / Ks[:, 0]
    MU = (Ks[:, 0] ** 2 + Ks[:, 1] ** 2 - Ks[:, 2] ** 2) / (2 * Ks[:, 0] * Ks[:, 1])
    n = g.n_mtbin
    kb = np.floor((Ks[:, 0] - g.LL_k1) / g.BW_k1).astype(int)
    tb = np.where(T == 1.0, n - 1, np.floor((T - 0.5) / (0.5 / n))).astype(int)
    mb = np.where(MU == 1.0, n - 1, np.floor((MU - 0.5) / (0.5 / n))).astype(int)
    flat = (kb * n + tb) * n + mb

    def near_edge(x):
        y = (x - 0.5) / (0.5 / n)
        return np.abs(y - np.round(y)) < 1e-9
    ok = ~(near_edge(T) & (T != 1.0)) & ~(near_edge(MU) & (MU != 1.0))
    assert ok.mean() > 0.9
    assert np.array_equal(flat[ok], c[ok, 7])
    assert np.all((MU >= 0.5 - 1e-12) & (MU <= 1.0 + 1e-12))


def test_checked_outputs_mark_missing_denominators(tmp_path):
    """bb_write_outputs_checked: rows as the reference writes them
    (monopole denominator > .1), every multipole whose own denominator is
    zero, negative or not finite written as "nan" and counted; the other
    entries identical to the unchecked files."""
    g = tiny()
    plan = B.Plan(g)
    rng = np.random.default_rng(1)
    nb = g.nbins * 16
    bs, nm, bi = rng.standard_normal(nb), np.abs(rng.standard_normal(nb)) + 0.2, rng.standard_normal(nb)
    nm4 = nm.reshape(-1, 16)
    nm4[::3, 2] = 0.0
    nm4[1::5, 7] = np.inf
    nm4[2::7, 11] = -1.0
    nm4[4::9, 0] = 0.05          # row dropped, as the reference drops it
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    assert plan.write_outputs(bs, nm, bi, str(tmp_path / "a"), 3) == 0
    n = plan.write_outputs(bs, nm, bi, str(tmp_path / "b"), 3, checked=True)
    rows = nm4[:, 0] > .1
    expect = 0
    for ii in range(16):
        d = nm4[:, ii]
        bad = rows & ~((d > 0) & np.isfinite(d))
        expect += int(bad.sum())
        name = f"BS_test_ii{ii}_N16_R3.txt"
        la = (tmp_path / "a" / name).read_text().splitlines()
        lb = (tmp_path / "b" / name).read_text().splitlines()
        assert len(la) == len(lb) == int(rows.sum())
        for row, x, y in zip(np.nonzero(rows)[0], la, lb):
            if bad[row]:
                fb = y.split("\t")
                assert fb[3] == "nan" and fb[4] == "nan", y
                assert fb[:3] == x.split("\t")[:3] and fb[5] == x.split("\t")[5]
            else:
                assert x == y
    assert n == expect and n > 0
