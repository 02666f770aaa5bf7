"""GPU parity: the HIP estimator against the CPU oracle on the same seeded
inputs, through the C ABI.

Tolerances (stated per check): FP64 path within 1e-10 of the oracle
(relative to the largest magnitude of the compared quantity), FP32 path
within 2e-5 norm-wise per multipole.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from geoms import tiny, small, medium, radix5, config1, oracle_params
from test_oracle import oracle_edges
import barbie_amd as B

pytestmark = pytest.mark.gpu

_ORC = {}


def oracle_run(gf, NRi, zero=False):
    key = (gf.__name__, NRi, zero)
    if key not in _ORC:
        g = gf()
        lo, hi, av = oracle_edges(g)
        rd = O.calcr(g.NN, lo, hi, av)
        field = np.zeros((g.NN, g.NN, g.NN + 2)) if zero else B.make_field(g.NN, NRi=NRi)
        bs, nm, bi, n, contr = O.estimate(oracle_params(g), lo, hi, av, rd, field, record=True)
        _ORC[key] = dict(g=g, field=field, rd=rd, bs=bs, nm=nm, bi=bi, contr=contr,
                         edges=(lo, hi, av))
    return _ORC[key]


def rel_err(a, b):
    scale = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / scale if scale > 0 else np.max(np.abs(a))


# Grid modes: the NN grid of the reference ("full"), one alias-free grid for
# every contraction ("one_grid") and the default grid levels ("levels").
MODES = {"full": dict(full_grid=True), "one_grid": dict(grid_levels=1), "levels": {}}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("gf", [tiny, small, radix5, medium, config1])
def test_fp64_matches_oracle(gf, mode):
    o = oracle_run(gf, NRi=1)
    g = o["g"]
    plan = B.Plan(g)
    est = B.Estimator(plan, "fp64", **MODES[mode])
    if mode == "full":
        assert est.grid == g.NN and est.levels == [g.NN]
    if mode == "one_grid":
        assert len(est.levels) == 1
    bs, nm, bi = est.estimate(o["field"])
    # triangle counts Num (calCR) within 1e-12
    assert rel_err(est.triangle_counts(), o["rd"][:, 3]) <= 1e-12
    # per-contraction sums within 1e-10 of the largest |S|
    S = est.contraction_sums()
    So = o["contr"][:, 6]
    assert S.shape == So.shape
    assert rel_err(S, So) <= 1e-10
    big = np.abs(So) > 1e-3 * np.max(np.abs(So))
    assert np.max(np.abs(S[big] - So[big]) / np.abs(So[big])) <= 1e-9
    # exact-zero rule decided identically
    assert np.array_equal(S == 0.0, So == 0.0)
    # bins, per multipole
    obs, onm, obi = (o[k].reshape(bs.shape) for k in ("bs", "nm", "bi"))
    for ii in range(16):
        assert rel_err(bs[..., ii], obs[..., ii]) <= 1e-10, ii
        assert rel_err(nm[..., ii], onm[..., ii]) <= 1e-12, ii
        assert rel_err(bi[..., ii], obi[..., ii]) <= 1e-10, ii


@pytest.mark.parametrize("gf", [tiny, radix5, medium])
def test_fp64_plane_ring_fields_match_oracle(gf):
    o = oracle_run(gf, NRi=1)
    g = o["g"]
    lo, hi, _ = o["edges"]
    plan = B.Plan(g)
    est = B.Estimator(plan, "fp64", full_grid=True)
    est.estimate(o["field"])
    pc = plan.contractions()
    used = {(int(p), int(r)) for p, r in zip(pc[:, 0], pc[:, 1])}
    used |= {(abs(int(p)), int(r)) for p, r in zip(pc[:, 2], pc[:, 3])}
    for p, r in sorted(used)[:: max(1, len(used) // 12)]:
        ref = O.df(o["field"], g.NN, p, lo[r], hi[r])
        got = est.field(0, p, r)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))), (p, r)


def resample(d_nn, NG):
    """The trigonometric polynomial sampled on the NN grid, resampled on NG
    points per side (numpy FFT): its modes must lie within |k| < NG/2."""
    NN = d_nn.shape[0]
    A = np.fft.fft2(d_nn) / NN ** 2
    k = np.arange(NN)
    s = np.where(k > NN // 2, k - NN, k)
    keep = np.abs(s) < NG // 2
    out = np.zeros((NG, NG), dtype=np.complex128)
    ii = np.nonzero(keep)[0]
    out[np.ix_(s[ii] % NG, s[ii] % NG)] = A[np.ix_(ii, ii)]
    tol = 1e-12 * np.max(np.abs(A))
    assert np.all(np.abs(A[~keep, :]) <= tol) and np.all(np.abs(A[:, ~keep]) <= tol)
    return np.fft.ifft2(out) * NG ** 2


@pytest.mark.parametrize("gf", [medium, config1])
def test_fp64_level_fields_match_oracle(gf):
    """Every level holds its fields on its own grid: the oracle's NN-grid
    field resampled onto that grid, within 1e-12."""
    o = oracle_run(gf, NRi=1)
    g = o["g"]
    lo, hi, _ = o["edges"]
    plan = B.Plan(g)
    est = B.Estimator(plan, "fp64")
    est.estimate(o["field"])
    grids, cl = plan.grid_levels("fp64")
    assert grids == est.levels and len(grids) > 1
    pc = plan.contractions()
    for lv, NG in enumerate(grids):
        sel = pc[cl == lv]
        used = {(int(p), int(r)) for p, r in zip(sel[:, 0], sel[:, 1])}
        used |= {(abs(int(p)), int(r)) for p, r in zip(sel[:, 4], sel[:, 5])}
        for p, r in sorted(used)[:: max(1, len(used) // 6)]:
            ref = resample(O.df(o["field"], g.NN, p, lo[r], hi[r]), NG)
            got = est.field(lv, p, r)
            assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))), (lv, p, r)


def test_levels_match_one_grid():
    """Grid levels and a single grid give the same sums (to FP64 rounding)."""
    g = config1()
    f = B.make_field(g.NN, NRi=3)
    plan = B.Plan(g)
    a = B.Estimator(plan, "fp64", grid_levels=1)
    b = B.Estimator(plan, "fp64")
    assert len(b.levels) > 1
    ra, rb = a.estimate(f), b.estimate(f)
    for x, y in zip(ra, rb):
        assert rel_err(y, x) <= 1e-12
    assert rel_err(b.contraction_sums(), a.contraction_sums()) <= 1e-12


def test_zero_signal_field():
    o = oracle_run(small, NRi=1, zero=True)
    g = o["g"]
    est = B.Estimator(B.Plan(g), "fp64")
    bs, nm, bi = est.estimate(o["field"])
    assert np.all(bs == 0) and np.all(nm == 0) and np.all(bi == 0)
    assert np.all(o["nm"] == 0)


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_zero_ring_field_and_back(prec):
    """A field with the modes of one ring zeroed in every plane: its nonzero
    flags differ from a generic field's, so its norm and analytic sums are
    accumulated for it instead of taken from the estimator's
    field-independent sums (the zero rule removes every contraction with a
    field of that ring).  Bins within 1e-10 (FP64) of the oracle, Num-weighted
    norms within 1e-12; then the same estimator gives a generic field its
    oracle bins again, and the first field its bins bit for bit."""
    g = medium()
    o = oracle_run(medium, NRi=1)
    lo, hi, av = o["edges"]
    NN = g.NN
    ia = np.fft.fftfreq(NN, 1.0 / NN)
    rad = np.sqrt(ia[:, None] ** 2 + ia[None, :] ** 2)
    f = o["field"].copy()
    f[(rad >= lo[3]) & (rad < hi[3]), :] = 0.0
    obs, onm, obi, _, _ = O.estimate(oracle_params(g), lo, hi, av, o["rd"], f)
    est = B.Estimator(B.Plan(g), prec)
    tol = 1e-10 if prec == "fp64" else 1e-5
    for run in range(2):
        bs, nm, bi = est.estimate(f)
        assert rel_err(nm, onm.reshape(nm.shape)) <= 1e-12
        assert rel_err(bs, obs.reshape(bs.shape)) <= tol
        assert rel_err(bi, obi.reshape(bi.shape)) <= 1e-10
        assert np.any(nm != B.Estimator(B.Plan(g), prec).estimate(o["field"])[1])
        if run == 0:
            first = (bs, nm, bi)
        else:
            for x, y in zip(first, (bs, nm, bi)):
                assert np.array_equal(x, y)
        bs2, nm2, bi2 = est.estimate(o["field"])   # a generic field in between
        assert rel_err(nm2, o["nm"].reshape(nm2.shape)) <= 1e-12
        assert rel_err(bs2, o["bs"].reshape(bs2.shape)) <= tol


def test_fp64_deterministic_and_estimator_reuse():
    g = medium()
    est = B.Estimator(B.Plan(g), "fp64")
    f1 = B.make_field(g.NN, NRi=1)
    f2 = B.make_field(g.NN, NRi=2)
    a = est.estimate(f1)
    b = est.estimate(f2)
    c = est.estimate(f1)
    for x, y in zip(a, c):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], b[0])


# Norm-wise relative error per multipole, ||est - orc||_2 / ||orc||_2 over
# all bins: FP32 within 2e-6, BF16x3 within 1e-5 of the FP64 oracle.
@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 1e-5)])
@pytest.mark.parametrize("gf", [medium, config1])
def test_reduced_precision_within_tolerance(gf, prec, tol):
    o = oracle_run(gf, NRi=1)
    g = o["g"]
    est = B.Estimator(B.Plan(g), prec)
    bs, nm, bi = est.estimate(o["field"])
    obs, onm = (o[k].reshape(bs.shape) for k in ("bs", "nm"))
    errs = []
    for ii in range(16):
        d = np.linalg.norm(bs[..., ii] - obs[..., ii]) / np.linalg.norm(obs[..., ii])
        errs.append(d)
        assert rel_err(nm[..., ii], onm[..., ii]) <= 1e-12
    print(prec, gf.__name__, "per-multipole rel err", ["%.1e" % e for e in errs])
    assert max(errs) <= tol, errs


@pytest.mark.parametrize("x_splits", [1, 3, 7])
def test_split_count_does_not_change_result(x_splits):
    g = medium()
    f = B.make_field(g.NN, NRi=4)
    ref = B.Estimator(B.Plan(g), "fp64", x_splits=1).estimate(f)
    got = B.Estimator(B.Plan(g), "fp64", x_splits=x_splits).estimate(f)
    for a, b in zip(got, ref):
        assert rel_err(a, b) <= 1e-13


def test_cli_end_to_end(tmp_path):
    """`barbie NN NRi UL LL BW` in a directory holding field_N<NN>_R<NRi>.txt
    writes the 16 files the reference writes; values match the oracle's own
    writer to the printed precision and the FP64 dump to 1e-10."""
    o = oracle_run(medium, NRi=1)
    g = o["g"]
    B.write_field(str(tmp_path / f"field_N{g.NN}_R7.txt"), o["field"])
    env = dict(os.environ, BARBIE_NMAX=str(g.Nmax), BARBIE_NR=str(g.Nr))
    r = subprocess.run([B.BIN_PATH, str(g.NN), "7", str(g.UL_k1), str(g.LL_k1), str(g.BW_k1)],
                       cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    outdir = tmp_path / f"ResFiles_N{g.NN}"
    refdir = tmp_path / "oracle"
    refdir.mkdir()
    O.write_outputs(oracle_params(g), o["bs"], o["nm"], o["bi"], str(refdir), 7)
    for ii in range(16):
        name = f"BS_test_ii{ii}_N{g.NN}_R7.txt"
        a = np.loadtxt(outdir / name, ndmin=2)
        b = np.loadtxt(refdir / name, ndmin=2)
        assert a.shape == b.shape and a.shape[0] > 0
        assert np.array_equal(a[:, :3], b[:, :3])
        fin = np.isfinite(b[:, 3])
        assert np.array_equal(fin, np.isfinite(a[:, 3]))
        # 7 printed digits; m != 0 bins that cancel to ~0 compare against
        # the column's scale
        for col in (3, 4, 5):
            scale = np.max(np.abs(b[fin, col])) if fin.any() else 0.0
            assert np.allclose(a[fin, col], b[fin, col], rtol=2e-6, atol=1e-9 * scale), col
    dump = np.fromfile(outdir / f"BS_bins_N{g.NN}_R7.f64").reshape(3, -1)
    assert rel_err(dump[0], o["bs"]) <= 1e-10
    assert rel_err(dump[1], o["nm"]) <= 1e-12
    # timing log in the reference's format (barbie.c:368-377): one line per
    # plane p1, then the loop's total
    lines = (outdir / f"time_{g.NN}_R7.txt").read_text().strip().splitlines()
    assert len(lines) == g.Nmax + 1 and lines[-1].startswith("Total time ")
    t = [float(ln.split("in  time ")[1]) for ln in lines[:-1]]
    for q, ln in enumerate(lines[:-1]):
        assert ln.startswith(f"loop {q} out of {g.Nmax} in  time "), ln
    assert abs(sum(t) - float(lines[-1].split()[-1])) <= 1e-5 + 1e-3 * sum(t)


def test_cli_split_precision_with_checked_denominators(tmp_path):
    """`barbie` with BARBIE_PRECISION=bf16x3 and BARBIE_CHECK_DENOMINATORS=1:
    the FP64 dump within the BF16x3 bar (1e-5 per multipole, norm-wise) of
    the oracle, Num to 1e-12, and in the 16 files exactly the written rows'
    multipoles whose own oracle denominator is not positive read "nan"."""
    o = oracle_run(medium, NRi=1)
    g = o["g"]
    B.write_field(str(tmp_path / f"field_N{g.NN}_R3.txt"), o["field"])
    env = dict(os.environ, BARBIE_NMAX=str(g.Nmax), BARBIE_NR=str(g.Nr),
               BARBIE_PRECISION="bf16x3", BARBIE_CHECK_DENOMINATORS="1")
    r = subprocess.run([B.BIN_PATH, str(g.NN), "3", str(g.UL_k1), str(g.LL_k1), str(g.BW_k1)],
                       cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    outdir = tmp_path / f"ResFiles_N{g.NN}"
    dump = np.fromfile(outdir / f"BS_bins_N{g.NN}_R3.f64").reshape(3, -1, 16)
    obs, onm = o["bs"].reshape(-1, 16), o["nm"].reshape(-1, 16)
    for ii in range(16):
        assert np.linalg.norm(dump[0][:, ii] - obs[:, ii]) <= 1e-5 * np.linalg.norm(obs[:, ii])
    assert rel_err(dump[1], onm) <= 1e-12
    rows = onm[:, 0] > .1
    expect = 0
    for ii in range(16):
        lines = (outdir / f"BS_test_ii{ii}_N{g.NN}_R3.txt").read_text().splitlines()
        assert len(lines) == int(rows.sum())
        bad = ~(onm[rows, ii] > 0)
        expect += int(bad.sum())
        for line, isbad in zip(lines, bad):
            f = line.split("\t")
            assert (f[3] == "nan" and f[4] == "nan") == bool(isbad), line
    if expect:
        assert f"{expect} (bin, multipole) entries" in r.stderr
    else:
        assert "nan" not in r.stderr


def test_tuned_layout_matches_oracle():
    """bb_tune picks among layouts that all give the reference's sums."""
    o = oracle_run(medium, NRi=1)
    g = o["g"]
    plan = B.Plan(g)
    opts, ms = plan.tune("fp64", repeats=1)
    assert opts["waves"] in (4, 8) and opts["grid_levels"] in (4, 5, 6, 8) and ms > 0
    est = B.Estimator(plan, "fp64", **opts)
    bs, nm, bi = est.estimate(o["field"])
    obs = o["bs"].reshape(bs.shape)
    assert rel_err(bs, obs) <= 1e-10
    four = B.Estimator(plan, "fp64", waves=4, grid_levels=3).estimate(o["field"])
    assert rel_err(four[0], obs) <= 1e-10


@pytest.mark.parametrize("gf", [medium, config1])
def test_mixed_bf16_fp32_layout_within_tolerance(gf):
    """BF16x3 with 40% of the cost on the concurrent FP32 kernel: within the
    BF16x3 bar (1e-5 norm-wise per multipole) of the FP64 oracle; Num exact."""
    o = oracle_run(gf, NRi=1)
    g = o["g"]
    plan = B.Plan(g)
    est = B.Estimator(plan, "bf16x3", mix_fp32=40)
    wl = est.worklist()
    bs, nm, bi = est.estimate(o["field"])
    obs, onm = (o[k].reshape(bs.shape) for k in ("bs", "nm"))
    errs = [np.linalg.norm(bs[..., i] - obs[..., i]) / np.linalg.norm(obs[..., i]) for i in range(16)]
    assert max(errs) <= 1e-5, errs
    assert rel_err(nm, onm) <= 1e-12
    S = est.contraction_sums()
    So = o["contr"][:, 6]
    assert np.array_equal(S == 0.0, So == 0.0)
    # the mixed estimator's units include both kernels' units
    assert wl["units"] > B.Estimator(plan, "bf16x3").worklist()["units"] // 4


# (waves, occupancy) -> most staged rings per item of that BF16x3 kernel shape
BF16X3_SHAPES = {(4, 3): 96, (4, 0): 120, (8, 0): 240}


@pytest.mark.parametrize("shape", sorted(BF16X3_SHAPES))
@pytest.mark.parametrize("gf", [medium, config1])
def test_bf16x3_kernel_shapes_within_tolerance(gf, shape):
    """Each BF16x3 kernel shape (4 waves and 6 tile windows per wave sized for
    3 workgroups per CU, at most 96 staged rings per item; 4 waves and 8
    windows, 120 rings; 8 waves, 240 rings): within the BF16x3 bar of the FP64
    oracle, and the same zero pattern of the sums."""
    o = oracle_run(gf, NRi=1)
    g = o["g"]
    plan = B.Plan(g)
    waves, occ = shape
    est = B.Estimator(plan, "bf16x3", waves=waves, occupancy=occ)
    wl = est.worklist()
    assert wl["max_rings"] <= BF16X3_SHAPES[shape] and wl["waves"] == waves
    bs, nm, bi = est.estimate(o["field"])
    obs, onm = (o[k].reshape(bs.shape) for k in ("bs", "nm"))
    errs = [np.linalg.norm(bs[..., i] - obs[..., i]) / np.linalg.norm(obs[..., i]) for i in range(16)]
    assert max(errs) <= 1e-5, errs
    assert rel_err(nm, onm) <= 1e-12
    S = est.contraction_sums()
    So = o["contr"][:, 6]
    assert np.array_equal(S == 0.0, So == 0.0)


def wide_rings():   # NN=96, 6 rings of width 8: about 2200 modes per outer ring
    return B.Geometry(NN=96, UL_k1=49, LL_k1=1, BW_k1=4, Nmax=6, Nr=6,
                      ring_rule=B.RINGS_UNIFORM, ring_width=8.0, last_ring_to_nmax=0)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("bf16x3", 1e-5)])
def test_wide_rings_stage_by_block(prec, tol):
    """Rings of width 8: the outer rings' fields hold more mode pairs than
    the ring pass keeps in LDS for a whole field (640 in FP64, 1024 in FP32),
    so their modes are staged again for every column block, in groups of
    rows.  Fields within 1e-12 (FP64) of the oracle's Df; bins within 1e-10
    (FP64) or 1e-5 per multipole (BF16x3) of the oracle."""
    o = oracle_run(wide_rings, NRi=1)
    g = o["g"]
    lo, hi, av = o["edges"]
    ia = np.fft.fftfreq(g.NN, 1.0 / g.NN)
    rad = np.sqrt(ia[:, None] ** 2 + ia[None, :] ** 2)
    assert np.count_nonzero((rad >= lo[5]) & (rad < hi[5])) > 2 * 1024
    est = B.Estimator(B.Plan(g), prec)
    bs, nm, bi = est.estimate(o["field"])
    obs, onm, obi = (o[k].reshape(bs.shape) for k in ("bs", "nm", "bi"))
    assert rel_err(nm, onm) <= 1e-12
    assert rel_err(bi, obi) <= 1e-10
    for ii in range(16):
        if np.linalg.norm(obs[..., ii]) > 0:
            assert (np.linalg.norm(bs[..., ii] - obs[..., ii]) / np.linalg.norm(obs[..., ii])
                    <= tol), ii
    if prec == "fp64":
        fields = B.Estimator(B.Plan(g), "fp64", full_grid=True)
        fields.estimate(o["field"])
        for p, r in ((1, 5), (0, 4), (3, 5), (0, 0)):
            ref = O.df(o["field"], g.NN, p, lo[r], hi[r])
            got = fields.field(0, p, r)
            assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))), (p, r)


# Grid sizes whose ring-pass FFT plans cover every stage radix (largest odd
# factor first): 30 = 5 x 6, 60 = 5 x 12, 72 = 9 x 8, 90 = 15 x 6,
# 100 = 5 x 5 x 4, 160 = 10 x 16, 96 = 6 x 16.
@pytest.mark.parametrize("NN", [30, 60, 72, 90, 96, 100, 160])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_ring_pass_fields_every_radix(NN, prec, tol):
    """Plane-ring fields of the ring pass on the NN grid against the
    oracle's Df (a direct masked inverse 2-D FFT), relative to the largest
    magnitude: FP64 within 1e-12, FP32 within 2e-6."""
    g = B.Geometry(NN=NN, UL_k1=NN // 2, LL_k1=1, BW_k1=1, Nmax=4, Nr=4,
                   ring_rule=B.RINGS_UNIFORM, ring_width=NN / 10, last_ring_to_nmax=0)
    lo, hi, av = O.edges_uniform(g.Nr, g.ring_width, g.Nmax, g.last_ring_to_nmax)
    field = B.make_field(NN, NRi=3)
    plan = B.Plan(g)
    est = B.Estimator(plan, prec, full_grid=True)
    assert est.levels == [NN]
    est.estimate(field)
    pc = plan.contractions()
    used = {(int(p), int(r)) for p, r in zip(pc[:, 0], pc[:, 1])}
    used |= {(abs(int(p)), int(r)) for p, r in zip(pc[:, 2], pc[:, 3])}
    used |= {(abs(int(p)), int(r)) for p, r in zip(pc[:, 4], pc[:, 5])}
    assert len(used) >= 8
    for p, r in sorted(used):
        ref = O.df(field, NN, p, lo[r], hi[r])
        got = est.field(0, p, r)
        assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref))), (p, r)
