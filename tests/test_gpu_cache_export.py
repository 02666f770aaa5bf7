"""bb_export_cache: the reference's on-disk caches (closed_rings_<NN>,
rcountFor_<NN>, dataPlaneRing_<NN>_<p>_<r>_R<NRi>; funcs.c:328-415, 537-566)
written from a field, checked against the CPU oracle: the ring triples and
triangle counts against orc_calcr, every plane-ring field against orc_df
(the FFTW-backward, unnormalised D_{p,r} on the NN grid) to 1e-12 of the
field's scale."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from geoms import small, medium, radix5
from test_oracle import oracle_edges
import barbie_amd as B

pytestmark = pytest.mark.gpu


def check_cache(d, g, field, NRi):
    NN = g.NN
    lo, hi, av = oracle_edges(g)
    ref = O.calcr(NN, lo, hi, av)
    assert int((d / f"rcountFor_{NN}").read_text()) == len(ref)
    rd = np.fromfile(d / f"closed_rings_{NN}").reshape(-1, 4)
    assert rd.shape == ref.shape
    assert np.array_equal(rd[:, :3], ref[:, :3])
    assert np.allclose(rd[:, 3], ref[:, 3], rtol=1e-12, atol=1e-9)
    worst = 0.0
    for p in range(g.Nmax):
        for r in range(g.Nr):
            D = np.fromfile(d / f"dataPlaneRing_{NN}_{p}_{r}_R{NRi}")
            assert D.size == 2 * NN * NN
            D = D.view(np.complex128).reshape(NN, NN)
            o = O.df(field, NN, p, lo[r], hi[r])
            scale = max(np.abs(o).max(), 1e-300)
            worst = max(worst, np.abs(D - o).max() / scale if np.abs(o).max() > 0 else np.abs(D).max())
    assert worst <= 1e-12, worst
    cfg = (d / f"cache_config_{NN}_R{NRi}.txt").read_text().split("\n")
    assert cfg[0] == f"NN {NN}" and any(line.startswith("key ") for line in cfg)


@pytest.mark.parametrize("gf", [small, radix5, medium])
def test_export_matches_oracle(tmp_path, gf):
    g = gf()
    field = B.make_field(g.NN, NRi=2)
    est = B.Estimator(B.Plan(g), "bf16x3")   # the export is FP64 whatever the precision
    est.export_cache(field, str(tmp_path), 2)
    check_cache(tmp_path, g, field, 2)


def test_cli_writes_the_reference_caches(tmp_path):
    g = medium()
    field = B.make_field(g.NN, NRi=4)
    B.write_field(str(tmp_path / f"field_N{g.NN}_R4.txt"), field)
    env = dict(os.environ, BARBIE_NMAX=str(g.Nmax), BARBIE_NR=str(g.Nr), BARBIE_WRITE_CACHE="1")
    r = subprocess.run([B.BIN_PATH, str(g.NN), "4", str(g.UL_k1), str(g.LL_k1), str(g.BW_k1)],
                       cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = tmp_path / f"ResFiles_N{g.NN}" / "data_files"
    assert len(list(d.glob(f"dataPlaneRing_{g.NN}_*_R4"))) == g.Nmax * g.Nr
    check_cache(d, g, field, 4)


def test_import_rings_checks_and_uses_the_file(tmp_path):
    """bb_import_rings: an exported cache reads back (0 mismatches) and gives
    the same bins; a changed Num, a foreign geometry's key, or another
    geometry's triples are refused and change nothing; a cache without the
    key file (the reference's own) is accepted when its content agrees."""
    g = medium()
    plan = B.Plan(g)
    field = B.make_field(g.NN, NRi=3)
    est = B.Estimator(plan, "fp64")
    ref = est.estimate(field)
    est.export_cache(field, str(tmp_path), 3)
    assert est.import_rings(str(tmp_path), 3) == 0
    got = est.estimate(field)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    NN = g.NN
    rings = tmp_path / f"closed_rings_{NN}"
    good = rings.read_bytes()
    rd = np.frombuffer(good).reshape(-1, 4).copy()
    rd[5, 3] *= 1.01
    rd.tofile(rings)
    with pytest.raises(B.BarbieError, match="triangle counts differ"):
        est.import_rings(str(tmp_path), 3)
    for a, b in zip(est.estimate(field), ref):
        assert np.array_equal(a, b)
    rings.write_bytes(good)
    # a geometry with the same NN but other rings: its key differs
    g2 = B.Geometry(NN=NN, UL_k1=g.UL_k1, LL_k1=1, BW_k1=1, Nmax=g.Nmax, Nr=10,
                    ring_rule=B.RINGS_REFERENCE)
    e2 = B.Estimator(B.Plan(g2), "fp64")
    with pytest.raises(B.BarbieError, match="another geometry"):
        e2.import_rings(str(tmp_path), 3)
    # without the key file the content decides: other triples are refused,
    # the matching cache is accepted
    (tmp_path / f"cache_config_{NN}_R3.txt").unlink()
    with pytest.raises(B.BarbieError):
        e2.import_rings(str(tmp_path), 3)
    assert est.import_rings(str(tmp_path), 3) == 0
    (tmp_path / f"rcountFor_{NN}").unlink()
    with pytest.raises(B.BarbieError):
        est.import_rings(str(tmp_path), 3)


def test_check_plane_cache_detects_a_foreign_field(tmp_path):
    g = small()
    plan = B.Plan(g)
    f1, f2 = B.make_field(g.NN, NRi=1), B.make_field(g.NN, NRi=2)
    est = B.Estimator(plan, "fp64")
    est.export_cache(f1, str(tmp_path), 1)
    assert est.check_plane_cache(f1, str(tmp_path), 1) <= 1e-14
    assert est.check_plane_cache(f1, str(tmp_path), 1, stride=5) <= 1e-14
    assert est.check_plane_cache(f2, str(tmp_path), 1) > 0.1
    (tmp_path / f"dataPlaneRing_{g.NN}_0_0_R1").unlink()
    with pytest.raises(B.BarbieError):
        est.check_plane_cache(f1, str(tmp_path), 1)


def test_cli_reads_the_caches_back(tmp_path):
    """BARBIE_READ_CACHE=1 after a BARBIE_WRITE_CACHE=1 run: the same bins;
    a stale plane-ring cache (another field in the same file name) is named
    on stderr; a foreign ring cache exits with status 2."""
    g = medium()
    NN = g.NN
    f1, f2 = B.make_field(NN, NRi=4), B.make_field(NN, NRi=5)
    env = dict(os.environ, BARBIE_NMAX=str(g.Nmax), BARBIE_NR=str(g.Nr))
    args = [B.BIN_PATH, str(NN), "4", str(g.UL_k1), str(g.LL_k1), str(g.BW_k1)]
    B.write_field(str(tmp_path / f"field_N{NN}_R4.txt"), f1)
    r = subprocess.run(args, cwd=tmp_path, env=dict(env, BARBIE_WRITE_CACHE="1"),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = tmp_path / f"ResFiles_N{NN}"
    first = np.fromfile(out / f"BS_bins_N{NN}_R4.f64")
    r = subprocess.run(args, cwd=tmp_path, env=dict(env, BARBIE_READ_CACHE="1"),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "read from" in r.stdout and "warning" not in r.stderr
    assert np.array_equal(np.fromfile(out / f"BS_bins_N{NN}_R4.f64"), first)
    # the same file name now holds another realization: the cache is stale
    B.write_field(str(tmp_path / f"field_N{NN}_R4.txt"), f2)
    r = subprocess.run(args, cwd=tmp_path, env=dict(env, BARBIE_READ_CACHE="1"),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "does not belong to this field" in r.stderr
    # a ring cache of another geometry with the same NN is refused
    r = subprocess.run(args, cwd=tmp_path, env=dict(env, BARBIE_READ_CACHE="1", BARBIE_NR="10"),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "another geometry" in r.stderr
