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
