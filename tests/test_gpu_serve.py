"""Request server on the GPU: a field's bins are the same bits alone and in a
mixed, batched, concurrent workload (the server's lanes hold estimators of
its max_batch, so the references are estimators of that batch size);
admission under a small HBM budget evicts idle lanes and refuses a request
that cannot fit."""
import numpy as np
import pytest

from geoms import small, medium, config1
import barbie_amd as B

pytestmark = pytest.mark.gpu


def alone(gf, NRi, prec, max_batch=1):
    g = gf()
    return B.Estimator(B.Plan(g), prec, max_batch=max_batch).estimate(B.make_field(g.NN, NRi=NRi))


@pytest.mark.parametrize("prec", ["bf16x3", "fp64"])
def test_results_same_alone_and_in_the_mix(prec):
    cases = [(medium, 1), (config1, 2), (small, 3), (medium, 4), (config1, 5), (medium, 6),
             (small, 7), (config1, 8), (medium, 9)]
    ref = {c: alone(c[0], c[1], prec, max_batch=3) for c in cases}
    srv = B.Server(prec, max_inflight=3, max_batch=3)
    # a request alone
    t = srv.submit(medium(), B.make_field(medium().NN, NRi=1))
    bs, nm, bi, lat = srv.wait(t)
    for x, y in zip((bs, nm, bi), ref[(medium, 1)]):
        assert np.array_equal(x, y)
    assert lat[2] > 0
    # the mix: interleaved geometries, batched and concurrent (queued as one
    # burst, so that same-geometry requests share batches)
    fields = {c: B.make_field(c[0]().NN, NRi=c[1]) for c in cases}
    srv.pause(True)
    tickets = [(c, srv.submit(c[0](), fields[c])) for c in cases * 2]
    assert all(srv.poll(t) == B.native.REQ_QUEUED for _, t in tickets)
    srv.pause(False)
    for c, t in tickets:
        bs, nm, bi, lat = srv.wait(t)
        for x, y in zip((bs, nm, bi), ref[c]):
            assert np.array_equal(x, y), c
    st = srv.stats()
    assert st["completed"] == 1 + 2 * len(cases) and st["failed"] == 0
    assert st["batched_fields"] > st["batches"]   # some batches held several fields
    srv.close()


def test_device_resident_fields():
    torch = pytest.importorskip("torch")
    g = config1()
    f = B.make_field(g.NN, NRi=3)
    ref = alone(config1, 3, "bf16x3", max_batch=4)   # the server's default batch
    d = torch.from_numpy(f).to("cuda")
    torch.cuda.synchronize()
    srv = B.Server("bf16x3")
    bs, nm, bi, _ = srv.wait(srv.submit(g, d.data_ptr()))
    assert np.array_equal(bs, ref[0]) and np.array_equal(nm, ref[1])
    srv.close()


def test_admission_evicts_and_refuses():
    gm, gc = medium(), config1()
    need = max(B.Plan(gm).footprint("fp64", max_batch=1),
               B.Plan(gc).footprint("fp64", max_batch=1))
    field_bytes = max(g.NN * g.NN * (g.NN + 2) * 8 for g in (gm, gc))
    budget = need + field_bytes + (1 << 20)    # room for one lane at a time
    srv = B.Server("fp64", hbm_budget=budget, max_inflight=1, max_batch=1)
    ts = [(gf, k, srv.submit(gf(), B.make_field(gf().NN, NRi=k)))
          for k, gf in enumerate([medium, config1, medium, config1], start=1)]
    for gf, k, t in ts:
        bs, nm, bi, _ = srv.wait(t)
        assert np.array_equal(bs, alone(gf, k, "fp64")[0])
    st = srv.stats()
    assert st["lanes_evicted"] >= 2 and st["peak_bytes"] <= budget
    srv.close()
    tiny = B.Server("fp64", hbm_budget=1 << 20)
    t = tiny.submit(gm, B.make_field(gm.NN, NRi=1))
    with pytest.raises(B.BarbieError):
        tiny.wait(t)
    tiny.close()


def test_short_requests_beside_long_ones_with_tuning():
    """tune=1 with reserved CUs (the default 1/8): the tuned layouts keep
    every lane on its CU-masked server stream, so short requests (NN=128) that
    arrive while long ones (NN=384) run finish well before them (compared by
    median: a stream of the process can share a hardware queue with a long
    lane's, which holds up the odd short request), and their bins are the
    same bits as when they ran alone on the same server."""
    torch = pytest.importorskip("torch")
    gl = B.Geometry(NN=384, UL_k1=51, LL_k1=1, BW_k1=1)
    gs = config1()
    dl = B.make_field_device(gl.NN, NRi=2)
    ds = torch.from_numpy(B.make_field(gs.NN, NRi=4)).to("cuda")
    torch.cuda.synchronize()
    srv = B.Server("bf16x3", max_inflight=4, max_batch=2, tune=True)
    srv.prepare(gl, 2)
    srv.prepare(gs, 2)
    for g, d in ((gl, dl), (gs, ds)):   # service times learned before the mix
        for _ in range(2):
            ref = srv.wait(srv.submit(g, d.data_ptr()))
    longs = [srv.submit(gl, dl.data_ptr()) for _ in range(6)]
    shorts = [srv.submit(gs, ds.data_ptr()) for _ in range(4)]
    lat_s = []
    for t in shorts:
        bs, nm, bi, lat = srv.wait(t)
        assert np.array_equal(bs, ref[0]) and np.array_equal(nm, ref[1])
        lat_s.append(lat[2])
    lat_l = [srv.wait(t)[3][2] for t in longs]
    print("short latencies (ms)", ["%.2f" % x for x in lat_s], "long",
          ["%.2f" % x for x in lat_l])
    assert np.median(lat_s) < 0.5 * np.median(lat_l)
    assert srv.stats()["failed"] == 0
    srv.close()
