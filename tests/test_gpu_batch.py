"""Batched and asynchronous estimates on the GPU: every field of a batch
gets exactly (bit for bit) the bins it gets alone, because a batch runs the
same work lists and the same summation order per field."""
import numpy as np
import pytest

from geoms import medium, config1
import barbie_amd as B

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
@pytest.mark.parametrize("gf", [medium, config1])
def test_batch_equals_single(gf, prec):
    g = gf()
    plan = B.Plan(g)
    fields = np.stack([B.make_field(g.NN, NRi=k) for k in (1, 2, 3)])
    single = B.Estimator(plan, prec)
    alone = [single.estimate(f) for f in fields]
    batched = B.Estimator(plan, prec, max_batch=4)
    bs, nm, bi = batched.estimate_batch(fields)
    for k in range(3):
        assert np.array_equal(bs[k], alone[k][0]), k
        assert np.array_equal(nm[k], alone[k][1]), k
        assert np.array_equal(bi[k], alone[k][2]), k
    # a smaller batch on the same estimator
    bs2, _, _ = batched.estimate_batch(fields[1:2])
    assert np.array_equal(bs2[0], alone[1][0])


def test_batch_larger_than_estimator_is_refused():
    g = medium()
    est = B.Estimator(B.Plan(g), "fp64", max_batch=2)
    fields = np.stack([B.make_field(g.NN, NRi=k) for k in (1, 2, 3)])
    with pytest.raises(B.BarbieError):
        est.estimate_batch(fields)


def test_async_batch_matches_sync():
    torch = pytest.importorskip("torch")
    g = config1()
    plan = B.Plan(g)
    fields = np.stack([B.make_field(g.NN, NRi=k) for k in (5, 6)])
    ref = B.Estimator(plan, "bf16x3", max_batch=2).estimate_batch(fields)
    est = B.Estimator(plan, "bf16x3", max_batch=2)
    d_f = torch.from_numpy(fields).to("cuda")
    nb = g.nbins * 16
    out = torch.zeros(3, 2, nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    est.estimate_batch_device(2, d_f.data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                              out[2].data_ptr(), wait=False)
    while not est.done():
        pass
    est.wait()
    got = out.cpu().numpy()
    for k, r in enumerate(ref):
        assert np.array_equal(got[k], r.reshape(2, nb))
