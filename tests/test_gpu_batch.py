"""Batched and asynchronous estimates on the GPU: every field of a batch
gets exactly (bit for bit) the bins it gets alone on the same estimator,
because a batch runs the same work lists and the same summation order per
field.  Estimators of different batch sizes split the work into x-ranges
differently (about 2048 workgroups per launch), so their bins agree to
rounding only."""
import numpy as np
import pytest

from geoms import medium, config1
import barbie_amd as B

pytestmark = pytest.mark.gpu


# Norm-wise per multipole agreement of estimators of different batch sizes
# (different x-ranges): FP64 rounding, and the FP32 accumulation groups of
# the split paths.
CROSS_TOL = {"fp64": 1e-12, "bf16x3": 1e-6}


def close_per_multipole(a, b, tol):
    a = np.asarray(a).reshape(-1, 16)
    b = np.asarray(b).reshape(-1, 16)
    for i in range(16):
        nb = np.linalg.norm(b[:, i])
        assert np.linalg.norm(a[:, i] - b[:, i]) <= tol * max(nb, 1e-300), i


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
@pytest.mark.parametrize("gf", [medium, config1])
def test_batch_equals_single(gf, prec):
    g = gf()
    plan = B.Plan(g)
    fields = np.stack([B.make_field(g.NN, NRi=k) for k in (1, 2, 3)])
    batched = B.Estimator(plan, prec, max_batch=4)
    alone = [batched.estimate(f) for f in fields]
    bs, nm, bi = batched.estimate_batch(fields)
    for k in range(3):
        assert np.array_equal(bs[k], alone[k][0]), k
        assert np.array_equal(nm[k], alone[k][1]), k
        assert np.array_equal(bi[k], alone[k][2]), k
    # a smaller batch on the same estimator
    bs2, _, _ = batched.estimate_batch(fields[1:2])
    assert np.array_equal(bs2[0], alone[1][0])
    # an estimator of one field per batch: other x-ranges, the same sums
    single = B.Estimator(plan, prec)
    for k in range(3):
        r = single.estimate(fields[k])
        close_per_multipole(bs[k], r[0], CROSS_TOL[prec])
        assert np.array_equal(nm[k], r[1]) and np.array_equal(bi[k], r[2])


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
    torch.cuda.synchronize()   # the estimator runs on its own stream
    est.estimate_batch_device(2, d_f.data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                              out[2].data_ptr(), wait=False)
    while not est.done():
        pass
    est.wait()
    got = out.cpu().numpy()
    for k, r in enumerate(ref):
        assert np.array_equal(got[k], r.reshape(2, nb))


def test_async_batches_queue_in_order():
    """Asynchronous batches enqueued back to back, without waiting, run in
    order on the estimator's stream: each output equals the synchronous
    result of its own field (bit for bit), including a batch that reuses the
    input buffer of an earlier one after it is overwritten on the host's
    stream only once everything has finished."""
    torch = pytest.importorskip("torch")
    g = config1()
    plan = B.Plan(g)
    fields = np.stack([B.make_field(g.NN, NRi=k) for k in (7, 8, 9)])
    single = B.Estimator(plan, "bf16x3")
    ref = [single.estimate(f) for f in fields]
    est = B.Estimator(plan, "bf16x3")
    d_f = torch.from_numpy(fields).to("cuda")
    nb = g.nbins * 16
    out = torch.zeros(3, 3, nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # the estimator runs on its own stream
    for k in range(3):
        est.estimate_batch_device(1, d_f[k].data_ptr(), out[0, k].data_ptr(),
                                  out[1, k].data_ptr(), out[2, k].data_ptr(), wait=False)
    est.wait()
    got = out.cpu().numpy()
    for k in range(3):
        for a in range(3):
            assert np.array_equal(got[a, k], ref[k][a].reshape(nb)), (k, a)


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_graph_replays_equal_eager(prec):
    """use_graph: the first call runs eagerly and captures the launch
    sequence; later calls with the same buffers replay it.  Every replay,
    with other eager work and synchronisation in between, gives the eager
    bits."""
    torch = pytest.importorskip("torch")
    g = config1()
    plan = B.Plan(g)
    f1 = torch.from_numpy(B.make_field(g.NN, NRi=1)).to("cuda")
    f2 = torch.from_numpy(B.make_field(g.NN, NRi=2)).to("cuda")
    nb = g.nbins * 16
    eager = B.Estimator(plan, prec)
    ref = {}
    for k, f in ((1, f1), (2, f2)):
        o = torch.zeros(3, nb, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()   # the estimator runs on its own stream
        eager.estimate_device(f.data_ptr(), o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr())
        ref[k] = o.cpu().numpy()
    est = B.Estimator(plan, prec, use_graph=True)
    out = torch.zeros(3, nb, dtype=torch.float64, device="cuda")
    other = torch.zeros(3, nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # the estimator runs on its own stream
    args = (f1.data_ptr(), out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr())
    for rnd in range(2):
        for rep in range(4):
            out.zero_()
            torch.cuda.synchronize()   # the estimator runs on its own stream
            est.estimate_device(*args)      # rep 0 eager + capture, then replays
            assert np.array_equal(out.cpu().numpy(), ref[1]), (rnd, rep)
            t = est.timings()
            assert t[4] > 0 and (rep == 0 or t[2] == 0.0)
            other.zero_()                   # other eager work in between
            torch.cuda.synchronize()
        # different buffers: eager again, and a new capture for them
        est.estimate_device(f2.data_ptr(), other[0].data_ptr(), other[1].data_ptr(),
                            other[2].data_ptr())
        assert np.array_equal(other.cpu().numpy(), ref[2]), rnd
    # host path: the estimator's own buffers, replayed after the first call
    host = B.Estimator(plan, prec, use_graph=True)
    hf = B.make_field(g.NN, NRi=1)
    for rep in range(3):
        bs, nm, bi = host.estimate(hf)
        assert np.array_equal(bs.ravel(), ref[1][0]) and np.array_equal(nm.ravel(), ref[1][1])


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
@pytest.mark.parametrize("NN", [192, 256])
def test_batch_equals_single_at_batching_sizes(NN, prec):
    """The batching workload's geometries (Nmax = NN/4, Nr = round(NN/10),
    UL = floor(NN/7)): four device fields in one batch get the bins each
    gets alone, bit for bit."""
    import torch
    g = B.Geometry(NN=NN, UL_k1=NN // 7, LL_k1=1, BW_k1=1, Nmax=NN // 4,
                   Nr=int(round(NN / 10)), n_mtbin=5)
    plan = B.Plan(g)
    nb = g.nbins * 16
    fields = torch.stack([B.make_field_device(NN, NRi=k, fnl=0.3) for k in range(1, 5)])
    out = torch.zeros(3, 4 * nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # the estimator runs on its own stream
    batched = B.Estimator(plan, prec, max_batch=4)
    batched.estimate_batch_device(4, fields.data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                                  out[2].data_ptr())
    got = out.cpu().numpy().reshape(3, 4, nb)
    single = B.Estimator(plan, prec)
    one = torch.zeros(3 * nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # the estimator runs on its own stream
    p = one.data_ptr()
    for k in range(4):
        torch.cuda.synchronize()
        batched.estimate_device(fields[k].data_ptr(), p, p + 8 * nb, p + 16 * nb)
        alone = one.cpu().numpy().reshape(3, nb)
        for a in range(3):
            assert np.array_equal(got[a, k], alone[a]), (k, a)
        single.estimate_device(fields[k].data_ptr(), p, p + 8 * nb, p + 16 * nb)
        other = one.cpu().numpy().reshape(3, nb)
        close_per_multipole(got[0, k], other[0], CROSS_TOL[prec])
        assert np.array_equal(got[1, k], other[1])
    assert np.abs(got[0]).max() > 0


def test_batch_of_1024_fields_on_the_nn_grid():
    """max_batch 1024 at NN=64 on the NN grid: the plane gather's launch
    rows (64 per field) exceed one launch's 65535, so the batch is gathered by
    two launches; every field's bins equal the single-field estimate's on the
    same estimator, and those of a one-field estimator to FP64 rounding."""
    torch = pytest.importorskip("torch")
    g = B.Geometry(NN=64, UL_k1=10, LL_k1=1, BW_k1=1, Nmax=16, Nr=8, ring_rule=B.RINGS_UNIFORM,
                   ring_width=1.0, last_ring_to_nmax=0)
    plan = B.Plan(g)
    nf = 1024
    f = [B.make_field(g.NN, NRi=k) for k in (1, 2, 3)]
    est = B.Estimator(plan, "fp64", full_grid=True, max_batch=nf)
    ref = [est.estimate(x) for x in f]
    for x, r in zip(f, ref):
        close_per_multipole(r[0], B.Estimator(plan, "fp64", full_grid=True).estimate(x)[0], 1e-12)
    assert est.grid == 64 and 64 * nf > 65535
    d_f = torch.from_numpy(np.stack([f[k % 3] for k in range(nf)])).to("cuda")
    nb = g.nbins * 16
    out = torch.zeros(3, nf, nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # the estimator runs on its own stream
    est.estimate_batch_device(nf, d_f.data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                              out[2].data_ptr())
    got = out.cpu().numpy()
    for k in (0, 1, 2, 1021, 1022, 1023):
        for a in range(3):
            assert np.array_equal(got[a, k], ref[k % 3][a].reshape(-1)), (k, a)
    est.close()


@pytest.mark.parametrize("full_grid", [False, True])
def test_host_field_upload_reads_only_the_modes_it_needs(full_grid):
    """bb_estimate with a host field uploads only kz < Nmax, |kx|, |ky| <=
    kmax: the modes outside that box may hold anything (here NaN) and the
    bins are the device path's bits for the clean field; batches too."""
    torch = pytest.importorskip("torch")
    g = config1()
    plan = B.Plan(g)
    est = B.Estimator(plan, "fp64", full_grid=full_grid, max_batch=2)
    f = B.make_field(g.NN, NRi=6)
    d = torch.from_numpy(f).to("cuda")
    nb = g.nbins * 16
    out = torch.zeros(3, nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    est.estimate_device(d.data_ptr(), out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr())
    ref = out.cpu().numpy()
    lo, hi, av = plan.ring_edges()
    km = int(np.floor(hi.max()))
    dirty = f.copy().view(np.complex128)
    ia = np.fft.fftfreq(g.NN, 1.0 / g.NN).astype(int)
    outside = (np.abs(ia)[:, None, None] > km) | (np.abs(ia)[None, :, None] > km) | \
              (np.arange(g.NN // 2 + 1)[None, None, :] >= g.Nmax)
    dirty[np.broadcast_to(outside, dirty.shape)] = np.nan
    dirty = dirty.view(np.float64)
    got = est.estimate(dirty)
    for k in range(3):
        assert np.array_equal(got[k].ravel(), ref[k]), k
    two = est.estimate_batch(np.stack([dirty, f]))
    for k in range(3):
        assert np.array_equal(two[k][0].ravel(), ref[k]) and np.array_equal(two[k][1].ravel(), ref[k])


def test_mixed_geometries_in_one_launch():
    """bb_estimate_mixed: fields of three geometries (NN 48, 64, 128) with
    their contractions in one persistent launch give each field the bits its
    own estimator gives it; estimators of another precision or layout shape,
    or one estimator twice, are refused."""
    torch = pytest.importorskip("torch")
    geoms = [medium(), B.Geometry(NN=64, UL_k1=10, LL_k1=1, BW_k1=1, Nmax=16, Nr=8,
                                  ring_rule=B.RINGS_UNIFORM, ring_width=1.0, last_ring_to_nmax=0),
             config1()]
    nfs = [3, 2, 2]
    opts = dict(waves=4, occupancy=3)
    ests, fields, outs, refs = [], [], [], []
    for g, nf in zip(geoms, nfs):
        e = B.Estimator(B.Plan(g), "bf16x3", max_batch=nf, **opts)
        f = torch.from_numpy(np.stack([B.make_field(g.NN, NRi=k + 1) for k in range(nf)])).to("cuda")
        nb = g.nbins * 16
        o = torch.zeros(3, nf, nb, dtype=torch.float64, device="cuda")
        r = torch.zeros(3, nf, nb, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        e.estimate_batch_device(nf, f.data_ptr(), r[0].data_ptr(), r[1].data_ptr(), r[2].data_ptr())
        ests.append(e)
        fields.append(f)
        outs.append(o)
        refs.append(r.cpu().numpy())
    for rep in range(2):   # the second call reuses the uploaded unit list
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        B.estimate_mixed(ests, nfs, [f.data_ptr() for f in fields], [o[0].data_ptr() for o in outs],
                         [o[1].data_ptr() for o in outs], [o[2].data_ptr() for o in outs])
        for o, r in zip(outs, refs):
            assert np.array_equal(o.cpu().numpy(), r), rep
    # A replaced estimator (possibly at the old one's address) gets a fresh
    # unit list: the same call shape must not reuse the old device addresses.
    ests[1].close()
    ests[1] = B.Estimator(B.Plan(geoms[1]), "bf16x3", max_batch=nfs[1], **opts)
    torch.cuda.synchronize()
    for o in outs:
        o.zero_()
    torch.cuda.synchronize()
    B.estimate_mixed(ests, nfs, [f.data_ptr() for f in fields], [o[0].data_ptr() for o in outs],
                     [o[1].data_ptr() for o in outs], [o[2].data_ptr() for o in outs])
    for o, r in zip(outs, refs):
        assert np.array_equal(o.cpu().numpy(), r), "after replacing an estimator"
    e64 = B.Estimator(B.Plan(geoms[0]), "fp64")
    args = lambda es: (es, [1] * len(es), [fields[0].data_ptr()] * len(es),
                       [outs[0][0].data_ptr()] * len(es), [outs[0][1].data_ptr()] * len(es),
                       [outs[0][2].data_ptr()] * len(es))
    with pytest.raises(B.BarbieError):
        B.estimate_mixed(*args([ests[0], e64]))
    with pytest.raises(B.BarbieError):
        B.estimate_mixed(*args([ests[0], ests[0]]))
    e8 = B.Estimator(B.Plan(geoms[1]), "bf16x3", waves=8)
    with pytest.raises(B.BarbieError):
        B.estimate_mixed(*args([ests[0], e8]))
