#!/usr/bin/env python3
"""Benchmark: NN=384 field estimates per hour on one MI355X.

Workload (value): research-size estimate, NN=384 seeded non-Gaussian RSD
field (r2c float64 384x384x386), Nmax=100 planes, Nr=40 piecewise rings,
UL_k1=51, LL_k1=1, BW_k1=1, n_mtbin=5 (2 756 382 contractions per field).
A step is one full estimate of one field already resident in HBM: plane
gather, plane-ring IFFT caches, contraction, binning (bb_estimate_device).

Prints one JSON line.  `value` is the fields/hour of the split path
(--precision, default bf16x3); the FP64 and FP32 GPU paths are timed in the
same run and reported beside it, with the per-multipole error of the split
and FP32 paths against FP64 at fNL-style coefficients 0.1, 0.3 and 1.0.
`roofline` is the contraction kernel on the work it performs (the
reference's NN-grid work beside it); `host_field_path` is the rate with the
field in host memory (upload included); `cpu_baseline` is the CPU oracle on
a bounded sample of the same workload (one thread, all threads, and the
reference's on-disk pattern); `c1` times the NN=128 CPU configuration in
full.  Further legs: `batched` (64 realizations per NN at NN = 128, 192,
256, batched launches) and `serving` / `serving_fp64` (request server,
Poisson arrivals over an NN = 128/256/384/512 mix weighted 4:3:2:1, p50/p99
latency at 50/80/95% of the measured saturation rate).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15 / 3}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="bf16x3", choices=["fp32", "bf16x3", "fp64"])
    ap.add_argument("--full-grid", action="store_true",
                    help="run the transforms and sums on the NN grid (the reference's grid)")
    ap.add_argument("--NN", type=int, default=384)
    ap.add_argument("--skip-fp64", action="store_true",
                    help="time only --precision (no FP64 and FP32 legs)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target CPU time of the single-thread oracle sample (0 disables it)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the fNL 0.1 / 1.0 fields")
    ap.add_argument("--no-power", action="store_true",
                    help="skip the power-spectrum multipoles (calpow_mom) leg")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the default layout instead of the autotuned one (bb_tune)")
    ap.add_argument("--layout", default="",
                    help="fixed layout options for --precision instead of tuning, e.g. "
                         "'grid_levels=4,waves=4,mix_fp32=0' (as bb_tune reports them)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a reference point for "
                         "HBM byte counters)")
    return ap.parse_args()


def geometry(B, NN):
    """NN=384: the reference geometry.  Other sizes: Nmax = NN/4,
    Nr = round(NN/10), UL = floor(NN/7), reference ring rule truncated at Nr;
    NN=512: Nmax = 134, Nr = 54 uniform rings of width Nmax/Nr, UL = 73."""
    if NN == 384:
        return B.Geometry(NN=384, UL_k1=51, LL_k1=1, BW_k1=1, Nmax=100, Nr=40, n_mtbin=5)
    if NN == 512:
        return B.Geometry(NN=512, UL_k1=73, LL_k1=1, BW_k1=1, Nmax=134, Nr=54, n_mtbin=5,
                          ring_rule=B.RINGS_UNIFORM, ring_width=134 / 54, last_ring_to_nmax=1)
    return B.Geometry(NN=NN, UL_k1=NN // 7, LL_k1=1, BW_k1=1, Nmax=NN // 4,
                      Nr=int(round(NN / 10)), n_mtbin=5)


def time_path(B, torch, plan, precision, d_field, steps, warmup, full_grid=False, opts=None):
    est = B.Estimator(plan, precision, full_grid=full_grid, **(opts or {}))
    wl = est.worklist()
    levels = est.levels
    nb = plan.geom.nbins * 16
    out = torch.zeros(3 * nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # the estimator runs on its own stream
    ptr = out.data_ptr()
    args = (d_field.data_ptr(), ptr, ptr + 8 * nb, ptr + 16 * nb)
    for _ in range(warmup):
        est.estimate_device(*args)
    torch.cuda.synchronize()
    # Timed: the steps are enqueued back to back (each runs after the
    # previous on the estimator's stream), as a caller with a stream of
    # fields would; one wait at the end.
    t0 = time.perf_counter()
    for _ in range(steps):
        est.estimate_batch_device(1, *args, wait=False)
    est.wait()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # Stage times (HIP events on the estimator's stream) from the same
    # number of estimates run one at a time.
    stage = []
    for _ in range(steps):
        est.estimate_device(*args)
        stage.append(est.timings())
    import numpy as np
    st = np.mean(np.array(stage), axis=0)
    bins = out.cpu().numpy()
    est.close()
    return (t1 - t0) * 1e3 / steps, st, bins, levels, wl


def unsupported_entries(bins, nbins):
    """(bin, multipole) entries of the written rows (monopole denominator
    > .1, barbie.c:392) whose own denominator is not positive: what
    bb_write_outputs_checked writes as nan."""
    import numpy as np
    nm = bins[nbins * 16:2 * nbins * 16].reshape(-1, 16)
    rows = nm[:, 0] > .1
    return int((rows[:, None] & ~(nm > 0)).sum())


def per_multipole_err(bins, ref, nb):
    import numpy as np
    a = bins[:nb].reshape(-1, 16)
    b = ref[:nb].reshape(-1, 16)
    return [float(np.linalg.norm(a[:, i] - b[:, i]) / max(np.linalg.norm(b[:, i]), 1e-300))
            for i in range(16)]


def estimate_once(B, torch, plan, precision, d_field):
    est = B.Estimator(plan, precision)
    nb = plan.geom.nbins * 16
    out = torch.zeros(3 * nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # the estimator runs on its own stream
    ptr = out.data_ptr()
    est.estimate_device(d_field.data_ptr(), ptr, ptr + 8 * nb, ptr + 16 * nb)
    r = out.cpu().numpy()
    est.close()
    return r


def accuracy_leg(B, torch, plan, g, precision):
    """Split path against the FP64 GPU path for fNL-style 0.1 and 1.0."""
    out = {}
    nb = g.nbins * 16
    for fnl in (0.1, 1.0):
        d = B.make_field_device(g.NN, NRi=1, fnl=fnl)
        torch.cuda.synchronize()
        a = estimate_once(B, torch, plan, precision, d)
        b = estimate_once(B, torch, plan, "fp64", d)
        out[str(fnl)] = max(per_multipole_err(a, b, nb))
        del d
    return out


def batched_leg(B, torch, precision, nfields=64, distinct=8, reps=3):
    """64 realizations per NN as batched launches, with the layout bb_tune
    picks for the geometry (timed on one of the fields); fields/hour batched
    and one at a time on the same estimator."""
    res = {}
    for NN in (128, 192, 256):
        g = geometry(B, NN)
        plan = B.Plan(g)
        nb = g.nbins * 16
        fsz = NN * NN * (NN + 2)
        base = torch.stack([B.make_field_device(NN, NRi=k + 1) for k in range(distinct)])
        fields = base.repeat(nfields // distinct, 1, 1, 1).contiguous()
        del base
        torch.cuda.synchronize()
        opts, _ = plan.tune(precision, d_field=fields.data_ptr())
        est = B.Estimator(plan, precision, max_batch=nfields, **opts)
        out = torch.zeros(3, nfields, nb, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()   # the estimator runs on its own stream
        ptrs = (fields.data_ptr(), out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr())
        est.estimate_batch_device(nfields, *ptrs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            est.estimate_batch_device(nfields, *ptrs)
        torch.cuda.synchronize()
        tb = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for k in range(nfields):
            est.estimate_batch_device(1, fields.data_ptr() + 8 * fsz * k, ptrs[1], ptrs[2],
                                      ptrs[3])
        torch.cuda.synchronize()
        t1 = time.perf_counter() - t0
        res[str(NN)] = {"fields_per_hour_batched": nfields * 3600.0 / tb,
                        "fields_per_hour_one_at_a_time": nfields * 3600.0 / t1,
                        "batch": nfields, "ms_per_batch": tb * 1e3, "layout": opts,
                        "ncontr": plan.stats["ncontr"],
                        "realizations": f"{distinct} seeded fields x {nfields // distinct} "
                                        "copies (cost does not depend on the values)"}
        est.close()
        del fields, out
        torch.cuda.empty_cache()
    return res


def mixed_leg(B, torch, counts=(22, 21, 21), reps=3):
    """Seeded realizations of mixed NN in {128, 192, 256} (counts per size;
    the C4 workload is 64 fields as 22/21/21) in one call: the contractions
    of all of them in one persistent launch (bb_estimate_mixed, BF16x3),
    against the same fields as one batched call per NN, one after the other.
    Fields/hour of both, and whether every field's bins are the same bits
    both ways."""
    import numpy as np
    sizes = (128, 192, 256)
    ests, fields, outs = [], [], []
    for NN, nf in zip(sizes, counts):
        g = geometry(B, NN)
        plan = B.Plan(g)
        ests.append(B.Estimator(plan, "bf16x3", max_batch=nf, waves=4, occupancy=3,
                                grid_levels=4))
        f = torch.empty((nf, NN, NN, NN + 2), dtype=torch.float64, device="cuda")
        for k in range(nf):
            B.generate_field(NN, B.seed_for(k + 1), out=f[k].data_ptr())
        fields.append(f)
        outs.append(torch.zeros(2, 3, nf, g.nbins * 16, dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()

    def separate(o):
        for e, f, out, nf in zip(ests, fields, outs, counts):
            e.estimate_batch_device(nf, f.data_ptr(), out[o, 0].data_ptr(), out[o, 1].data_ptr(),
                                    out[o, 2].data_ptr())

    def mixed(o):
        B.estimate_mixed(ests, counts, [f.data_ptr() for f in fields],
                         [x[o, 0].data_ptr() for x in outs], [x[o, 1].data_ptr() for x in outs],
                         [x[o, 2].data_ptr() for x in outs])

    res = {}
    for name, fn, o in (("separate", separate, 0), ("mixed", mixed, 1)):
        fn(o)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn(o)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[name] = {"fields_per_hour": sum(counts) * 3600.0 / dt, "ms_per_call": dt * 1e3}
    same = all(bool(torch.equal(x[0], x[1])) for x in outs)
    for e in ests:
        e.close()
    return {"fields": dict(zip([str(n) for n in sizes], counts)),
            "layout": "grid_levels=4,waves=4,occupancy=3 (one shape for all sizes)",
            "separate_per_NN_batches": res["separate"], "mixed_one_launch": res["mixed"],
            "speedup": res["separate"]["ms_per_call"] / res["mixed"]["ms_per_call"],
            "bins_identical": same}


def serving_leg(B, torch, precision, n_sat=200, n_load=300, seed=7):
    """Request server over the NN mix 128/256/384/512 (weights 4:3:2:1), one
    device-resident field per NN.  Saturation: a burst of n_sat requests.
    Then Poisson arrivals at 50/80/95% of that rate, n_load requests each."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes, weights = [128, 256, 384, 512], np.array([4, 3, 2, 1], dtype=float)
    geoms = {NN: geometry(B, NN) for NN in sizes}
    fields = {NN: B.make_field_device(NN, NRi=1) for NN in sizes}
    torch.cuda.synchronize()
    srv = B.Server(precision, max_inflight=4, max_batch=4, tune=True)
    t0 = time.perf_counter()
    for NN in sizes:   # plans, autotuning and lanes before timing
        srv.prepare(geoms[NN], 4)
        srv.wait(srv.submit(geoms[NN], fields[NN].data_ptr()))
    setup_s = time.perf_counter() - t0
    alone = {}
    for NN in sizes:   # one request at a time: service time per size
        lat = [srv.wait(srv.submit(geoms[NN], fields[NN].data_ptr()))[3][2] for _ in range(3)]
        alone[str(NN)] = float(np.median(lat))

    def draw(n):
        return rng.choice(sizes, size=n, p=weights / weights.sum())

    mix = draw(n_sat)
    srv.pause(True)
    tickets = [srv.submit(geoms[NN], fields[NN].data_ptr()) for NN in mix]
    t0 = time.perf_counter()
    srv.pause(False)
    for t in tickets:
        srv.wait(t)
    sat = n_sat / (time.perf_counter() - t0)
    out = {"saturation_req_per_s": sat, "mix": "NN 128/256/384/512 weighted 4:3:2:1",
           "max_inflight": 4, "max_batch": 4, "setup_seconds": setup_s,
           "alone_ms_by_NN": alone, "loads": {}}
    for load in (0.5, 0.8, 0.95):
        lam = load * sat
        mix = draw(n_load)
        gaps = rng.exponential(1.0 / lam, size=n_load)
        tickets = []
        t_next = time.perf_counter()
        for NN, gap in zip(mix, gaps):
            t_next += gap
            while True:
                dt = t_next - time.perf_counter()
                if dt <= 0:
                    break
                if dt > 0.0005:
                    time.sleep(dt - 0.0004)
            tickets.append(srv.submit(geoms[NN], fields[NN].data_ptr()))
        lat = np.array([srv.wait(t)[3][2] for t in tickets])
        by_nn = {str(NN): {"p50_ms": float(np.percentile(lat[mix == NN], 50)),
                           "p99_ms": float(np.percentile(lat[mix == NN], 99))}
                 for NN in sizes if np.any(mix == NN)}
        out["loads"][str(load)] = {"rate_req_per_s": lam, "p50_ms": float(np.percentile(lat, 50)),
                                   "p99_ms": float(np.percentile(lat, 99)),
                                   "mean_ms": float(lat.mean()), "by_NN": by_nn}
    st = srv.stats()
    out["batches"] = st["batches"]
    out["batched_fields"] = st["batched_fields"]
    out["lanes_created"] = st["lanes_created"]
    out["lanes_evicted"] = st["lanes_evicted"]
    out["lane_bytes"] = st["bytes"]
    srv.close()
    return out


def power_leg(B, torch, d_field, NN, reps=20):
    """Power-spectrum multipoles (calpow_mom) of the same field, device
    resident: wall time per call (launches, the field pass, the sum over
    workgroups and the copy of 40 sums back; the geometry pass of the first
    call per (NN, bins) is outside the timed calls) and the HBM rate of the
    algorithmic bytes, 16 B per summed mode.  Returns the leg's record and
    the rows of the last call (for the CPU-baseline leg's comparison)."""
    ptr = d_field.data_ptr()
    B.power_multipoles(ptr, NN)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        got = B.power_multipoles(ptr, NN)
    dt = (time.perf_counter() - t0) / reps
    nbytes = 16.0 * float(got[:, 5].sum())
    return ({"ms_per_field": dt * 1e3, "fields_per_s": 1.0 / dt,
             "summed_modes": float(got[:, 5].sum()), "bytes_algorithmic": nbytes,
             "hbm_GBps": nbytes / dt / 1e9, "hbm_frac": nbytes / dt / 8e12,
             "note": "wall time per bb_power_multipoles call on a device field"}, got)


def cpu_power_baseline(field, NN, got):
    """The CPU oracle of calpow_mom on the same field, one host thread, and
    its agreement with the GPU rows (mode counts exact, monopole)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    t0 = time.perf_counter()
    ref = O.calpow_mom(field, NN)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "fields/s", "cores": 1, "kind": "port",
            "sample": f"calpow_mom of the NN={NN} field (whole workload) in {dt * 1e3:.0f} ms",
            "num_equal_to_gpu": bool(np.array_equal(got[:, 5], ref[:, 5])),
            "max_rel_diff_P0_gpu": float(np.max(np.abs(got[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])))}


def host_info():
    """The host CPU model and the threads this process may use."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "os_cpu_count": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)), "threads_used": cpu_threads()}


def cpu_threads():
    """All cores of the host share this run may use (at most 16)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def oracle_setup(plan):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    g = plan.geom
    lo, hi, av = plan.ring_edges()
    t0 = time.perf_counter()
    rd = O.calcr(g.NN, lo, hi, av)
    t_calcr = time.perf_counter() - t0
    params = {kk: getattr(g, kk) for kk in ("NN", "Nmax", "Nr", "n_mtbin", "UL_k1", "LL_k1",
                                            "BW_k1", "LL", "ff", "beta")}
    return O, params, (lo, hi, av), rd, t_calcr


def fields_read(plan, p1_begin, p1_end):
    """[Nmax, Nr] mask of the plane-ring fields the reference's loop reads for
    p1 in [p1_begin, p1_end): buffer1 = D_{p1,r1} for every ring triple
    (barbie.c:274-277, read whether or not the triple has a contraction at
    this p1) and the second and third fields of each contraction."""
    import numpy as np
    g = plan.geom
    pc = plan.contractions()
    sel = pc[(pc[:, 0] >= p1_begin) & (pc[:, 0] < p1_end)]
    want = np.zeros((g.Nmax, g.Nr), dtype=np.uint8)
    r1 = np.unique(plan.triples()[:, 0])
    for p1 in range(p1_begin, p1_end):
        want[p1, r1] = 1
    for pcol, rcol in ((2, 3), (4, 5)):
        want[np.abs(sel[:, pcol]), sel[:, rcol]] = 1
    return want


def disk_leg(O, params, edges, rd, field, plan, p1_begin, p1_end, want, tmpdir):
    """The reference's own access pattern: the fields written to disk by
    wrapid_fast with 4 FFT threads (barbie.c:68,151-153, funcs.c:537-566),
    then the loop reading buffer1 per ring triple and buffer2/buffer3 per
    contraction from those files, one thread (barbie.c:274-324).  Returns
    (seconds writing, files written, seconds in the loop, contractions)."""
    import numpy as np
    g = plan.geom
    lo, hi, av = edges
    t0 = time.perf_counter()
    nw = O.write_plane_cache(field, g.NN, g.Nmax, lo, hi, tmpdir, 1, nthreads=4,
                             want=None if want is None else want)
    t_w = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, _, _, n, _ = O.estimate_disk(params, lo, hi, av, rd, tmpdir, 1, p1_begin=p1_begin,
                                    p1_end=p1_end)
    t_l = time.perf_counter() - t0
    return t_w, nw, t_l, n


def cpu_baseline(B, plan, field, target_s):
    """CPU oracle (FP64, own FFT) on a bounded sample of the NN=384 workload,
    extrapolated by contraction count:
    * `value`: the in-memory oracle on one thread over the p1 prefix whose
      contraction count makes about target_s of work;
    * `all_cores`: the same sample with the ring triples of each p1 spread
      over all threads of the host share;
    * `reference_faithful`: the reference's pattern (on-disk plane-ring
      cache built with 4 FFT threads, single-thread loop reading two cache
      files per contraction) on the slice of the last p1 plane with
      contractions, the cache build scaled by files (Nmax x Nr are written
      per field) and the loop by contractions."""
    import numpy as np
    import tempfile
    import shutil
    O, params, edges, rd, t_calcr = oracle_setup(plan)
    lo, hi, av = edges
    g = plan.geom
    pc = plan.contractions()
    counts = np.bincount(pc[:, 0], minlength=g.Nmax)
    ncontr = len(pc)
    per = 3.0e-9 * g.NN * g.NN   # ~0.45 ms per contraction at NN=384 on one core
    want = max(1, int(target_s / per))
    k = int(np.searchsorted(np.cumsum(counts), want)) + 1
    k = max(1, min(k, g.Nmax))
    t0 = time.perf_counter()
    _, _, _, n, _ = O.estimate(params, lo, hi, av, rd, field, p1_begin=0, p1_end=k, nthreads=1)
    dt = time.perf_counter() - t0
    est_s = dt * ncontr / n + t_calcr
    threads = cpu_threads()
    t0 = time.perf_counter()
    _, _, _, nm, _ = O.estimate(params, lo, hi, av, rd, field, p1_begin=0, p1_end=k,
                                nthreads=threads)
    dtm = time.perf_counter() - t0
    est_m = dtm * ncontr / nm + t_calcr
    out = {"value": 3600.0 / est_s, "unit": "fields/hour", "cores": 1, "kind": "port",
           "sample": f"CPU oracle (FP64, own FFT, in-memory fields, 1 thread) over p1 in [0,{k}): "
                     f"{n} of {ncontr} contractions in {dt:.1f} s, extrapolated by contraction "
                     f"count (+ calcr {t_calcr:.1f} s) to {est_s:.0f} s/field",
           "host": host_info(),
           "all_cores": {"value": 3600.0 / est_m, "cores": threads,
                         "sample": f"same p1 prefix with {threads} OpenMP threads: {nm} "
                                   f"contractions in {dtm:.1f} s, extrapolated to "
                                   f"{est_m:.0f} s/field"}}
    nzp = np.nonzero(counts)[0]
    p1s = int(nzp.max())
    mask = fields_read(plan, p1s, p1s + 1)
    tmp = tempfile.mkdtemp(prefix="barbie_ref_cache_")
    try:
        t_w, nw, t_l, nd = disk_leg(O, params, edges, rd, field, plan, p1s, p1s + 1, mask, tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    est_d = t_w * (g.Nmax * g.Nr) / nw + t_l * ncontr / nd + t_calcr
    out["reference_faithful"] = {
        "value": 3600.0 / est_d, "unit": "fields/hour", "cores": 4,
        "sample": f"p1 = {p1s}: {nw} cache files written with 4 FFT threads in {t_w:.1f} s "
                  f"(scaled to the {g.Nmax * g.Nr} files of a field), {nd} contractions reading "
                  f"their cache files in {t_l:.1f} s on 1 thread (scaled to {ncontr}); "
                  f"{est_d:.0f} s/field",
        "seconds_per_field": est_d}
    return out


def c1_leg(B, torch):
    """The CPU configuration, NN=128 (Nmax=32, Nr=12 rings of width 1, UL_k1=21),
    timed in full: the oracle on one thread, on all threads of the host share
    and in the reference's on-disk pattern; the FP64 GPU estimate of the same
    field beside it, with its largest bin difference from the oracle."""
    import numpy as np
    import tempfile
    import shutil
    g = B.Geometry(NN=128, UL_k1=21, LL_k1=1, BW_k1=1, Nmax=32, Nr=12, n_mtbin=5,
                   ring_rule=B.RINGS_UNIFORM, ring_width=1.0, last_ring_to_nmax=0)
    plan = B.Plan(g)
    field = B.make_field(g.NN, NRi=1, fnl=0.3)
    O, params, edges, rd, t_calcr = oracle_setup(plan)
    lo, hi, av = edges
    t0 = time.perf_counter()
    bs1, nm1, _, n1, _ = O.estimate(params, lo, hi, av, rd, field, nthreads=1)
    t1 = time.perf_counter() - t0 + t_calcr
    threads = cpu_threads()
    t0 = time.perf_counter()
    O.estimate(params, lo, hi, av, rd, field, nthreads=threads)
    tm = time.perf_counter() - t0 + t_calcr
    tmp = tempfile.mkdtemp(prefix="barbie_ref_cache_")
    try:
        t_w, nw, t_l, nd = disk_leg(O, params, edges, rd, field, plan, 0, g.Nmax, None, tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    td = t_w + t_l + t_calcr
    est = B.Estimator(plan, "fp64")
    bs, nm, _ = est.estimate(field)
    d_f = torch.from_numpy(field).to("cuda")
    nb = g.nbins * 16
    out = torch.zeros(3 * nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ptr = out.data_ptr()
    est.estimate_device(d_f.data_ptr(), ptr, ptr + 8 * nb, ptr + 16 * nb)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        est.estimate_device(d_f.data_ptr(), ptr, ptr + 8 * nb, ptr + 16 * nb)
    tg = (time.perf_counter() - t0) / reps
    est.close()
    err = float(np.max(np.abs(bs.ravel() - bs1)) / np.max(np.abs(bs1)))
    return {"workload": "estimate NN=128 Nmax=32 Nr=12 (width 1) UL_k1=21 LL_k1=1 BW_k1=1",
            "ncontr": int(n1),
            "cpu_fields_per_hour": {"oracle_1_thread": 3600.0 / t1,
                                    f"oracle_{threads}_threads": 3600.0 / tm,
                                    "reference_faithful": 3600.0 / td},
            "cpu_seconds": {"oracle_1_thread": t1, f"oracle_{threads}_threads": tm,
                            "reference_faithful": td,
                            "reference_faithful_parts": {"calcr": t_calcr, "cache_write": t_w,
                                                         "loop_from_disk": t_l}},
            "gpu_fp64_fields_per_hour": 3600.0 / tg, "gpu_fp64_ms": tg * 1e3,
            "gpu_fp64_max_rel_diff_bins_vs_oracle": err,
            "note": "CPU times are whole-field (no extrapolation); reference_faithful writes "
                    f"{nw} dataPlaneRing files with 4 FFT threads, then reads 2 per "
                    "contraction on 1 thread"}


def host_path_rate(B, plan, precision, opts, field, reps=5):
    """Fields/hour through bb_estimate with the field in pageable host memory
    (the CLI's path): upload included."""
    est = B.Estimator(plan, precision, **(opts or {}))
    est.estimate(field)
    t0 = time.perf_counter()
    for _ in range(reps):
        est.estimate(field)
    dt = (time.perf_counter() - t0) / reps
    est.close()
    return {"fields_per_hour": 3600.0 / dt, "ms_per_field": dt * 1e3,
            "note": "bb_estimate with the r2c field in pageable host memory; the library "
                    "uploads what the estimate reads"}


def traffic_from_profile(precision, NN):
    """HBM bytes per contraction launch, from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE counters recorded under profiles/ (not measured by this run)."""
    prof = os.path.join(ROOT, "profiles", f"contraction_{precision}_nn{NN}.txt")
    if os.path.exists(prof):
        for line in open(prof):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1]), os.path.relpath(prof, ROOT)
    return None, None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {a.gpus}); the estimator "
              "has no multi-GPU path to measure.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import barbie_amd as B

    torch.cuda.set_device(0)
    g = geometry(B, a.NN)
    plan = B.Plan(g)
    s = plan.stats
    field = B.make_field(g.NN, NRi=1, fnl=0.3)
    d_field = torch.from_numpy(field).to("cuda")

    if a.copy_calibration:
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    others = [] if a.skip_fp64 else [p for p in ("fp64", "fp32") if p != a.precision]
    tuned = {}
    for prec in [a.precision] + others:
        if a.layout and prec == a.precision:
            kv = dict(x.split("=") for x in a.layout.split(","))
            tuned[prec] = ({k: int(v) for k, v in kv.items()}, None, 0.0)
        elif a.no_tune:
            tuned[prec] = ({}, None, 0.0)
        else:
            t0 = time.perf_counter()
            opts, tms = plan.tune(prec, full_grid=a.full_grid, d_field=d_field.data_ptr())
            tuned[prec] = (opts, tms, time.perf_counter() - t0)
    ms, st, bins, levels, wl = time_path(B, torch, plan, a.precision, d_field, a.steps,
                                         a.warmup, a.full_grid, tuned[a.precision][0])
    res = {}
    for prec in others:
        res[prec] = time_path(B, torch, plan, prec, d_field, a.steps, a.warmup, a.full_grid,
                              tuned[prec][0])

    # Contraction work per field.  Performed: 4 flops (two real FMAs of
    # Re(P C)) per point of each contraction's own grid, 4 * sum_c NG(c)^2;
    # the reference's: the same over the NN grid, 4 * NN^2 * ncontr (the grid
    # levels give the same sums exactly on fewer points).
    w_ref = 4.0 * g.NN * g.NN * s["ncontr"]
    w_done = 4.0 * wl["points_summed"]
    t_contr = st[2] * 1e-3
    peak = PEAK[a.precision]
    value = 3600e3 / ms
    traffic, traffic_src = traffic_from_profile(a.precision, g.NN)
    out = {
        "metric": "NN=384 field estimates/hour, split-FP32 vs FP64 GPU; p50/p99 request latency",
        "value": value,
        "unit": "fields/hour",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "fp32", "bf16x3": "bf16x3+fp32acc", "fp64": "fp64"}[a.precision],
        "data": "synthetic (seeded Gaussian P(k)=1e3 k^-1.5, fNL-style 0.3, Kaiser beta=1)",
        "config": {"workload": f"estimate NN={g.NN} Nmax={g.Nmax} Nr={g.Nr} UL_k1={g.UL_k1} "
                               f"LL_k1={g.LL_k1} BW_k1={g.BW_k1} n_mtbin={g.n_mtbin}",
                   "ncontr": s["ncontr"], "fields_used": s["nfields_used"],
                   "precision": a.precision, "grid_levels": levels},
        "stages_ms": {"gather": st[0], "fft_cache": st[1], "contraction": st[2],
                      "binning": st[3]},
        "tuned": {p: {"options": o, "tune_ms_per_estimate": m, "tune_seconds": ts}
                  for p, (o, m, ts) in tuned.items()},
        "worklist": wl,
        "multipoles_without_denominator": unsupported_entries(bins, g.nbins),
        "roofline": {"bound": "mfma", "kernel": "contraction",
                     "achieved": w_done / t_contr / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": w_done / t_contr / peak, "traffic": traffic,
                     "traffic_source": (f"{traffic_src}: rocprofv3 --pmc FETCH_SIZE x2 + "
                                        "WRITE_SIZE of bench.py (not measured by this run)")
                                       if traffic_src else None,
                     "achieved_reference_work": w_ref / t_contr / 1e12,
                     "work": f"performed W = 4 * sum_c NG(c)^2 = {w_done:.4g} flop per launch "
                             f"(grid levels {levels}); the reference's NN-grid work "
                             f"4 * NN^2 * ncontr = {w_ref:.4g} (the same sums) gives "
                             "achieved_reference_work; bf16x3 peak taken as 2.5 PF/3 "
                             "(three bf16 products per FP32 product)",
                     "kernel_ms": st[2]},
    }
    nb = g.nbins * 16
    for prec, (msp, stp, binsp, levp, wlp) in res.items():
        out[prec] = {"fields_per_hour": 3600e3 / msp, "ms_per_step": msp,
                     "stages_ms": {"gather": stp[0], "fft_cache": stp[1],
                                   "contraction": stp[2], "binning": stp[3]},
                     "grid_levels": levp,
                     "contraction_roofline": {
                         "achieved": 4.0 * wlp["points_summed"] / (stp[2] * 1e-3) / 1e12,
                         "peak": PEAK[prec] / 1e12, "unit": "TFLOP/s",
                         "frac": 4.0 * wlp["points_summed"] / (stp[2] * 1e-3) / PEAK[prec],
                         "achieved_reference_work": w_ref / (stp[2] * 1e-3) / 1e12},
                     "worklist": wlp}
    if "fp64" in res:
        bins64 = res["fp64"][2]
        out["speedup_vs_fp64"] = res["fp64"][0] / ms
        out["rel_err_vs_fp64_per_multipole"] = per_multipole_err(bins, bins64, nb)
        if "fp32" in res:
            out["fp32"]["speedup_vs_fp64"] = res["fp64"][0] / res["fp32"][0]
            out["fp32"]["rel_err_vs_fp64_per_multipole"] = per_multipole_err(res["fp32"][2],
                                                                             bins64, nb)
        if not a.no_accuracy:
            acc = accuracy_leg(B, torch, plan, g, a.precision)
            acc["0.3"] = max(out["rel_err_vs_fp64_per_multipole"])
            out["max_rel_err_vs_fp64_by_fnl"] = acc
            if "fp32" in res:
                acc32 = accuracy_leg(B, torch, plan, g, "fp32")
                acc32["0.3"] = max(out["fp32"]["rel_err_vs_fp64_per_multipole"])
                out["fp32"]["max_rel_err_vs_fp64_by_fnl"] = acc32
    out["host_field_path"] = host_path_rate(B, plan, a.precision, tuned[a.precision][0], field)
    if not a.no_power:
        out["power_multipoles"], power_rows = power_leg(B, torch, d_field, g.NN)
    del d_field
    torch.cuda.empty_cache()
    if not a.no_batched:
        out["batched"] = batched_leg(B, torch, a.precision)
        if a.precision == "bf16x3":
            out["batched"]["mixed_NN"] = mixed_leg(B, torch)
            # one field of each size: separate launches leave the chip partly idle
            out["batched"]["mixed_NN_one_each"] = mixed_leg(B, torch, counts=(1, 1, 1), reps=10)
    if not a.no_serving:
        sv = serving_leg(B, torch, a.precision)
        out["serving"] = sv
        out["p50_ms"] = sv["loads"]["0.8"]["p50_ms"]
        out["p99_ms"] = sv["loads"]["0.8"]["p99_ms"]
        if not a.skip_fp64 and a.precision != "fp64":
            out["serving_fp64"] = serving_leg(B, torch, "fp64")
    if a.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(B, plan, field, a.cpu_seconds)
        if not a.no_power:
            out["cpu_baseline"]["power_multipoles"] = cpu_power_baseline(field, g.NN, power_rows)
        out["c1"] = c1_leg(B, torch)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
