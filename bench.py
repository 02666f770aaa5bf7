#!/usr/bin/env python3
"""Benchmark: NN=384 field estimates per hour on one MI355X.

Workload: research-size estimate, NN=384 seeded non-Gaussian RSD field
(r2c float64 384x384x386), Nmax=100 planes, Nr=40 piecewise rings,
UL_k1=51, LL_k1=1, BW_k1=1, n_mtbin=5 (2 756 382 contractions per field).
A step is one full estimate of one field already resident in HBM: plane
gather, plane-ring IFFT cache, contraction, binning (bb_estimate_device).

Prints one JSON line.  `value` is the fields/hour of the fast path
(--precision, default fp32); the FP64 GPU path is timed in the same run and
reported beside it.  `roofline` is the contraction kernel; `cpu_baseline`
is the CPU oracle (single thread) on a bounded sample of the same workload,
extrapolated by contraction count.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15 / 3}
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16x3", choices=["fp32", "bf16x3", "fp64"])
    ap.add_argument("--full-grid", action="store_true",
                    help="run the transforms and sums on the NN grid instead of the compact grid")
    ap.add_argument("--NN", type=int, default=384)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target CPU time of the oracle sample (0 disables it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a reference point for "
                         "HBM byte counters)")
    return ap.parse_args()


def geometry(NN):
    import barbie_amd as B
    if NN == 384:
        return B.Geometry(NN=384, UL_k1=51, LL_k1=1, BW_k1=1, Nmax=100, Nr=40, n_mtbin=5)
    # other sizes: Nmax = NN/4, Nr = round(NN/10), UL = NN/7 (batched-form rule)
    return B.Geometry(NN=NN, UL_k1=NN // 7, LL_k1=1, BW_k1=1, Nmax=NN // 4,
                      Nr=int(round(NN / 10)), n_mtbin=5)


def time_path(B, torch, plan, precision, d_field, steps, warmup, full_grid=False):
    est = B.Estimator(plan, precision, full_grid=full_grid)
    grid = est.levels
    wl = est.worklist()
    nb = plan.geom.nbins * 16
    out = torch.zeros(3 * nb, dtype=torch.float64, device="cuda")
    ptr = out.data_ptr()
    args = (d_field.data_ptr(), ptr, ptr + 8 * nb, ptr + 16 * nb)
    for _ in range(warmup):
        est.estimate_device(*args)
    torch.cuda.synchronize()
    stage = []
    t0 = time.perf_counter()
    for _ in range(steps):
        est.estimate_device(*args)
        stage.append(est.timings())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    import numpy as np
    st = np.mean(np.array(stage), axis=0)
    bins = out.cpu().numpy()
    est.close()
    return (t1 - t0) * 1e3 / steps, st, bins, grid, wl


def cpu_baseline(B, plan, field, target_s):
    """CPU oracle, 1 thread, on the p1 prefix whose contraction count makes
    about target_s of work; extrapolated to the whole field."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    g = plan.geom
    pc = plan.contractions()
    counts = np.bincount(pc[:, 0], minlength=g.Nmax)
    ncontr = len(pc)
    # ~0.45 ms per contraction at NN=384 on one core: pick the p1 prefix
    per = 3.0e-9 * g.NN * g.NN
    want = max(1, int(target_s / per))
    k = int(np.searchsorted(np.cumsum(counts), want)) + 1
    k = max(1, min(k, g.Nmax))
    lo, hi, av = plan.ring_edges()
    rd = O.calcr(g.NN, lo, hi, av)
    params = {kk: getattr(g, kk) for kk in ("NN", "Nmax", "Nr", "n_mtbin", "UL_k1", "LL_k1",
                                            "BW_k1", "LL", "ff", "beta")}
    t0 = time.perf_counter()
    _, _, _, n, _ = O.estimate(params, lo, hi, av, rd, field, p1_begin=0, p1_end=k)
    dt = time.perf_counter() - t0
    est_s = dt * ncontr / n
    return {"value": 3600.0 / est_s, "unit": "fields/hour", "cores": 1, "kind": "port",
            "sample": f"CPU oracle (FP64, own FFT, 1 thread) over p1 in [0,{k}): {n} of {ncontr} "
                      f"contractions in {dt:.1f} s, extrapolated by contraction count to "
                      f"{est_s:.0f} s/field"}


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
    g = geometry(a.NN)
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
    ms, st, bins, grid, wl = time_path(B, torch, plan, a.precision, d_field, a.steps, a.warmup,
                                       a.full_grid)
    res = {"fast": (ms, st, bins)}
    if not a.skip_fp64:
        r64 = time_path(B, torch, plan, "fp64", d_field, a.steps, a.warmup, a.full_grid)
        res["fp64"] = r64[:3]
        wl64 = r64[4]

    # algorithmic contraction flops: Re(P C) = 2 FMAs per point per
    # contraction, summed over the points of each contraction's grid level
    w_contr = 4.0 * wl["points_summed"]
    t_contr = st[2] * 1e-3
    achieved = w_contr / t_contr
    peak = PEAK[a.precision]
    value = 3600e3 / ms
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
                   "precision": a.precision, "grid_levels": grid},
        "stages_ms": {"gather": st[0], "fft_cache": st[1], "contraction": st[2],
                      "binning": st[3]},
        "worklist": wl,
        "roofline": {"bound": "mfma", "kernel": "contraction",
                     "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "work": f"4 * sum_c NG(c)^2 = {w_contr:.4g} flop per launch (grid levels "
                             f"{grid}); bf16x3 peak taken as 2.5 PF/3"},
    }
    if "fp64" in res:
        ms64, st64, bins64 = res["fp64"]
        out["fp64"] = {"fields_per_hour": 3600e3 / ms64, "ms_per_step": ms64,
                       "stages_ms": {"gather": st64[0], "fft_cache": st64[1],
                                     "contraction": st64[2], "binning": st64[3]},
                       "contraction_TFLOPs": w_contr / (st64[2] * 1e-3) / 1e12,
                       "contraction_frac_fp64_peak": w_contr / (st64[2] * 1e-3) / PEAK["fp64"],
                       "worklist": wl64}
        out["speedup_vs_fp64"] = ms64 / ms
        nb = g.nbins * 16
        err = {}
        for name, k in (("bs", 0),):
            a_ = bins[k * nb:(k + 1) * nb].reshape(-1, 16)
            b_ = bins64[k * nb:(k + 1) * nb].reshape(-1, 16)
            err[name] = [float(np.linalg.norm(a_[:, ii] - b_[:, ii]) /
                               max(np.linalg.norm(b_[:, ii]), 1e-300)) for ii in range(16)]
        out["rel_err_vs_fp64_per_multipole"] = err["bs"]
    if a.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(B, plan, field, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
