"""ctypes bindings of libbarbie_amd.so (the C ABI declared in include/barbie.h).

The product path is the HIP library; this module only marshals arguments.
Loading fails loudly when the library has not been built -- there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field as dc_field
from typing import Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libbarbie_amd.so")
BIN_PATH = os.path.join(PKG_DIR, "bin", "barbie")
CALPOW_PATH = os.path.join(PKG_DIR, "bin", "calpow_mom")

NPOLES = 16
BB_FP64, BB_FP32, BB_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": BB_FP64, "fp32": BB_FP32, "bf16x3": BB_BF16X3}
RINGS_REFERENCE, RINGS_UNIFORM, RINGS_EXPLICIT = 0, 1, 2

# C-ABI symbols declared in include/barbie.h (checked by the CPU tests).
EXPORTED = [
    "bb_last_error", "bb_version", "bb_geometry_reference", "bb_plan_create",
    "bb_plan_destroy", "bb_plan_stats_get", "bb_plan_ring_edges", "bb_plan_triples",
    "bb_plan_contractions", "bb_plan_grid_levels", "bb_plan_worklist", "bb_tune",
    "bb_estimator_create", "bb_estimator_create_opts",
    "bb_estimator_destroy", "bb_estimator_grid", "bb_estimator_levels",
    "bb_estimator_footprint", "bb_estimator_footprint_opts", "bb_estimate", "bb_estimate_device",
    "bb_estimate_batch", "bb_estimate_batch_device", "bb_estimate_batch_async", "bb_estimate_mixed",
    "bb_estimator_query", "bb_estimator_wait", "bb_estimator_stream", "bb_estimator_set_stream",
    "bb_estimator_contraction_sums", "bb_estimator_triangle_counts", "bb_estimator_field",
    "bb_estimator_timings", "bb_estimator_worklist", "bb_write_outputs", "bb_write_outputs_checked", "bb_export_cache", "bb_import_rings", "bb_check_plane_cache",
    "bb_power_multipoles", "bb_power_multipoles_noalloc", "bb_generate_field",
    "bb_server_create", "bb_server_destroy", "bb_submit", "bb_poll", "bb_wait", "bb_server_pause",
    "bb_server_prepare",
    "bb_server_stats",
]
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

# bb_estimator_worklist / bb_plan_worklist entries, in order.
WORKLIST_KEYS = ("items", "units", "tiles", "max_rings", "staged_rings", "staged_bytes", "grid",
                 "chunk", "waves", "wave_balance_x1000", "levels", "partial_tiles",
                 "points_summed", "slots", "total_cost", "cache_points", "band_rows",
                 "pair_points", "pair_points_used", "window_points")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class BBGeometry(ctypes.Structure):
    _fields_ = [("NN", ctypes.c_int), ("Nmax", ctypes.c_int), ("Nr", ctypes.c_int),
                ("ring_rule", ctypes.c_int), ("ring_width", ctypes.c_double),
                ("last_ring_to_nmax", ctypes.c_int), ("ring_low", _dp), ("ring_high", _dp),
                ("n_mtbin", ctypes.c_int), ("UL_k1", ctypes.c_int), ("LL_k1", ctypes.c_int),
                ("BW_k1", ctypes.c_int), ("LL", ctypes.c_double), ("ff", ctypes.c_double),
                ("beta", ctypes.c_double)]


class BBPlanStats(ctypes.Structure):
    _fields_ = [("NN", ctypes.c_int), ("Nmax", ctypes.c_int), ("Nr", ctypes.c_int),
                ("n_mtbin", ctypes.c_int), ("n_k1bin", ctypes.c_int), ("rcount", ctypes.c_int),
                ("ncontr", ctypes.c_longlong), ("npairs", ctypes.c_longlong),
                ("ngroups", ctypes.c_int), ("nfields_used", ctypes.c_int),
                ("nbins", ctypes.c_int)]


class BBOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("full_grid", ctypes.c_int),
                ("x_splits", ctypes.c_int), ("grid_levels", ctypes.c_int),
                ("max_batch", ctypes.c_int), ("waves", ctypes.c_int),
                ("mix_fp32", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("occupancy", ctypes.c_int)]


class BBServerOptions(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_inflight", ctypes.c_int),
                ("max_batch", ctypes.c_int), ("precision", ctypes.c_int), ("tune", ctypes.c_int),
                ("reserved_cus", ctypes.c_int)]


class BBFieldParams(ctypes.Structure):
    _fields_ = [("amp", ctypes.c_double), ("index", ctypes.c_double), ("kunit", ctypes.c_double),
                ("kcut", ctypes.c_double), ("fnl", ctypes.c_double), ("beta", ctypes.c_double),
                ("LL", ctypes.c_double), ("stage", ctypes.c_int)]


class BarbieError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """Loads libbarbie_amd.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C barbie_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        L.bb_last_error.restype = ctypes.c_char_p
        L.bb_version.restype = ctypes.c_char_p
        L.bb_geometry_reference.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(BBGeometry)]
        L.bb_plan_create.argtypes = [ctypes.POINTER(BBGeometry), ctypes.POINTER(ctypes.c_void_p)]
        L.bb_plan_destroy.argtypes = [ctypes.c_void_p]
        L.bb_plan_destroy.restype = None
        L.bb_plan_stats_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(BBPlanStats)]
        L.bb_plan_ring_edges.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.bb_plan_triples.argtypes = [ctypes.c_void_p, _ip]
        L.bb_plan_contractions.argtypes = [ctypes.c_void_p, _ip]
        L.bb_estimator_create.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_void_p)]
        L.bb_estimator_create_opts.argtypes = [ctypes.c_void_p, ctypes.POINTER(BBOptions),
                                               ctypes.POINTER(ctypes.c_void_p)]
        L.bb_estimator_grid.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.bb_estimator_destroy.argtypes = [ctypes.c_void_p]
        L.bb_estimator_destroy.restype = None
        L.bb_estimator_footprint.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_size_t)]
        L.bb_estimate.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.bb_estimate_device.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
        L.bb_estimate_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.bb_estimate_batch_device.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 4
        L.bb_estimate_batch_async.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 4
        L.bb_estimator_query.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.bb_estimator_wait.argtypes = [ctypes.c_void_p]
        L.bb_estimator_stream.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        L.bb_estimator_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.bb_server_create.argtypes = [ctypes.POINTER(BBServerOptions),
                                       ctypes.POINTER(ctypes.c_void_p)]
        L.bb_server_destroy.argtypes = [ctypes.c_void_p]
        L.bb_server_destroy.restype = None
        L.bb_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(BBGeometry), ctypes.c_void_p,
                                ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]
        L.bb_poll.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_int)]
        L.bb_wait.argtypes = [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp, _dp, _dp]
        L.bb_server_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        L.bb_server_pause.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.bb_server_prepare.argtypes = [ctypes.c_void_p, ctypes.POINTER(BBGeometry), ctypes.c_int]
        L.bb_estimator_contraction_sums.argtypes = [ctypes.c_void_p, _dp]
        L.bb_estimator_triangle_counts.argtypes = [ctypes.c_void_p, _dp]
        L.bb_estimator_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _dp]
        L.bb_estimator_levels.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _ip]
        L.bb_estimator_footprint_opts.argtypes = [ctypes.c_void_p, ctypes.POINTER(BBOptions),
                                                  ctypes.POINTER(ctypes.c_size_t)]
        L.bb_tune.argtypes = [ctypes.c_void_p, ctypes.POINTER(BBOptions), ctypes.c_void_p,
                              ctypes.c_int, ctypes.POINTER(BBOptions), _dp]
        L.bb_plan_worklist.argtypes = [ctypes.c_void_p, ctypes.POINTER(BBOptions),
                                       ctypes.POINTER(ctypes.c_longlong)]
        L.bb_plan_grid_levels.argtypes = [ctypes.c_void_p, ctypes.POINTER(BBOptions),
                                          ctypes.POINTER(ctypes.c_int), _ip, _ip]
        L.bb_estimator_timings.argtypes = [ctypes.c_void_p, _dp]
        L.bb_estimator_worklist.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        L.bb_export_cache.argtypes = [ctypes.c_void_p, _dp, ctypes.c_char_p, ctypes.c_int]
        _vpp = ctypes.POINTER(ctypes.c_void_p)
        L.bb_estimate_mixed.argtypes = [ctypes.c_int, _vpp, ctypes.POINTER(ctypes.c_int), _vpp,
                                        _vpp, _vpp, _vpp]
        L.bb_import_rings.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]
        L.bb_check_plane_cache.argtypes = [ctypes.c_void_p, _dp, ctypes.c_char_p, ctypes.c_int,
                                           ctypes.c_int, _dp]
        L.bb_power_multipoles.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, _dp]
        L.bb_power_multipoles_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_double, ctypes.c_void_p, _dp]
        L.bb_generate_field.argtypes = [ctypes.c_int, ctypes.c_ulonglong,
                                        ctypes.POINTER(BBFieldParams), ctypes.c_void_p,
                                        ctypes.c_int]
        L.bb_write_outputs.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_char_p,
                                       ctypes.c_int]
        L.bb_write_outputs_checked.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_char_p,
                                               ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise BarbieError(rc, lib().bb_last_error().decode())


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


@dataclass
class Geometry:
    """Geometry and binning of one estimate (bb_geometry)."""
    NN: int
    UL_k1: int
    LL_k1: int = 1
    BW_k1: int = 1
    Nmax: int = 100
    Nr: int = 40
    ring_rule: int = RINGS_REFERENCE
    ring_width: float = 1.0
    last_ring_to_nmax: int = 1
    ring_low: Optional[np.ndarray] = None
    ring_high: Optional[np.ndarray] = None
    n_mtbin: int = 5
    LL: float = 1.0
    ff: float = 0.1
    beta: float = 1.0
    _keep: list = dc_field(default_factory=list, repr=False)

    def to_c(self) -> BBGeometry:
        g = BBGeometry()
        g.NN, g.Nmax, g.Nr = self.NN, self.Nmax, self.Nr
        g.ring_rule, g.ring_width = self.ring_rule, self.ring_width
        g.last_ring_to_nmax = self.last_ring_to_nmax
        if self.ring_low is not None:
            lo = np.ascontiguousarray(self.ring_low, dtype=np.float64)
            hi = np.ascontiguousarray(self.ring_high, dtype=np.float64)
            self._keep[:] = [lo, hi]
            g.ring_low, g.ring_high = _p(lo), _p(hi)
        g.n_mtbin = self.n_mtbin
        g.UL_k1, g.LL_k1, g.BW_k1 = self.UL_k1, self.LL_k1, self.BW_k1
        g.LL, g.ff, g.beta = self.LL, self.ff, self.beta
        return g

    @property
    def n_k1bin(self) -> int:
        return int((self.UL_k1 - self.LL_k1) / (1.0 * self.BW_k1))

    @property
    def nbins(self) -> int:
        return self.n_k1bin * self.n_mtbin * self.n_mtbin


class Plan:
    """Host-side geometry plan (bb_plan)."""

    def __init__(self, geom: Geometry):
        self.geom = geom
        self._g = geom.to_c()
        h = ctypes.c_void_p()
        _check(lib().bb_plan_create(ctypes.byref(self._g), ctypes.byref(h)))
        self.handle = h
        s = BBPlanStats()
        _check(lib().bb_plan_stats_get(h, ctypes.byref(s)))
        self.stats = {k: getattr(s, k) for k, _ in BBPlanStats._fields_}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.bb_plan_destroy(h)
            self.handle = None

    def ring_edges(self):
        Nr = self.stats["Nr"]
        lo, hi, av = np.zeros(Nr), np.zeros(Nr), np.zeros(Nr)
        _check(lib().bb_plan_ring_edges(self.handle, _p(lo), _p(hi), _p(av)))
        return lo, hi, av

    def triples(self) -> np.ndarray:
        out = np.zeros((self.stats["rcount"], 3), dtype=np.int32)
        _check(lib().bb_plan_triples(self.handle, out.ctypes.data_as(_ip)))
        return out

    def contractions(self) -> np.ndarray:
        """(ncontr, 8) int32: p1, r1, p2, r2, p3, r3, rp, flat_bin."""
        out = np.zeros((self.stats["ncontr"], 8), dtype=np.int32)
        _check(lib().bb_plan_contractions(self.handle, out.ctypes.data_as(_ip)))
        return out

    def footprint(self, precision: str, full_grid: bool = False, x_splits: int = 0,
                  grid_levels: int = 0, max_batch: int = 0) -> int:
        """Device bytes an Estimator with these options holds."""
        b = ctypes.c_size_t()
        o = BBOptions(PRECISIONS[precision], int(full_grid), int(x_splits), int(grid_levels),
                      int(max_batch), 0)
        _check(lib().bb_estimator_footprint_opts(self.handle, ctypes.byref(o), ctypes.byref(b)))
        return b.value

    def grid_levels(self, precision: str = "fp64", full_grid: bool = False,
                    grid_levels: int = 0):
        """(grids, level of each contraction) an Estimator with these options
        uses; host only."""
        n = ctypes.c_int()
        grids = (ctypes.c_int * 16)()
        cl = np.zeros(self.stats["ncontr"], dtype=np.int32)
        o = BBOptions(PRECISIONS[precision], int(full_grid), 0, int(grid_levels), 0, 0)
        _check(lib().bb_plan_grid_levels(self.handle, ctypes.byref(o), ctypes.byref(n), grids,
                                         cl.ctypes.data_as(_ip)))
        return [grids[i] for i in range(n.value)], cl

    def worklist(self, precision: str = "fp64", full_grid: bool = False, x_splits: int = 0,
                 grid_levels: int = 0, waves: int = 0, occupancy: int = 0) -> dict:
        """Work-list figures of an Estimator with these options (host only)."""
        o = (ctypes.c_longlong * len(WORKLIST_KEYS))()
        opt = BBOptions(PRECISIONS[precision], int(full_grid), int(x_splits), int(grid_levels), 0,
                        int(waves), 0, 0, int(occupancy))
        _check(lib().bb_plan_worklist(self.handle, ctypes.byref(opt), o))
        return dict(zip(WORKLIST_KEYS, list(o)))

    def tune(self, precision: str = "bf16x3", full_grid: bool = False, d_field: int = 0,
             repeats: int = 2, max_levels: int = 0):
        """bb_tune: times candidate layouts on the current device; returns
        (options dict for Estimator(**options), ms per estimate)."""
        base = BBOptions(PRECISIONS[precision], int(full_grid), 0, int(max_levels), 0, 0)
        best = BBOptions()
        ms = ctypes.c_double()
        _check(lib().bb_tune(self.handle, ctypes.byref(base), ctypes.c_void_p(d_field or None),
                             int(repeats), ctypes.byref(best), ctypes.byref(ms)))
        return ({"grid_levels": best.grid_levels, "waves": best.waves, "mix_fp32": best.mix_fp32,
                 "occupancy": best.occupancy}, ms.value)

    def write_outputs(self, bs, norm, biana, outdir: str, NRi: int, checked: bool = False):
        """The 16 multipole files (barbie.c:383-397).  checked: multipoles
        whose own denominator is not positive and finite are written as
        "nan"; returns their count (0 when not checked)."""
        bs, norm, biana = (np.ascontiguousarray(a, dtype=np.float64) for a in (bs, norm, biana))
        if not checked:
            _check(lib().bb_write_outputs(self.handle, _p(bs), _p(norm), _p(biana),
                                          outdir.encode(), NRi))
            return 0
        n = ctypes.c_longlong()
        _check(lib().bb_write_outputs_checked(self.handle, _p(bs), _p(norm), _p(biana),
                                              outdir.encode(), NRi, ctypes.byref(n)))
        return n.value


class Estimator:
    """Device estimator (bb_estimator) for one plan and precision."""

    def __init__(self, plan: Plan, precision: str = "fp64", full_grid: bool = False,
                 x_splits: int = 0, grid_levels: int = 0, max_batch: int = 0, waves: int = 0,
                 use_graph: bool = False, mix_fp32: int = 0, occupancy: int = 0):
        self.plan = plan
        self.precision = precision
        self.max_batch = max(1, int(max_batch))
        h = ctypes.c_void_p()
        o = BBOptions(PRECISIONS[precision], int(full_grid), int(x_splits), int(grid_levels),
                      int(max_batch), int(waves), int(mix_fp32), int(use_graph), int(occupancy))
        _check(lib().bb_estimator_create_opts(plan.handle, ctypes.byref(o), ctypes.byref(h)))
        self.handle = h
        ng = ctypes.c_int()
        _check(lib().bb_estimator_grid(h, ctypes.byref(ng)))
        self.grid = ng.value
        n = ctypes.c_int()
        grids = (ctypes.c_int * 16)()
        _check(lib().bb_estimator_levels(h, ctypes.byref(n), grids))
        self.levels = [grids[i] for i in range(n.value)]

    def close(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.bb_estimator_destroy(h)
        self.handle = None

    def __del__(self):
        self.close()

    def estimate(self, field_r2c: np.ndarray):
        """field_r2c: float64 (NN, NN, NN+2).  Returns (bs, norm, biana),
        each (n_k1bin, n_mtbin, n_mtbin, 16) float64, un-normalised."""
        g = self.plan.geom
        f = np.ascontiguousarray(field_r2c, dtype=np.float64)
        assert f.size == g.NN * g.NN * (g.NN + 2), "field must be NN x NN x (NN+2) doubles"
        shape = (g.n_k1bin, g.n_mtbin, g.n_mtbin, NPOLES)
        bs, nm, bi = (np.zeros(shape) for _ in range(3))
        _check(lib().bb_estimate(self.handle, _p(f), _p(bs), _p(nm), _p(bi)))
        return bs, nm, bi

    def estimate_device(self, d_field: int, d_bs: int, d_norm: int, d_biana: int):
        """All arguments are device addresses (e.g. torch tensor data_ptr())."""
        _check(lib().bb_estimate_device(self.handle, ctypes.c_void_p(d_field),
                                        ctypes.c_void_p(d_bs), ctypes.c_void_p(d_norm),
                                        ctypes.c_void_p(d_biana)))

    def estimate_batch(self, fields_r2c: np.ndarray):
        """fields_r2c: float64 (nf, NN, NN, NN+2), nf <= max_batch.  Returns
        (bs, norm, biana), each (nf, n_k1bin, n_mtbin, n_mtbin, 16)."""
        g = self.plan.geom
        f = np.ascontiguousarray(fields_r2c, dtype=np.float64)
        nf = f.shape[0]
        assert f.size == nf * g.NN * g.NN * (g.NN + 2), "fields must be nf x NN x NN x (NN+2)"
        shape = (nf, g.n_k1bin, g.n_mtbin, g.n_mtbin, NPOLES)
        bs, nm, bi = (np.zeros(shape) for _ in range(3))
        _check(lib().bb_estimate_batch(self.handle, nf, _p(f), _p(bs), _p(nm), _p(bi)))
        return bs, nm, bi

    def estimate_batch_device(self, nf: int, d_fields: int, d_bs: int, d_norm: int,
                              d_biana: int, wait: bool = True):
        """Device addresses; outputs back to back per field.  wait=False
        enqueues only (poll with done(), block with wait())."""
        fn = lib().bb_estimate_batch_device if wait else lib().bb_estimate_batch_async
        _check(fn(self.handle, int(nf), ctypes.c_void_p(d_fields), ctypes.c_void_p(d_bs),
                  ctypes.c_void_p(d_norm), ctypes.c_void_p(d_biana)))

    def done(self) -> bool:
        d = ctypes.c_int()
        _check(lib().bb_estimator_query(self.handle, ctypes.byref(d)))
        return bool(d.value)

    def wait(self):
        _check(lib().bb_estimator_wait(self.handle))

    def export_cache(self, field: np.ndarray, directory: str, NRi: int):
        """bb_export_cache: the reference's cache files of `field` (host
        [NN][NN][NN+2] doubles) in `directory`."""
        f = np.ascontiguousarray(field, dtype=np.float64)
        _check(lib().bb_export_cache(self.handle, _p(f), directory.encode(), int(NRi)))

    def import_rings(self, directory: str, NRi: int, rel_tol: float = 1e-12,
                     use_num: bool = True) -> int:
        """bb_import_rings: read closed_rings / rcountFor back (checked);
        returns the number of mismatching triangle counts (0 on success)."""
        n = ctypes.c_longlong()
        _check(lib().bb_import_rings(self.handle, directory.encode(), int(NRi), float(rel_tol),
                                     int(use_num), ctypes.byref(n)))
        return n.value

    def check_plane_cache(self, field: np.ndarray, directory: str, NRi: int,
                          stride: int = 1) -> float:
        """bb_check_plane_cache: largest relative difference between the
        dataPlaneRing files and the field's own plane-ring fields."""
        f = np.ascontiguousarray(field, dtype=np.float64)
        r = ctypes.c_double()
        _check(lib().bb_check_plane_cache(self.handle, _p(f), directory.encode(), int(NRi),
                                          int(stride), ctypes.byref(r)))
        return r.value

    def contraction_sums(self) -> np.ndarray:
        S = np.zeros(self.plan.stats["ncontr"])
        _check(lib().bb_estimator_contraction_sums(self.handle, _p(S)))
        return S

    def triangle_counts(self) -> np.ndarray:
        n = np.zeros(self.plan.stats["rcount"])
        _check(lib().bb_estimator_triangle_counts(self.handle, _p(n)))
        return n

    def field(self, level: int, plane: int, ring: int) -> np.ndarray:
        """Plane-ring field D_{plane,ring} on the grid of a level
        (NG x NG, NG = levels[level])."""
        NG = self.levels[level]
        out = np.zeros((NG, NG), dtype=np.complex128)
        _check(lib().bb_estimator_field(self.handle, level, plane, ring,
                                        _p(out.view(np.float64))))
        return out

    def worklist(self) -> dict:
        o = (ctypes.c_longlong * len(WORKLIST_KEYS))()
        _check(lib().bb_estimator_worklist(self.handle, o))
        return dict(zip(WORKLIST_KEYS, list(o)))

    def timings(self) -> np.ndarray:
        """[gather, fft, contraction, binning, total] ms of the last estimate."""
        t = np.zeros(5)
        _check(lib().bb_estimator_timings(self.handle, _p(t)))
        return t


def normalise(geom: Geometry, bs, norm, biana):
    """barbie.c:390-391: B = bs / (vol * norm), Biana / norm (per multipole)."""
    vol = (geom.NN * geom.LL) ** 3
    with np.errstate(divide="ignore", invalid="ignore"):
        return bs / (vol * norm), biana / norm


SERVER_STATS_KEYS = ("submitted", "completed", "failed", "batches", "batched_fields",
                     "lanes_created", "lanes_evicted", "peak_bytes", "queued", "running",
                     "lanes", "bytes", "budget")


def power_multipoles(field, NN: int, Nbin: int = 10, LL: float = 0.56) -> np.ndarray:
    """bb_power_multipoles (calpow_mom on the GPU): per bin {kmean, P0, P2,
    P4, P6, num}, shape (Nbin, 6).  `field` is a host array [NN][NN][NN+2] or
    a device address (int, e.g. a torch tensor's data_ptr())."""
    out = np.zeros((Nbin, 6))
    if isinstance(field, int):
        src, dev = ctypes.c_void_p(field), 1
    else:
        f = np.ascontiguousarray(field, dtype=np.float64)
        src, dev = f.ctypes.data_as(ctypes.c_void_p), 0
    _check(lib().bb_power_multipoles(src, dev, int(NN), int(Nbin), float(LL), _p(out)))
    return out


class Server:
    """Request server (bb_server): submit fields of any geometry, wait for
    their bins.  Fields are numpy arrays (host, copied at submit) or device
    addresses (int, e.g. a torch tensor's data_ptr(), kept valid until the
    request is done)."""

    def __init__(self, precision: str = "bf16x3", hbm_budget: int = 0, max_inflight: int = 0,
                 max_batch: int = 0, tune: bool = False, reserved_cus: int = 0):
        o = BBServerOptions(int(hbm_budget), int(max_inflight), int(max_batch),
                            PRECISIONS[precision], int(tune), int(reserved_cus))
        h = ctypes.c_void_p()
        _check(lib().bb_server_create(ctypes.byref(o), ctypes.byref(h)))
        self.handle = h
        self._geoms = {}

    def close(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.bb_server_destroy(h)
        self.handle = None

    def __del__(self):
        self.close()

    def submit(self, geom: Geometry, field) -> int:
        g = geom.to_c()
        t = ctypes.c_longlong()
        if isinstance(field, int):
            _check(lib().bb_submit(self.handle, ctypes.byref(g), ctypes.c_void_p(field), 1,
                                   ctypes.byref(t)))
        else:
            f = np.ascontiguousarray(field, dtype=np.float64)
            assert f.size == geom.NN * geom.NN * (geom.NN + 2), "field must be NN x NN x (NN+2)"
            _check(lib().bb_submit(self.handle, ctypes.byref(g), _p(f), 0, ctypes.byref(t)))
        self._geoms[t.value] = geom
        return t.value

    def prepare(self, geom: Geometry, nlanes: int = 1):
        """Create lanes of a geometry ahead of its traffic."""
        g = geom.to_c()
        _check(lib().bb_server_prepare(self.handle, ctypes.byref(g), int(nlanes)))

    def pause(self, paused: bool = True):
        """Start no new batches while paused (a burst then batches together)."""
        _check(lib().bb_server_pause(self.handle, int(paused)))

    def poll(self, ticket: int) -> int:
        st = ctypes.c_int()
        _check(lib().bb_poll(self.handle, ticket, ctypes.byref(st)))
        return st.value

    def wait(self, ticket: int):
        """Returns (bs, norm, biana, latency_ms[queued, running, total])."""
        geom = self._geoms.pop(ticket)
        shape = (geom.n_k1bin, geom.n_mtbin, geom.n_mtbin, NPOLES)
        bs, nm, bi = (np.zeros(shape) for _ in range(3))
        lat = np.zeros(3)
        _check(lib().bb_wait(self.handle, ticket, _p(bs), _p(nm), _p(bi), _p(lat)))
        return bs, nm, bi, lat

    def stats(self) -> dict:
        o = (ctypes.c_longlong * len(SERVER_STATS_KEYS))()
        _check(lib().bb_server_stats(self.handle, o))
        return dict(zip(SERVER_STATS_KEYS, list(o)))


def generate_field(NN: int, seed: int, out=None, amp: float = 1e3, index: float = -1.5,
                   kunit: float = 1.0, kcut=None, fnl: float = 0.3, beta: float = 1.0,
                   LL: float = 1.0, stage: int = 0):
    """bb_generate_field: the seeded non-Gaussian redshift-space field on the
    GPU, in the r2c layout [NN][NN][NN+2] doubles.  out: None (returns a
    numpy array) or a device address (int) to write into.  kcut None: NN/3."""
    p = BBFieldParams(float(amp), float(index), float(kunit),
                      float(NN / 3.0 if kcut is None else kcut), float(fnl), float(beta),
                      float(LL), int(stage))
    if out is None:
        host = np.empty((NN, NN, NN + 2))
        _check(lib().bb_generate_field(int(NN), int(seed), ctypes.byref(p),
                                       host.ctypes.data_as(ctypes.c_void_p), 0))
        return host
    _check(lib().bb_generate_field(int(NN), int(seed), ctypes.byref(p), ctypes.c_void_p(out), 1))
    return None


def estimate_mixed(estimators, nfields, d_fields, d_bs, d_norm, d_biana):
    """bb_estimate_mixed: estimators (Estimator objects, BF16x3, one layout
    shape), per estimator the field count and the device addresses of its
    fields and of its three outputs."""
    n = len(estimators)
    arr = lambda xs: (ctypes.c_void_p * n)(*[ctypes.c_void_p(int(x)) for x in xs])
    _check(lib().bb_estimate_mixed(n, arr([e.handle.value for e in estimators]),
                                   (ctypes.c_int * n)(*[int(x) for x in nfields]), arr(d_fields),
                                   arr(d_bs), arr(d_norm), arr(d_biana)))
