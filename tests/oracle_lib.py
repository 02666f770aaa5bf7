"""ctypes bindings to the CPU oracle (oracle/_build/liborc.so).

Test infrastructure: only the tests, the smoke check and bench.py's
CPU-baseline leg load this, and only as the checker.
"""
import ctypes
import os
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liborc.so")

_dp = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [("NN", ctypes.c_int), ("Nmax", ctypes.c_int), ("Nr", ctypes.c_int),
                ("n_mtbin", ctypes.c_int), ("UL_k1", ctypes.c_int), ("LL_k1", ctypes.c_int),
                ("BW_k1", ctypes.c_int), ("LL", ctypes.c_double), ("ff", ctypes.c_double),
                ("beta", ctypes.c_double)]


class OrcFieldParams(ctypes.Structure):
    _fields_ = [("amp", ctypes.c_double), ("index", ctypes.c_double), ("kunit", ctypes.c_double),
                ("kcut", ctypes.c_double), ("fnl", ctypes.c_double), ("beta", ctypes.c_double),
                ("LL", ctypes.c_double), ("stage", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_fft1d.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_ifft2.argtypes = [ctypes.c_int, _dp]
        L.orc_edges_reference.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_edges_uniform.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                        _dp, _dp, _dp]
        L.orc_calcr.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_df.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_sort.argtypes = [_dp]
        L.orc_cal_coeff.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_estimate.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp, _dp, ctypes.c_int, _dp, _dp,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                   ctypes.c_long]
        L.orc_estimate.restype = ctypes.c_long
        L.orc_write_outputs.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp, _dp, ctypes.c_char_p,
                                        ctypes.c_int]
        L.orc_calpow_mom.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_philox4x64.argtypes = [ctypes.POINTER(ctypes.c_ulonglong),
                                     ctypes.POINTER(ctypes.c_ulonglong),
                                     ctypes.POINTER(ctypes.c_ulonglong)]
        L.orc_gen_log.argtypes = [ctypes.c_double]
        L.orc_gen_log.restype = ctypes.c_double
        L.orc_gen_sincos_2pi.argtypes = [ctypes.c_double, _dp, _dp]
        L.orc_gen_normal_pair.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong, _dp]
        L.orc_generate_field.argtypes = [ctypes.c_int, ctypes.c_ulonglong,
                                         ctypes.POINTER(OrcFieldParams), _dp]
        L.orc_input_ps.argtypes = [ctypes.c_double]
        L.orc_input_ps.restype = ctypes.c_double
        L.orc_triangle_params.argtypes = [_dp, _dp]
        L.orc_write_plane_cache.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                            _dp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p]
        L.orc_write_plane_cache.restype = ctypes.c_long
        L.orc_estimate_disk.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp, _dp, ctypes.c_int,
                                        _dp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_long]
        L.orc_estimate_disk.restype = ctypes.c_long
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def fft1d(x, sign=+1):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    assert lib().orc_fft1d(a.size, sign, _p(a.view(np.float64))) == 0
    return a


def ifft2(x):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    assert lib().orc_ifft2(a.shape[0], _p(a.view(np.float64))) == 0
    return a


def edges_reference(Nr, Nmax):
    lo, hi, av = (np.zeros(Nr) for _ in range(3))
    assert lib().orc_edges_reference(Nr, Nmax, _p(lo), _p(hi), _p(av)) == 0
    return lo, hi, av


def edges_uniform(Nr, width, Nmax, last_to_nmax):
    lo, hi, av = (np.zeros(Nr) for _ in range(3))
    assert lib().orc_edges_uniform(Nr, width, Nmax, int(last_to_nmax), _p(lo), _p(hi), _p(av)) == 0
    return lo, hi, av


def calcr(NN, lo, hi, av, cap=200000):
    Nr = lo.size
    rdata = np.zeros(cap * 4)
    n = lib().orc_calcr(NN, Nr, _p(lo), _p(hi), _p(av), _p(rdata), cap)
    assert n >= 0, n
    return rdata[: 4 * n].reshape(n, 4).copy()


def df(field, NN, plane, lo, hi):
    out = np.zeros((NN, NN), dtype=np.complex128)
    f = np.ascontiguousarray(field, dtype=np.float64)
    lib().orc_df(NN, _p(f), plane, float(lo), float(hi), _p(out.view(np.float64)))
    return out


def cal_coeff(px, pz):
    bs, nm = np.zeros(16), np.zeros(16)
    lib().orc_cal_coeff(float(px), float(pz), _p(bs), _p(nm))
    return bs, nm


def sort3(K):
    a = np.ascontiguousarray(K, dtype=np.float64).reshape(3, 2).copy()
    lib().orc_sort(_p(a))
    return a


def estimate(params, lo, hi, av, rdata, field, p1_begin=0, p1_end=None, nthreads=1,
             record=False, cap=None):
    """Runs the oracle estimator loop.  Returns (bs, norm, biana, count, contr)."""
    P = OrcParams(**params)
    nb = int((P.UL_k1 - P.LL_k1) / (1.0 * P.BW_k1)) * P.n_mtbin * P.n_mtbin
    bs, nm, bi = np.zeros(nb * 16), np.zeros(nb * 16), np.zeros(nb * 16)
    if p1_end is None:
        p1_end = P.Nmax
    rd = np.ascontiguousarray(rdata, dtype=np.float64).reshape(-1)
    f = np.ascontiguousarray(field, dtype=np.float64).reshape(-1)
    contr = None
    cptr = None
    if record:
        cap = cap or 2_000_000
        contr = np.zeros((cap, 8))
        cptr = _p(contr)
    n = lib().orc_estimate(ctypes.byref(P), _p(lo), _p(hi), _p(av), rd.size // 4, _p(rd), _p(f),
                           p1_begin, p1_end, nthreads, _p(bs), _p(nm), _p(bi), cptr,
                           cap if record else 0)
    if n < 0:
        raise RuntimeError(f"orc_estimate failed: {n}")
    if record:
        contr = contr[:n].copy()
    return bs, nm, bi, n, contr


def write_outputs(params, bs, nm, bi, outdir, NRi):
    P = OrcParams(**params)
    assert lib().orc_write_outputs(ctypes.byref(P), _p(bs), _p(nm), _p(bi),
                                   outdir.encode(), NRi) == 0


def calpow_mom(field, NN, Nbin=10, LL=0.56):
    """funcs.c:860-1097: per bin {kmean, P0, P2, P4, P6, num}, shape (Nbin, 6)."""
    out = np.zeros((Nbin, 6))
    f = np.ascontiguousarray(field, dtype=np.float64)
    assert lib().orc_calpow_mom(NN, _p(f), Nbin, float(LL), _p(out)) == 0
    return out


def write_plane_cache(field, NN, Nmax, lo, hi, directory, NRi, nthreads=4, want=None):
    """wrapid_fast (funcs.c:537-566): dataPlaneRing_<NN>_<p>_<r>_R<NRi> files
    in `directory` for every (p, r), or those with want[p, r] true."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    w = None
    if want is not None:
        w = np.ascontiguousarray(want, dtype=np.uint8).reshape(-1)
        assert w.size == Nmax * lo.size
    n = lib().orc_write_plane_cache(NN, _p(f), Nmax, lo.size, _p(lo), _p(hi), directory.encode(),
                                    NRi, nthreads, None if w is None else w.ctypes.data)
    assert n >= 0, n
    return n


def estimate_disk(params, lo, hi, av, rdata, directory, NRi, p1_begin=0, p1_end=None,
                  record=False, cap=None):
    """The oracle loop reading its fields from the on-disk cache as
    barbie.c:274-324 does (one thread).  Returns (bs, norm, biana, count, contr)."""
    P = OrcParams(**params)
    nb = int((P.UL_k1 - P.LL_k1) / (1.0 * P.BW_k1)) * P.n_mtbin * P.n_mtbin
    bs, nm, bi = np.zeros(nb * 16), np.zeros(nb * 16), np.zeros(nb * 16)
    if p1_end is None:
        p1_end = P.Nmax
    rd = np.ascontiguousarray(rdata, dtype=np.float64).reshape(-1)
    contr, cptr = None, None
    if record:
        cap = cap or 2_000_000
        contr = np.zeros((cap, 8))
        cptr = _p(contr)
    n = lib().orc_estimate_disk(ctypes.byref(P), _p(lo), _p(hi), _p(av), rd.size // 4, _p(rd),
                                directory.encode(), NRi, p1_begin, p1_end, _p(bs), _p(nm), _p(bi),
                                cptr, cap if record else 0)
    if n < 0:
        raise RuntimeError(f"orc_estimate_disk failed: {n}")
    if record:
        contr = contr[:n].copy()
    return bs, nm, bi, n, contr


def input_ps(kk):
    return lib().orc_input_ps(float(kk))


def triangle_params(K):
    """barbie.c:293-297 on a sorted {(|k|, k_z)} x 3: (K1, T, MU, pz, px)."""
    a = np.ascontiguousarray(K, dtype=np.float64).reshape(3, 2)
    out = np.zeros(5)
    lib().orc_triangle_params(_p(a), _p(out))
    return out


def philox4x64(ctr, key):
    c = (ctypes.c_ulonglong * 4)(*[int(x) for x in ctr])
    k = (ctypes.c_ulonglong * 2)(*[int(x) for x in key])
    o = (ctypes.c_ulonglong * 4)()
    lib().orc_philox4x64(c, k, o)
    return [int(x) for x in o]


def gen_log(x):
    return lib().orc_gen_log(float(x))


def gen_sincos_2pi(u):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().orc_gen_sincos_2pi(float(u), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def gen_normal_pair(c, seed):
    g = np.zeros(2)
    lib().orc_gen_normal_pair(int(c), int(seed), _p(g))
    return g


def generate_field(NN, seed, amp=1e3, index=-1.5, kunit=1.0, kcut=None, fnl=0.3, beta=1.0,
                   LL=1.0, stage=0):
    """The generator's CPU restatement: [NN][NN][NN+2] doubles (r2c layout)."""
    P = OrcFieldParams(amp, index, kunit, NN / 3.0 if kcut is None else kcut, fnl, beta, LL,
                       stage)
    out = np.zeros((NN, NN, NN + 2))
    assert lib().orc_generate_field(NN, int(seed), ctypes.byref(P), _p(out)) == 0
    return out
