"""bb_power_multipoles_noalloc through the C ABI: the variant whose partial
sums live in the library's static device array gives bb_power_multipoles'
rows bit for bit, on two calls in a row (the second on a smaller field, so
that stale partials of the first would show), on the caller's stream, and
refuses sizes beyond its array."""
import ctypes

import numpy as np
import pytest

import barbie_amd as B

pytestmark = pytest.mark.gpu


def noalloc(d_field, NN, Nbin=10, LL=0.56, stream=None):
    out = np.zeros((Nbin, 6))
    rc = B.lib().bb_power_multipoles_noalloc(ctypes.c_void_p(d_field), NN, Nbin, LL,
                                              ctypes.c_void_p(stream),
                                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


def test_noalloc_power_matches_allocating_entry_point():
    torch = pytest.importorskip("torch")
    cases = [(192, 10, 0.56, 1), (64, 7, 1.0, 2)]   # second call: smaller field, fewer bins
    fields = {NN: torch.from_numpy(B.make_field(NN, NRi=k)).to("cuda") for NN, _, _, k in cases}
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    for NN, nbin, LL, _ in cases:
        ref = B.power_multipoles(fields[NN].data_ptr(), NN, Nbin=nbin, LL=LL)
        rc, got = noalloc(fields[NN].data_ptr(), NN, nbin, LL, stream=s.cuda_stream)
        assert rc == 0, B.lib().bb_last_error()
        assert np.array_equal(got, ref), (NN, nbin)
        assert got[:, 5].sum() > 0
    # the null stream, again after the smaller call
    NN, nbin, LL, _ = cases[0]
    rc, got = noalloc(fields[NN].data_ptr(), NN, nbin, LL)
    assert rc == 0 and np.array_equal(got, B.power_multipoles(fields[NN].data_ptr(), NN,
                                                               Nbin=nbin, LL=LL))
    # sizes beyond the array, and invalid arguments, fail with BB_E_INVALID
    for args in ((NN, 11, LL), (NN, 0, LL), (4098, 10, LL), (NN + 1, 10, LL), (NN, 10, 0.0)):
        rc, _ = noalloc(fields[NN].data_ptr(), *args)
        assert rc == -1, args
    rc, _ = noalloc(0, NN, 10, LL)
    assert rc == -1
