"""pf_nn_forward_noalloc (the inference entry point that allocates nothing)
against pf_nn_forward on two calls in a row, the second on a smaller grid.
Device arrays come from the HIP runtime the library itself is linked to."""
import ctypes as C

import numpy as np
import pytest

from oracle import nn_ref as nr


def _hip():
    return C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)


def _params(seed):
    p = nr.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    p += rng.normal(0.0, 0.02, p.shape) * (p != 0)
    return p.astype(np.float32)


@pytest.mark.gpu
def test_nn_forward_noalloc_matches_the_allocating_entry(pf):
    hip = _hip()
    net = pf.Net((8, 8, 8), max_batch=1)
    net.set_params(_params(3))
    for shape in ((32, 32, 32), (12, 10, 20)):
        S = int(np.prod(shape))
        rng = np.random.default_rng(S)
        x = rng.normal(0.0, 1.0, (1, S, nr.IN_CH)).astype(np.float32)
        ref = net.forward(x, shape)[0]
        d_in, d_out = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(d_in), C.c_size_t(x.nbytes)) == 0
        assert hip.hipMalloc(C.byref(d_out), C.c_size_t(3 * S * 4)) == 0
        try:
            assert hip.hipMemcpy(d_in, x.ctypes.data_as(C.c_void_p), C.c_size_t(x.nbytes), 1) == 0   # H2D
            rc = pf.lib().pf_nn_forward_noalloc(net._h, *shape, d_in, d_out)
            assert rc == 0, pf.PfError(rc, "")
            got = np.empty((3, S), dtype=np.float32)
            assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), d_out, C.c_size_t(got.nbytes), 2) == 0   # D2H
        finally:
            hip.hipFree(d_in)
            hip.hipFree(d_out)
        print(f"noalloc {shape}: max|diff|={np.abs(got - ref).max():.3e}", flush=True)
        assert np.array_equal(got, ref)
    # beyond the static scratch: refused before any launch
    rc = pf.lib().pf_nn_forward_noalloc(net._h, 64, 64, 65, C.c_void_p(1), C.c_void_p(1))
    assert rc == -1
    net.close()
