"""The entry point that allocates nothing (przm_gpu_encode_device_noalloc),
through the C ABI: its frames equal przm_gpu_encode_device's on the same
device input, on two calls in a row (the second smaller, so stale contents of
the shared static scratch would show), and inputs above its maximum are
refused with PRZM_GPU_ERR_CAPACITY."""
import ctypes

import numpy as np
import pytest

from oracle_lib import oracle

from przm_amd import datagen as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    from przm_amd import gpu as G
    if not G.IsAvailable():
        pytest.fail("gfx950 device not available (the HIP path must run; there is no fallback)")
    c = G.GPUCompressor(0)
    yield c
    c.Close()


def _encode(lib, g, fn, d_src, n, d_out, cap):
    olen, nf = ctypes.c_uint64(), ctypes.c_uint64()
    nb = max((n + 65535) // 65536, 1)
    sizes = np.zeros(nb, np.uint32)
    if fn == "noalloc":
        rc = lib.przm_gpu_encode_device_noalloc(g.handle, d_src, n, 6, d_out, cap, ctypes.byref(olen),
                                                ctypes.c_void_p(sizes.ctypes.data), nb, ctypes.byref(nf))
    else:
        rc = lib.przm_gpu_encode_device(g.handle, d_src, n, 6, d_out, cap, ctypes.byref(olen),
                                        ctypes.c_void_p(sizes.ctypes.data), nb, ctypes.byref(nf), None)
    assert rc == 0, (fn, rc)
    out = np.zeros(max(olen.value, 1), np.uint8)
    lib.przm_gpu_d2h(g.handle, ctypes.c_void_p(out.ctypes.data), d_out, olen.value)
    assert nf.value == nb
    return out[: olen.value].tobytes(), sizes


def test_noalloc_equals_encode_device(g):
    from przm_amd import _lib as L
    lib = L.lib()
    cap = 1 << 22
    d_src, d_a, d_b = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    for p in (d_src, d_a, d_b):
        assert lib.przm_gpu_malloc(g.handle, cap, ctypes.byref(p)) == 0
    try:
        # the full 1 MiB first, then a smaller buffer of other content, then an empty one
        cases = [D.mixture(1 << 20, 51), D.markov(300000, 52), np.zeros(0, np.uint8)]
        for a in cases:
            if a.size:
                lib.przm_gpu_h2d(g.handle, d_src, ctypes.c_void_p(a.ctypes.data), a.size)
            f0, s0 = _encode(lib, g, "noalloc", d_src, a.size, d_a, cap)
            f1, s1 = _encode(lib, g, "device", d_src, a.size, d_b, cap)
            assert np.array_equal(s0, s1) and f0 == f1, a.size
            of, osz = oracle().encode_buffer(a)
            assert f0 == of and np.array_equal(s0, osz), a.size
        # above the maximum: refused before any device work
        olen, nf = ctypes.c_uint64(), ctypes.c_uint64()
        sizes = np.zeros(32, np.uint32)
        rc = lib.przm_gpu_encode_device_noalloc(g.handle, d_src, (1 << 20) + 1, 6, d_a, cap, ctypes.byref(olen),
                                                ctypes.c_void_p(sizes.ctypes.data), 32, ctypes.byref(nf))
        assert rc == L.ERR_CAPACITY
    finally:
        for p in (d_src, d_a, d_b):
            lib.przm_gpu_free(g.handle, p)
