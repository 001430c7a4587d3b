"""The PRZS stream layout (przm_gpu_set_container) on the GPU, through the C ABI.

Every buffer's output equals the oracle's frames wrapped as the reference's multi-chunk stream
(tests/oracle_lib.py stream_of: decompress.go:849-866 header, analyzer.go:242-256 chunks): in the
batched graph path, in launches of a few blocks whose streams cross launch boundaries, in the
single-buffer entry points and in the entry point that allocates nothing; each decodes with the
library's decoder.  Switching back restores the bare frames.
"""
import ctypes

import numpy as np
import pytest

from oracle_lib import oracle, encode_many, stream_of

from przm_amd import datagen as D
from przm_amd import decode as X

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    from przm_amd import gpu as G
    if not G.IsAvailable():
        pytest.fail("gfx950 device not available (the HIP path must run; there is no fallback)")
    c = G.GPUCompressor(0)
    c.SetContainer(G.ContainerStream)
    yield c
    c.Close()


def _bufs():
    rng = np.random.default_rng(4242)
    return [np.zeros(0, np.uint8), D.markov(1, 1), D.markov(65536, 2), D.markov(65537, 3),
            rng.integers(0, 256, 1000007, dtype=np.uint8), D.mixture(700000, 5), D.markov(3 * 65536, 6),
            np.zeros(200000, np.uint8), D.mixture(5000, 7)]


def _spans(sizes):
    """the bytes each block of a stream occupies: header on the first, prefix on every one"""
    s = np.asarray(sizes, np.uint32).copy()
    if s.size > 1:
        s += 4
        s[0] += 17
    return s


def _check(res, bufs, what):
    want = encode_many(bufs)
    for i, ((out, sizes), (of, osz), a) in enumerate(zip(res, want, bufs)):
        exp = stream_of(of, osz, a.size)
        assert out == exp, f"{what}: buffer {i} ({a.size} bytes) differs from the oracle's stream"
        assert np.array_equal(sizes, _spans(osz)), f"{what}: buffer {i} block spans"
        assert X.Decompress(out) == a.tobytes(), f"{what}: buffer {i} round trip"


def test_batched_streams_match_oracle(g):
    bufs = _bufs()
    res, _ = g.EncodeBatch(bufs)
    _check(res, bufs, "batched")


def test_streams_across_launches(g):
    """launches of 8 blocks on 4 LZ workgroups: a stream's header block and its last block lie in
    different launches"""
    from przm_amd import _lib as L
    lib = L.lib()
    tuned = (ctypes.c_int32 * 4)()
    assert lib.przm_gpu_get_tuning(g.handle, tuned) == 0
    try:
        assert lib.przm_gpu_set_tuning(g.handle, 4, 8) == 0
        bufs = [D.mixture(40 * 65536 + 99, 31), D.markov(65536 * 5, 32), np.zeros(0, np.uint8),
                D.markov(9 * 65536 - 1, 33)]
        for graph in (1, 0):
            assert lib.przm_gpu_set_options(g.handle, graph) == 0
            res, _ = g.EncodeBatch(bufs)
            _check(res, bufs, f"chunked graph={graph}")
    finally:
        assert lib.przm_gpu_set_options(g.handle, 1) == 0
        assert lib.przm_gpu_set_tuning(g.handle, tuned[0], tuned[1]) == 0


def test_single_buffer_entry_points(g):
    from przm_amd import _lib as L
    lib = L.lib()
    for a in (D.mixture(300000, 41), D.markov(65536, 42)):
        out, sizes, _ = g.CompressFrames(a)
        of, osz = oracle().encode_buffer(a)
        assert out == stream_of(of, osz, a.size) and np.array_equal(sizes, _spans(osz)), a.size
    # the device entry points, the one that allocates nothing included (at most 1 MiB)
    cap = 1 << 22
    d_src, d_out = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.przm_gpu_malloc(g.handle, cap, ctypes.byref(d_src)) == 0
    assert lib.przm_gpu_malloc(g.handle, cap, ctypes.byref(d_out)) == 0
    try:
        for a in (D.mixture(1 << 20, 43), D.markov(100000, 44), D.markov(4000, 45)):
            lib.przm_gpu_h2d(g.handle, d_src, ctypes.c_void_p(a.ctypes.data), a.size)
            of, osz = oracle().encode_buffer(a)
            exp = stream_of(of, osz, a.size)
            assert lib.przm_gpu_stream_bound(a.size) >= len(exp)
            for fn in ("noalloc", "device"):
                olen, nf = ctypes.c_uint64(), ctypes.c_uint64()
                nb = max((a.size + 65535) // 65536, 1)
                sizes = np.zeros(nb, np.uint32)
                if fn == "noalloc":
                    rc = lib.przm_gpu_encode_device_noalloc(g.handle, d_src, a.size, 6, d_out, cap, ctypes.byref(olen),
                                                            ctypes.c_void_p(sizes.ctypes.data), nb, ctypes.byref(nf))
                else:
                    rc = lib.przm_gpu_encode_device(g.handle, d_src, a.size, 6, d_out, cap, ctypes.byref(olen),
                                                    ctypes.c_void_p(sizes.ctypes.data), nb, ctypes.byref(nf), None)
                assert rc == 0, (fn, rc)
                out = np.zeros(olen.value, np.uint8)
                lib.przm_gpu_d2h(g.handle, ctypes.c_void_p(out.ctypes.data), d_out, olen.value)
                assert out.tobytes() == exp, (fn, a.size)
                assert np.array_equal(sizes, _spans(osz)), (fn, a.size)
    finally:
        lib.przm_gpu_free(g.handle, d_src)
        lib.przm_gpu_free(g.handle, d_out)


def test_frames_layout_restored(g):
    from przm_amd import gpu as G
    bufs = [D.mixture(300000, 51), D.markov(70000, 52)]
    try:
        g.SetContainer(G.ContainerFrames)
        res, _ = g.EncodeBatch(bufs)
        for (out, sizes), (of, osz) in zip(res, encode_many(bufs)):
            assert out == of and np.array_equal(sizes, osz)
    finally:
        g.SetContainer(G.ContainerStream)
