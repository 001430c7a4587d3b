"""Section 3 as a distance-prefix container on the GPU (przm_gpu_set_offsets_mode,
PRZM_OFFSETS_PREFIX), through the C ABI: every frame equals the oracle's frame with
the same option (orc_encode_block_opts, PackageOffsetsSection of
core/distance/distance.go:825-849), on

  * edge inputs, the mixture and Markov content at several levels;
  * the seeded blocks of tests/golden/offsets_ties.json: one whose container would
    carry a code stream exactly as long as its match count (written flat), and one
    whose Mode-12 coder state meets its renormalisation bound;
  * the batched graph over several chunks, the device-stream path (TokenizeDevice +
    EncodeD2D), the entry point that allocates nothing, and the serving loop;

and the frames decode with the library's own decoder and the reference C99 decoder.
Switching the handle back to flat offsets gives the flat frames again."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec
from offsets_cases import TIES_PATH, copy_block

from przm_amd import datagen as D
from przm_amd import decode as X

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    from przm_amd import gpu as G
    if not G.IsAvailable():
        pytest.fail("gfx950 device not available (the HIP path must run; there is no fallback)")
    c = G.GPUCompressor(0)
    c.SetOffsetsMode(G.OffsetsPrefix)
    yield c
    c.Close()


def _inputs():
    rng = np.random.default_rng(71)
    yield "empty", np.zeros(0, np.uint8)
    for n in (1, 5, 13, 1000, 65535, 65536, 65537):
        yield f"markov{n}", D.markov(n, n)
    yield "zeros", np.zeros(150000, np.uint8)
    yield "ab", np.frombuffer(b"ab" * 70000, np.uint8).copy()
    yield "random", rng.integers(0, 256, 90000, dtype=np.uint8)
    yield "lowalpha", rng.integers(0, 3, 90000, dtype=np.uint8)
    yield "text", np.frombuffer(b"the quick brown fox jumps over the lazy dog; " * 4000, np.uint8).copy()
    yield "mixture", D.mixture(8 << 20, 72)
    yield "markov", D.markov(4 << 20, 73)
    T = json.load(open(TIES_PATH))
    for key in ("hazard_block", "renorm_block"):
        yield key, copy_block(T[key]["distances"], T[key]["seed"])


def _prefix_frames(frames, sizes):
    """How many frames carry a distance-prefix Section 3."""
    f = np.frombuffer(frames, np.uint8)
    pos, k = 0, 0
    for s in sizes:
        fr = f[pos: pos + int(s)]
        pos += int(s)
        if fr.size <= 12 or fr[4] & 0x04:
            continue
        ext = bool(fr[4] & 0x08)
        lc = int.from_bytes(fr[8:11].tobytes() if ext else fr[8:10].tobytes(), "little")
        lo = int.from_bytes(fr[11:14].tobytes() if ext else fr[10:12].tobytes(), "little")
        o3 = (14 if ext else 12) + 8 + lc
        k += int(lo > 0 and fr[o3] in (0x06, 0x0A, 0x0B))
    return k


def test_frames_equal_oracle(g):
    O = oracle()
    used = 0
    for name, a in _inputs():
        frames, sizes, _ = g.CompressFrames(a)
        of, osz = O.encode_buffer(a, 6, True)
        assert np.array_equal(sizes, osz), name
        assert frames == of, name
        assert X.DecompressFrames(frames, sizes) == a.tobytes(), name
        if have_refdec():
            assert refdec().decompress_frames(frames, sizes) == a.tobytes(), name
        used += _prefix_frames(frames, sizes)
    assert used > 50


@pytest.mark.parametrize("level", [1, 3, 9, 11])
def test_levels_equal_oracle(g, level):
    a = D.mixture(2 << 20, 80 + level)
    frames, sizes, _ = g.CompressFrames(a, level)
    assert frames == oracle().encode_buffer(a, level, True)[0]


def test_fixture_blocks(g):
    """The code-length rule keeps the hazard block's Section 3 flat (the frame equals the
    flat-option frame); the renorm block's container carries a Mode-12 code stream."""
    T = json.load(open(TIES_PATH))
    hb = copy_block(T["hazard_block"]["distances"], T["hazard_block"]["seed"])
    fr, _, _ = g.CompressFrames(hb)
    assert fr == oracle().encode_block(hb, 6, False)[0]
    rb = copy_block(T["renorm_block"]["distances"], T["renorm_block"]["seed"])
    fr, sz, _ = g.CompressFrames(rb)
    assert _prefix_frames(fr, sz) == 1
    assert fr == oracle().encode_block(rb, 6, True)[0]


def test_batched_multichunk_equals_oracle(g):
    """Buffers of mixed sizes in launches of 64 blocks on 16 LZ workgroups, graph replayed twice."""
    from przm_amd import _lib as L
    rng = np.random.default_rng(74)
    sizes = [0, 1, 65536, 65537, 300000, 2 << 20, 5 << 20, 12345, 131072, 7 << 20, 999999]
    bufs = [D.mixture(s, 100 + i) if i % 2 else D.markov(s, 200 + i) for i, s in enumerate(sizes)]
    bufs.append(rng.integers(0, 256, 200000, dtype=np.uint8))
    lib = L.lib()
    tuned = (ctypes.c_int32 * 4)()
    assert lib.przm_gpu_get_tuning(g.handle, tuned) == 0
    assert lib.przm_gpu_set_tuning(g.handle, 16, 64) == 0
    try:
        want = [oracle().encode_buffer(b, 6, True) for b in bufs]
        for rep in range(2):
            res, _ = g.EncodeBatch(bufs)
            for i, ((f, s), (of, osz)) in enumerate(zip(res, want)):
                assert np.array_equal(s, osz) and f == of, (rep, i, bufs[i].size)
    finally:
        assert lib.przm_gpu_set_tuning(g.handle, tuned[0], tuned[1]) == 0


def test_device_streams_and_encode_d2d(g):
    a = D.mixture(3 << 20, 75)
    st, _ = g.TokenizeDevice(a)
    frames, sizes, _ = g.EncodeD2D(st, a.size, 0)
    assert frames == oracle().encode_buffer(a, 6, True)[0]


def test_noalloc_with_prefix_offsets(g):
    from przm_amd import _lib as L
    lib = L.lib()
    for n, seed in ((1 << 20, 76), (300000, 77)):
        a = D.mixture(n, seed)
        d_src, d_out = ctypes.c_void_p(), ctypes.c_void_p()
        cap = int(lib.przm_gpu_frames_bound(a.size)) + 64
        assert lib.przm_gpu_malloc(g.handle, a.size + 64, ctypes.byref(d_src)) == 0
        assert lib.przm_gpu_malloc(g.handle, cap, ctypes.byref(d_out)) == 0
        try:
            assert lib.przm_gpu_h2d(g.handle, d_src, ctypes.c_void_p(a.ctypes.data), a.size) == 0
            nb = (a.size + 65535) // 65536
            sizes = np.zeros(nb, np.uint32)
            olen, nf = ctypes.c_uint64(), ctypes.c_uint64()
            assert lib.przm_gpu_encode_device_noalloc(g.handle, d_src, a.size, 6, d_out, cap, ctypes.byref(olen),
                                                      ctypes.c_void_p(sizes.ctypes.data), nb, ctypes.byref(nf)) == 0
            out = np.zeros(olen.value, np.uint8)
            assert lib.przm_gpu_d2h(g.handle, ctypes.c_void_p(out.ctypes.data), d_out, olen.value) == 0
            assert out.tobytes() == oracle().encode_buffer(a, 6, True)[0]
        finally:
            lib.przm_gpu_free(g.handle, d_src)
            lib.przm_gpu_free(g.handle, d_out)


def test_serving_with_prefix_offsets(g):
    from przm_amd import serve as S
    rng = np.random.default_rng(78)
    bufs = [D.mixture(int(s), 300 + i) for i, s in enumerate(np.exp(rng.uniform(np.log(4 << 10), np.log(4 << 20), 24)))]
    srv = S.Server(g, level=6, hbm_budget=8 << 30, max_batch_bytes=16 << 20, streams=2)
    try:
        ids = [srv.submit(b) for b in bufs]
        for b, rid in zip(bufs, ids):
            frames, sizes, _ = srv.wait(rid)
            assert frames == oracle().encode_buffer(b, 6, True)[0]
    finally:
        srv.close()


def test_back_to_flat(g):
    """The option is the handle's: flat again gives the flat frames (a new graph, too)."""
    from przm_amd import gpu as G
    a = D.mixture(1 << 20, 79)
    try:
        g.SetOffsetsMode(G.OffsetsFlat)
        frames, _, _ = g.CompressFrames(a)
        assert frames == oracle().encode_buffer(a, 6, False)[0]
        res, _ = g.EncodeBatch([a])
        assert res[0][0] == frames
    finally:
        g.SetOffsetsMode(G.OffsetsPrefix)
    frames, _, _ = g.CompressFrames(a)
    assert frames == oracle().encode_buffer(a, 6, True)[0]
    with pytest.raises(G.GPUError):
        g.SetOffsetsMode(2)
