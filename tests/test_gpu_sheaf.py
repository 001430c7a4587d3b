"""Section 4 with the sheaf trials on the GPU (przm_gpu_set_literals_mode,
PRZM_LITERALS_SHEAF), through the C ABI: every frame equals the oracle's frame with the
same option (orc_encode_block_opts, PackageLiteralsSection of
core/literals/literals.go:667-693), alone and together with the Section-3 prefix option,
on edge inputs, structured text where the sheaf containers win, the mixture and Markov
content, the batched graph over several chunks, the device-stream path, the entry point
that allocates nothing and the serving loop; the frames decode with the library's own
decoder and the reference C99 decoder."""
import ctypes

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec
from sheaf_cases import modal_text, unique_text

from przm_amd import datagen as D
from przm_amd import decode as X

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    from przm_amd import gpu as G
    if not G.IsAvailable():
        pytest.fail("gfx950 device not available (the HIP path must run; there is no fallback)")
    c = G.GPUCompressor(0)
    c.SetLiteralsMode(G.LiteralsSheaf)
    yield c
    c.Close()


def _json(n, seed):
    rng = np.random.default_rng(seed)
    return np.frombuffer(bytes(rng.choice(list(b'{"name": "value_1", "n": 12.5, "ok": true}, \n'), n).tolist()),
                         np.uint8).copy()


def _inputs():
    rng = np.random.default_rng(91)
    yield "empty", np.zeros(0, np.uint8)
    for n in (1, 5, 13, 1000, 65535, 65536, 65537):
        yield f"markov{n}", D.markov(n, n)
    yield "zeros", np.zeros(150000, np.uint8)
    yield "random", rng.integers(0, 256, 90000, dtype=np.uint8)
    yield "text", np.frombuffer(b"the quick brown fox jumps over the lazy dog; " * 4000, np.uint8).copy()
    yield "json", _json(600000, 92)
    for mode in (1, 2, 3):
        yield f"mode{mode}", modal_text(mode, 200000, 40 + mode)
    yield "mixture", D.mixture(8 << 20, 93)
    yield "markov", D.markov(4 << 20, 94)


def _kinds(frames, sizes):
    """Section-4 container kinds (0 raw or none, 1 V-ANS, 2 + mode sheaf) of compressed frames."""
    f = np.frombuffer(frames, np.uint8)
    pos, out = 0, []
    for s in sizes:
        fr = f[pos: pos + int(s)]
        pos += int(s)
        if fr.size <= 12 or fr[4] & 0x04 or not fr[4] & 0x80:
            continue
        ext = bool(fr[4] & 0x08)
        lc = int.from_bytes(fr[8:11].tobytes() if ext else fr[8:10].tobytes(), "little")
        lo = int.from_bytes(fr[11:14].tobytes() if ext else fr[10:12].tobytes(), "little")
        s4 = fr[(14 if ext else 12) + 8 + lc + lo:]
        out.append(2 + (int(s4[3]) >> 4) if s4[3] & 0x04 else 1)
    return out


@pytest.mark.parametrize("prefix", [False, True])
def test_frames_equal_oracle(g, prefix):
    from przm_amd import gpu as G
    g.SetOffsetsMode(G.OffsetsPrefix if prefix else G.OffsetsFlat)
    try:
        O = oracle()
        kinds = set()
        for name, a in _inputs():
            frames, sizes, _ = g.CompressFrames(a)
            of, osz = O.encode_buffer(a, 6, prefix, True)
            assert np.array_equal(sizes, osz), name
            assert frames == of, name
            assert X.DecompressFrames(frames, sizes) == a.tobytes(), name
            if have_refdec():
                assert refdec().decompress_frames(frames, sizes) == a.tobytes(), name
            kinds |= set(_kinds(frames, sizes))
        assert kinds >= {1, 2, 3, 4, 5}, kinds
    finally:
        g.SetOffsetsMode(G.OffsetsFlat)


@pytest.mark.parametrize("level", [1, 11])
def test_levels_equal_oracle(g, level):
    a = np.concatenate([D.mixture(1 << 20, 95 + level), _json(300000, 96 + level)])
    frames, sizes, _ = g.CompressFrames(a, level)
    assert frames == oracle().encode_buffer(a, level, False, True)[0]


def test_literal_round_boundaries(g):
    """Blocks of 64 KiB that repeat their first T bytes of a text with no repeated 4-byte
    string, so each has exactly T literals (checked with the oracle's tokenizer): literal
    counts at the split kernel's 4,096-literal rounds and around them, up to a block of
    literals only, in contexts that make each mode's containers win."""
    O = oracle()
    targets = (1, 15, 16, 17, 21, 22, 23, 4095, 4096, 4097, 8191, 8192, 8193, 20480, 40961, 65535, 65536)
    for mode in range(4):
        src = unique_text(mode, 65536, 800 + mode)
        blocks = [np.resize(src[:t], 65536) for t in targets]
        counts = [O.tokenize(b)["lit"].size for b in blocks]
        assert {4095, 4096, 4097, 65536} <= set(counts), (mode, counts)
        a = np.concatenate(blocks)
        frames, sizes, _ = g.CompressFrames(a)
        of, osz = O.encode_buffer(a, 6, False, True)
        assert np.array_equal(sizes, osz), mode
        assert frames == of, mode
        assert 2 + mode in set(_kinds(frames, sizes)), (mode, _kinds(frames, sizes))


def test_batched_multichunk_equals_oracle(g):
    from przm_amd import _lib as L
    sizes = [0, 1, 65536, 65537, 300000, 2 << 20, 12345, 131072, 999999]
    bufs = [_json(s, 400 + i) if i % 3 == 0 else D.mixture(s, 500 + i) for i, s in enumerate(sizes)]
    lib = L.lib()
    tuned = (ctypes.c_int32 * 4)()
    assert lib.przm_gpu_get_tuning(g.handle, tuned) == 0
    assert lib.przm_gpu_set_tuning(g.handle, 16, 64) == 0
    try:
        want = [oracle().encode_buffer(b, 6, False, True) for b in bufs]
        for rep in range(2):
            res, _ = g.EncodeBatch(bufs)
            for i, ((f, s), (of, osz)) in enumerate(zip(res, want)):
                assert np.array_equal(s, osz) and f == of, (rep, i, bufs[i].size)
    finally:
        assert lib.przm_gpu_set_tuning(g.handle, tuned[0], tuned[1]) == 0


def test_device_streams_and_encode_d2d(g):
    a = np.concatenate([_json(1 << 20, 97), D.mixture(1 << 20, 98)])
    st, _ = g.TokenizeDevice(a)
    frames, sizes, _ = g.EncodeD2D(st, a.size, 0)
    assert frames == oracle().encode_buffer(a, 6, False, True)[0]


def test_noalloc_with_sheaf(g):
    from przm_amd import _lib as L
    lib = L.lib()
    for a in (_json(1 << 20, 99), D.mixture(300000, 100)):
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
            assert out.tobytes() == oracle().encode_buffer(a, 6, False, True)[0]
        finally:
            lib.przm_gpu_free(g.handle, d_src)
            lib.przm_gpu_free(g.handle, d_out)


def test_serving_with_sheaf(g):
    from przm_amd import serve as S
    rng = np.random.default_rng(101)
    bufs = [_json(int(s), 600 + i) if i % 2 else D.mixture(int(s), 700 + i)
            for i, s in enumerate(np.exp(rng.uniform(np.log(4 << 10), np.log(4 << 20), 20)))]
    srv = S.Server(g, level=6, hbm_budget=8 << 30, max_batch_bytes=16 << 20, streams=2)
    try:
        ids = [srv.submit(b) for b in bufs]
        for b, rid in zip(bufs, ids):
            frames, sizes, _ = srv.wait(rid)
            assert frames == oracle().encode_buffer(b, 6, False, True)[0]
    finally:
        srv.close()


def test_back_to_vans(g):
    from przm_amd import gpu as G
    a = _json(500000, 102)
    try:
        g.SetLiteralsMode(G.LiteralsVANS)
        frames, _, _ = g.CompressFrames(a)
        assert frames == oracle().encode_buffer(a, 6, False, False)[0]
    finally:
        g.SetLiteralsMode(G.LiteralsSheaf)
    frames, _, _ = g.CompressFrames(a)
    assert frames == oracle().encode_buffer(a, 6, False, True)[0]
    with pytest.raises(G.GPUError):
        g.SetLiteralsMode(2)
