"""GPU parity: the HIP path, called through the C ABI, against the oracle.

Bit-exact for every stage (all arithmetic is integer / byte work):
match descriptors, the six token streams, histograms, V-ANS containers and
frames; frames also decode with the reference C99 decoder.  Full-size
cases use size-independent properties (decode round trip, frame count,
batched == plain == alone).
"""
import os

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec

from przm_amd import datagen as D

pytestmark = pytest.mark.gpu
STREAMS = ["cmd", "off", "lit", "op", "meta", "var"]


@pytest.fixture(scope="module")
def g():
    from przm_amd import gpu as G
    if not G.IsAvailable():
        pytest.fail("gfx950 device not available (the HIP path must run; there is no fallback)")
    c = G.GPUCompressor(0)
    yield c
    c.Close()


def edge_inputs():
    rng = np.random.default_rng(17)
    yield "empty", np.zeros(0, np.uint8)
    for n in (1, 3, 4, 5, 11, 12, 13, 64, 1023, 1024, 1025, 4097, 32767, 32768, 32769, 65535, 65536, 65537):
        yield f"markov{n}", D.markov(n, n)
    yield "zeros64k", np.zeros(65536, np.uint8)
    yield "zeros100k", np.zeros(100000, np.uint8)
    yield "ab", np.frombuffer(b"ab" * 40000, np.uint8).copy()
    yield "random64k", rng.integers(0, 256, 65536, dtype=np.uint8)
    yield "lowalpha", rng.integers(0, 3, 70000, dtype=np.uint8)
    yield "runs", np.repeat(rng.integers(0, 256, 3000, dtype=np.uint8), rng.integers(1, 40, 3000))[:90000].copy()
    yield "text", np.frombuffer(b"The quick brown fox jumps over the lazy dog. " * 3000, np.uint8).copy()
    yield "mixture", D.mixture(3 << 20, 5)
    yield "markov", D.markov(2 << 20, 6)


def test_device_info(g):
    assert g.Arch.startswith("gfx950")
    assert g.SMCount > 0 and g.TotalVRAMBytes > (100 << 30)


def test_match_descriptors_match_oracle(g):
    O = oracle()
    for name, a in edge_inputs():
        if a.size == 0:
            continue
        desc, _ = g.MatchDescriptors(a)
        for b in range((a.size + 65535) // 65536):
            blk = a[b * 65536:(b + 1) * 65536]
            want = O.match_block(blk)
            got = desc[b * 65536: b * 65536 + blk.size]
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, f"{name} block {b}: {bad.size} positions differ, first {bad[:5]}"


def test_results_do_not_carry_between_blocks(g):
    """Blocks with many matches alternate with blocks with none, more blocks than workgroups, so
    a workgroup's per-block scratch is reused: each block's results are its own."""
    O = oracle()
    rng = np.random.default_rng(23)
    nb = 640
    a = np.concatenate([D.markov(65536, 100 + b) if b % 2 == 0 else rng.integers(0, 256, 65536, dtype=np.uint8)
                        for b in range(nb)])
    desc, _ = g.MatchDescriptors(a)
    for b in range(nb):
        want = O.match_block(a[b * 65536:(b + 1) * 65536])
        bad = np.nonzero(desc[b * 65536:(b + 1) * 65536] != want)[0]
        assert bad.size == 0, f"block {b}: {bad.size} positions differ, first {bad[:5]}"


def test_token_streams_match_oracle(g):
    O = oracle()
    for name, a in edge_inputs():
        if a.size == 0:
            continue
        s, lens, _ = g.TokenizeStreams(a)
        pos = np.zeros(6, np.int64)
        for b in range((a.size + 65535) // 65536):
            ot = O.tokenize(a[b * 65536:(b + 1) * 65536])
            for i, k in enumerate(STREAMS):
                l = int(lens[b, i])
                got = s[k][pos[i]: pos[i] + l]
                pos[i] += l
                assert np.array_equal(got, ot[k]), f"{name} block {b} stream {k}: {l} vs {ot[k].size} bytes"
            assert int(lens[b, 6]) == ot["ntok"]


def test_frames_match_oracle_and_decode(g):
    O = oracle()
    R = refdec() if have_refdec() else None
    for name, a in edge_inputs():
        frames, sizes, _ = g.CompressFrames(a)
        of, osz = O.encode_buffer(a)
        assert np.array_equal(sizes, osz), f"{name}: frame sizes differ"
        assert frames == of, f"{name}: frame bytes differ"
        if R is not None and a.size:
            assert R.decompress_frames(frames, sizes) == a.tobytes(), name
        from przm_amd import decode as X
        assert X.DecompressFrames(frames, sizes) == a.tobytes(), name


@pytest.mark.parametrize("level", [1, 3, 5, 9, 10, 11])
def test_levels_match_oracle(g, level):
    """Every chain depth class: 2, 4, 8 (the short-chain search), 32, 64 (chains inside the
    staged ring) and 128 (chains past the ring, read from the block)."""
    O = oracle()
    a = D.mixture(1 << 20, level)
    frames, sizes, _ = g.CompressFrames(a, level)
    of, osz = O.encode_buffer(a, level)
    assert frames == of and np.array_equal(sizes, osz)


def test_deep_chains_on_edge_inputs(g):
    """Level 11 (chains of 128, most of them longer than the staged ring) on inputs whose groups
    are longest: all zeros, a period of 2, periodic text, a three-letter alphabet, byte runs.
    The degenerate ones are kept short: the oracle's chain walk costs depth x match length."""
    O = oracle()
    rng = np.random.default_rng(17)
    cases = {
        "zeros": np.zeros(8000, np.uint8),
        "ab": np.frombuffer(b"ab" * 4000, np.uint8).copy(),
        "text": np.frombuffer(b"The quick brown fox jumps over the lazy dog. " * 180, np.uint8).copy(),
        "lowalpha": rng.integers(0, 3, 70000, dtype=np.uint8),
        "runs": np.repeat(rng.integers(0, 256, 3000, dtype=np.uint8), rng.integers(1, 40, 3000))[:90000].copy(),
        "markov": D.markov(65537, 65537),
    }
    for name, a in cases.items():
        frames, sizes, _ = g.CompressFrames(a, 11)
        of, osz = O.encode_buffer(a, 11)
        assert np.array_equal(sizes, osz) and frames == of, name


def test_histogram_and_vans_match(g):
    O = oracle()
    rng = np.random.default_rng(3)
    for n in (1, 5, 1000, 65536, 1 << 20):
        a = rng.geometric(0.1, n).clip(0, 255).astype(np.uint8)
        counts, _ = g.ComputeHistogramHost(a)
        assert np.array_equal(counts, np.bincount(a, minlength=256))
        v, _ = g.VANSEncode4xHost(a)
        assert v == O.vans(a), n
    f, _ = g.PackStreamFlatHost(b"hello")
    assert f == b"\x00\x05\x00\x00\x00hello"


def test_vans_largest_input(g):
    """The stage coder takes up to 16 MiB (MaxBlockSize, api.go:22), where a one-symbol table
    reaches a count of 2^24; one byte more is refused."""
    from przm_amd import _lib as L
    from przm_amd.gpu import GPUError
    a = np.full(16 << 20, 7, np.uint8)
    v, _ = g.VANSEncode4xHost(a)
    assert v == oracle().vans(a)
    with pytest.raises(GPUError) as e:
        g.VANSEncode4xHost(np.zeros((16 << 20) + 1, np.uint8))
    assert e.value.code == L.ERR_ARG


def test_batched_equals_alone(g):
    """Each buffer's frames are the same alone and in a mixed batch."""
    arena, offs, sizes = D.batch(60, 1024, 1 << 20, 9)
    bufs = [arena[int(o): int(o) + int(s)] for o, s in zip(offs, sizes)]
    bufs.insert(7, np.zeros(0, np.uint8))
    res, prof = g.EncodeBatch(bufs)
    O = oracle()
    for i, b in enumerate(bufs):
        alone, asz, _ = g.CompressFrames(b)
        assert res[i][0] == alone and np.array_equal(res[i][1], asz), i
        if i % 10 == 0:
            of, osz = O.encode_buffer(b)
            assert of == alone


def test_full_size_markov_256mib_roundtrip(g):
    """256 MiB, 4096 blocks: the plain-GPU case at its real size; decode round trip."""
    from przm_amd import decode as X
    a = D.markov(256 << 20, 1)
    frames, sizes, prof = g.CompressFrames(a)
    assert sizes.size == 4096 and int(sizes.sum()) == len(frames)
    assert len(frames) < a.size
    assert X.DecompressFrames(frames, sizes, a.size) == a.tobytes()  # every frame, the library's own decoder
    if have_refdec():
        R = refdec()
        f = np.frombuffer(frames, np.uint8)
        offs = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))])
        for b in range(0, 4096, 37):  # every 37th frame through the reference decoder
            d = R.decompress(f[offs[b]: offs[b + 1]])
            assert d == a[b * 65536:(b + 1) * 65536].tobytes()
    # and a slice compared with the oracle byte for byte
    O = oracle()
    of, osz = O.encode_buffer(a[: 64 * 65536])
    assert np.array_equal(osz, sizes[:64])
    assert of == frames[: int(sizes[:64].sum())]


def test_copy_d2d(g):
    """przm_gpu_copy_d2d: aligned (vector) and unaligned (byte) copies of ragged sizes."""
    import ctypes
    from przm_amd import _lib as L
    lib = L.lib()
    rng = np.random.default_rng(8)
    n = (8 << 20) + 77
    src = rng.integers(0, 256, n, dtype=np.uint8)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.przm_gpu_malloc(g.handle, n + 64, ctypes.byref(a)) == 0
    assert lib.przm_gpu_malloc(g.handle, n + 64, ctypes.byref(b)) == 0
    try:
        lib.przm_gpu_h2d(g.handle, a, ctypes.c_void_p(src.ctypes.data), n)
        for off_s, off_d, m in [(0, 0, n), (0, 0, 15), (3, 16, 1 << 20), (16, 5, 12345), (0, 0, 0)]:
            out = np.zeros(m + 1, np.uint8)
            assert lib.przm_gpu_copy_d2d(g.handle, ctypes.c_void_p(b.value + off_d), ctypes.c_void_p(a.value + off_s), m) == 0
            if m:
                lib.przm_gpu_d2h(g.handle, ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(b.value + off_d), m)
                assert np.array_equal(out[:m], src[off_s: off_s + m]), (off_s, off_d, m)
    finally:
        lib.przm_gpu_free(g.handle, a)
        lib.przm_gpu_free(g.handle, b)


def test_mock_error_and_fallback(g):
    from przm_amd import gpu as G
    O = oracle()
    a = D.markov(100000, 2)
    G.SetMockMatchErrorForTest(RuntimeError("injected"))
    try:
        with pytest.raises(RuntimeError):
            g.CompressD2D(a)
        frames, path = G.compress_with_fallback(a, lambda s, lv: O.encode_buffer(s, lv)[0])
        assert path == "cpu"
    finally:
        G.SetMockMatchErrorForTest(None)
    frames, path = G.compress_with_fallback(a, lambda s, lv: O.encode_buffer(s, lv)[0])
    assert path == "gpu" and frames == O.encode_buffer(a)[0]


def test_serving_results_equal_batch(g):
    from przm_amd import serve as S
    arena, offs, sizes = D.batch(40, 4096, 2 << 20, 21)
    bufs = [arena[int(o): int(o) + int(s)] for o, s in zip(offs, sizes)]
    srv = S.Server(g, level=6, hbm_budget=8 << 30, max_batch_bytes=4 << 20, streams=3)
    ids = [srv.submit(b) for b in bufs]
    outs = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    for b, (frames, fs, res) in zip(bufs, outs):
        alone, asz, _ = g.CompressFrames(b)
        assert frames == alone and np.array_equal(fs, asz)
    assert st["done"] == len(bufs) and st["launches"] >= 2
