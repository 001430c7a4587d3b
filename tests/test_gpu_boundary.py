"""The rest of the drop-in boundary, through the C ABI: device-resident token
streams and their frames (TokenizeDevice / TokenizePointersDeviceOnly /
TokenizeDeviceOnly / DownloadDeviceBytes / EncodeD2D of gpu/bridge.go), the
device-buffer stages (ComputeHistogramGPU, PackStreamFlatGPU, VANSEncode4xGPU
with caller tables), the probes and the pinned buffer pools.

Device streams are checked against the host-copy path and the oracle
(device == host, as gpu/tokenizer_test.go:269-364 does for the reference), and
EncodeD2D's frames against przm_gpu_encode and the oracle.
"""
import ctypes

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec

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


@pytest.fixture(scope="module")
def lib():
    from przm_amd import _lib as L
    return L.lib()


def _inputs():
    rng = np.random.default_rng(61)
    yield "mixture", D.mixture(1 << 20, 61)
    yield "markov_ragged", D.markov(200003, 62)
    yield "random_then_text", np.concatenate([rng.integers(0, 256, 70000, dtype=np.uint8),
                                              np.frombuffer(b"the lazy dog " * 6000, np.uint8)])
    yield "tiny", np.frombuffer(b"abcabcabcabc", np.uint8).copy()


def _dmalloc(g, lib, n):
    p = ctypes.c_void_p()
    assert lib.przm_gpu_malloc(g.handle, max(n, 1), ctypes.byref(p)) == 0
    return p.value


def test_tokenize_device_streams_equal_host_and_oracle(g):
    O = oracle()
    for name, a in _inputs():
        st, prof = g.TokenizeDevice(a)
        host, lens, _ = g.TokenizeStreams(a)
        assert st.NumChunks == lens.shape[0] and prof.NumChunks == st.NumChunks
        for key, dp, ln in (("cmd", st.DFinalCmds, st.CmdLen), ("off", st.DFinalOffs, st.OffLen),
                            ("lit", st.DFinalLits, st.LitLen)):
            got = g.DownloadDeviceBytes(dp, ln)
            assert got == host[key].tobytes(), (name, key)
        tot = np.frombuffer(g.DownloadDeviceBytes(st.DTotals, 12), np.uint32)
        assert list(tot) == [st.CmdLen, st.OffLen, st.LitLen]
        bl = np.frombuffer(g.DownloadDeviceBytes(st.DBlockLens, 12 * st.NumChunks), np.uint32).reshape(-1, 3)
        assert np.array_equal(bl, lens[:, [0, 1, 2]]), name
        for b in range(st.NumChunks):  # each block's streams against the oracle
            ot = O.tokenize(a[b * 65536:(b + 1) * 65536])
            assert [int(x) for x in bl[b]] == [ot["cmd"].size, ot["off"].size, ot["lit"].size], (name, b)


def test_encode_d2d_frames_equal_plain_and_oracle(g):
    for name, a in _inputs():
        st, _ = g.TokenizeDevice(a)
        frames, sizes, _ = g.EncodeD2D(st, a.size, 0)
        of, osz = oracle().encode_buffer(a)
        assert frames == of and np.array_equal(sizes, osz), name
        again, sizes2, _ = g.EncodeD2D(st, a.size, 0)  # the streams stay valid for a second encode
        assert again == frames
        plain, psz, _ = g.CompressFrames(a)
        assert plain == frames
        if have_refdec():
            assert refdec().decompress_frames(frames, sizes) == a.tobytes()


def test_encode_d2d_refuses_stale_or_wrong_streams(g):
    from przm_amd.gpu import GPUError
    from przm_amd import _lib as L
    a = D.markov(100000, 63)
    st, _ = g.TokenizeDevice(a)
    with pytest.raises(GPUError) as e:
        g.EncodeD2D(st, a.size, 16)  # mode G sets each frame's window log itself
    assert e.value.code == L.ERR_ARG
    with pytest.raises(GPUError) as e:
        g.EncodeD2D(st, a.size - 1, 0)
    assert e.value.code == L.ERR_ARG
    g.CompressFrames(a[:1000])  # a later call reuses the workspace
    with pytest.raises(GPUError) as e:
        g.EncodeD2D(st, a.size, 0)
    assert e.value.code == L.ERR_ARG
    assert g.EncodeD2D(st, 0, 0)[0] == b""


def test_tokenize_pointers_device_only(g, lib):
    """Input staged in the compressor's input buffer, and input in a caller buffer at an odd
    offset: the same streams and frames as the host path."""
    a = D.mixture(300001, 64)
    d = g.InputBuffer(a.size)
    assert lib.przm_gpu_h2d(g.handle, ctypes.c_void_p(d), ctypes.c_void_p(a.ctypes.data), a.size) == 0
    cmd, off, lit, _ = g.TokenizeDeviceOnly(a.size)
    host, _, _ = g.TokenizeStreams(a)
    assert cmd == host["cmd"].tobytes() and off == host["off"].tobytes() and lit == host["lit"].tobytes()
    buf = _dmalloc(g, lib, a.size + 64)
    try:
        assert lib.przm_gpu_h2d(g.handle, ctypes.c_void_p(buf + 3), ctypes.c_void_p(a.ctypes.data), a.size) == 0
        st, _ = g.TokenizePointersDeviceOnly(a.size, buf + 3)
        frames, sizes, _ = g.EncodeD2D(st, a.size)
        assert frames == oracle().encode_buffer(a)[0]
    finally:
        lib.przm_gpu_free(g.handle, ctypes.c_void_p(buf))


def test_histogram_and_pack_flat_on_device_buffers(g, lib):
    rng = np.random.default_rng(65)
    n_max = (1 << 20) + 40
    src = (rng.geometric(0.05, n_max) % 256).astype(np.uint8)
    din, dout = _dmalloc(g, lib, n_max + 64), _dmalloc(g, lib, n_max + 64)
    try:
        assert lib.przm_gpu_h2d(g.handle, ctypes.c_void_p(din), ctypes.c_void_p(src.ctypes.data), n_max) == 0
        for off, n in [(0, 1), (1, 15), (3, 17), (0, 4096), (7, 1 << 20), (16, (1 << 20) + 3)]:
            counts, ms, _ = g.ComputeHistogramGPU(din + off, n)
            assert np.array_equal(counts, np.bincount(src[off:off + n], minlength=256)), (off, n)
            for doff in (0, 5, 11):
                ln, _, _ = g.PackStreamFlatGPU(din + off, dout + doff, n)
                assert ln == n + 5
                got = np.frombuffer(g.DownloadDeviceBytes(dout + doff, ln), np.uint8)
                assert got[0] == 0 and int.from_bytes(got[1:5].tobytes(), "little") == n
                assert np.array_equal(got[5:], src[off:off + n]), (off, n, doff)
        assert g.PackStreamFlatGPU(din, dout, 0)[0] == 0
        assert g.PackStreamFlatHost(b"")[0] == b"\0" * 5
    finally:
        lib.przm_gpu_free(g.handle, ctypes.c_void_p(din))
        lib.przm_gpu_free(g.handle, ctypes.c_void_p(dout))


def test_vans_encode_gpu_with_caller_tables(g, lib):
    from przm_amd import gpu as G
    from przm_amd.gpu import GPUError
    from przm_amd import _lib as L
    O = oracle()
    rng = np.random.default_rng(66)
    for n, off in ((1, 0), (5, 1), (1000, 3), (65536, 0), (300001, 9)):
        a = rng.geometric(0.1, n).clip(0, 255).astype(np.uint8)
        din, dout = _dmalloc(g, lib, n + 64), _dmalloc(g, lib, 2 * n + 700)
        try:
            lib.przm_gpu_h2d(g.handle, ctypes.c_void_p(din + off), ctypes.c_void_p(a.ctypes.data), n)
            # the adaptive table (NormalizeFreq / BuildCDF / SerializeMatrix, bridge.go:2777-2886)
            freq = G.NormalizeFreq(np.bincount(a, minlength=256).astype(np.uint32))
            ln, _, _ = g.VANSEncode4xGPU(din + off, dout, n, freq, G.BuildCDF(freq), G.SerializeMatrix(freq))
            got = g.DownloadDeviceBytes(dout, ln)
            assert got == O.vans(a, freq) == O.vans(a), n
            # a table of the caller's own: every byte 16
            flat = np.full(256, 16, np.uint16)
            ln, _, _ = g.VANSEncode4xGPU(din + off, dout, n, flat, G.BuildCDF(flat), G.SerializeMatrix(flat))
            got = g.DownloadDeviceBytes(dout, ln)
            assert got == O.vans(a, flat), n
            if have_refdec():
                assert refdec().vans_decode(got, n) == a.tobytes()
            # a byte of the input with frequency 0, and a table that does not sum to 4096
            bad = freq.copy()
            bad[int(a[0])] = 0
            with pytest.raises(GPUError) as e:
                g.VANSEncode4xGPU(din + off, dout, n, bad, G.BuildCDF(bad), G.SerializeMatrix(bad))
            assert e.value.code == L.ERR_ARG
        finally:
            lib.przm_gpu_free(g.handle, ctypes.c_void_p(din))
            lib.przm_gpu_free(g.handle, ctypes.c_void_p(dout))


def test_probes(g):
    v = g.BenchmarkVRAMBandwidth(256)
    e = g.BenchmarkD2DEntropy(64, 5)
    c = g.BenchmarkEntropyVRAM(16, 3)
    print(f"vram {v:.0f} GB/s, d2d entropy {e:.0f} GB/s, coder {c:.1f} GB/s")
    assert 500 < v < 20000 and 50 < e < 20000 and 1 < c < 20000


def test_pinned_pools(g):
    from przm_amd import gpu as G
    pool = G.NewPinnedBufferPool(4096, 2)
    a, b = pool.Get(), pool.Get()
    assert a.Ptr and b.Ptr and a.Ptr != b.Ptr and a.Bytes.size == 4096
    a.Bytes[:] = 7
    pool.Put(a)
    c = pool.Get()
    assert c.Ptr == a.Ptr and c.Bytes[100] == 7  # reused, not reallocated
    pool.Put(b)
    pool.Put(c)
    pool.Close()
    buf, p2 = g.GetPinned(5000)
    assert buf.Bytes.size == 5000 and p2.bufSize == 8192
    buf.Bytes[:] = 3
    p2.Put(buf)
    buf2, p3 = g.GetPinned(6000)
    assert buf2.Ptr == buf.Ptr  # the same power-of-two pool of the compressor
    p3.Put(buf2)
    raw, ptr = G.AllocPinnedHost(100)
    raw[:] = 1
    G.FreePinnedHost(ptr)


def test_streaming_match_equals_single_pass(g):
    """MatchLZ77Streaming's double-buffered chunks give the single pass's descriptors (which the
    parity tests pin to the oracle), with chunks of 4 MiB (6 chunks) and the default size."""
    O = oracle()
    a = D.mixture((20 << 20) + 12345, 67)
    whole, _ = g.MatchDescriptors(a)
    for chunk in (4 << 20, 0):
        st, prof = g.MatchDescriptorsStreaming(a, chunk)
        assert np.array_equal(st, whole), chunk
        assert prof.PipelinedChunks == (6 if chunk else 2)
    for b in range(0, (a.size + 65535) // 65536, 9):
        blk = a[b * 65536:(b + 1) * 65536]
        assert np.array_equal(whole[b * 65536: b * 65536 + blk.size], O.match_block(blk)), b
    m, _ = g.MatchLZ77Streaming(a[:3 << 20], 1 << 20)
    assert len(m) == int(np.count_nonzero(whole[:3 << 20]))
