"""GPU parity at the workload shapes the benchmark runs, through the C ABI.

  * the multi-chunk batched path: blocks processed in several launches of
    `chunk` blocks (per-chunk block counters, the frame-offset scan chained
    across chunks), with and without the captured graph;
  * a batch shaped like the batched workload (buffers log-uniform in
    [1 KiB, 8 MiB], mixture content) in several chunks, every buffer against
    the oracle;
  * the 1 GiB mixture buffer (16,384 blocks) through the plain entry point;
  * serving: Poisson arrivals with sizes log-uniform in [4 KiB, 16 MiB], every
    result against the oracle; and an HBM budget that admits one launch slot,
    so that requests wait for memory;
  * a frame whose raw literal section would be misread by the decoders (the
    uncompressed frame the encoder emits instead).
"""
import ctypes
import threading
import time

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec, encode_many

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


def _tuning(g, grid, chunk):
    from przm_amd import _lib as L
    assert L.lib().przm_gpu_set_tuning(g.handle, grid, chunk) == 0


def _graph(g, on):
    from przm_amd import _lib as L
    assert L.lib().przm_gpu_set_options(g.handle, 1 if on else 0) == 0


def _check_batch(g, bufs, what):
    res, _ = g.EncodeBatch(bufs)
    want = encode_many(bufs)
    for i, ((f, s), (of, osz)) in enumerate(zip(res, want)):
        assert np.array_equal(s, osz), f"{what}: buffer {i} ({bufs[i].size} bytes) frame sizes differ"
        assert f == of, f"{what}: buffer {i} ({bufs[i].size} bytes) frames differ"


def test_multichunk_batch_matches_oracle(g):
    """About 400 blocks in launches of 64 blocks (7 chunks) on 16 LZ workgroups, buffers of mixed
    sizes including empty ones and exact block multiples, graph and direct launches."""
    rng = np.random.default_rng(31)
    sizes = [0, 1, 65536, 65537, 3 * 65536, 1000, 0] + [int(x) for x in rng.integers(1, 7 * 65536, 110)]
    mix = D.mixture(sum(sizes) + 65536, 31)
    bufs, o = [], 0
    for s in sizes:
        bufs.append(mix[o:o + s].copy())
        o += s
    nblocks = sum(max((b.size + 65535) // 65536, 1) for b in bufs)
    assert nblocks > 6 * 64
    try:
        _tuning(g, 16, 64)
        _check_batch(g, bufs, "graph, chunk 64")
        _graph(g, False)
        _check_batch(g, bufs, "direct, chunk 64")
    finally:
        _graph(g, True)
        _tuning(g, g.SMCount, 8192)


def test_batched_workload_shape_matches_oracle(g):
    """150 buffers log-uniform in [1 KiB, 8 MiB] of mixture content in launches of 1024 blocks."""
    arena, offs, sizes = D.batch(150, 1024, 8 << 20, 13)
    bufs = [arena[int(o): int(o) + int(s)] for o, s in zip(offs, sizes)]
    nblocks = sum(max((b.size + 65535) // 65536, 1) for b in bufs)
    assert nblocks > 2 * 1024, nblocks
    try:
        _tuning(g, g.SMCount, 1024)
        _check_batch(g, bufs, "150 buffers, chunk 1024")
    finally:
        _tuning(g, g.SMCount, 8192)


def test_mixture_1gib_plain(g):
    """The 1 GiB mixture buffer: 16,384 frames whose sizes add up, all through the library's own
    decoder, every 37th through the reference decoder, every 8th against the oracle and the
    first 64 byte for byte."""
    from przm_amd import decode as X
    a = D.mixture(1 << 30, 2)
    frames, sizes, prof = g.CompressFrames(a)
    assert sizes.size == 16384 and int(sizes.sum()) == len(frames)
    assert X.DecompressFrames(frames, sizes, a.size) == a.tobytes()  # every frame, the library's own decoder
    f = np.frombuffer(frames, np.uint8)
    offs = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))])
    if have_refdec():
        R = refdec()
        for b in range(0, 16384, 37):
            assert R.decompress(f[offs[b]: offs[b + 1]]) == a[b * 65536:(b + 1) * 65536].tobytes(), b
    idx = list(range(0, 16384, 8))
    want = encode_many([a[b * 65536:(b + 1) * 65536] for b in idx])
    for b, (of, osz) in zip(idx, want):
        assert osz[0] == sizes[b] and f[offs[b]: offs[b + 1]].tobytes() == of, f"block {b}"
    of, osz = oracle().encode_buffer(a[: 64 * 65536])
    assert np.array_equal(osz, sizes[:64]) and of == frames[: int(sizes[:64].sum())]


def test_serving_poisson_matches_oracle(g):
    """Open-loop Poisson arrivals (150 req/s for 2 s), sizes log-uniform in [4 KiB, 16 MiB]: each
    served result equals the oracle's frames for that request alone."""
    from przm_amd import serve as S
    pool = D.mixture(64 << 20, 41)
    rng = np.random.default_rng(41)
    n = 300
    arrivals = np.cumsum(rng.exponential(1 / 150.0, n))
    sizes = np.exp(rng.uniform(np.log(4 << 10), np.log(16 << 20), n)).astype(np.int64)
    offs = (rng.uniform(0, 1, n) * (pool.size - sizes)).astype(np.int64)
    reqs = [pool[o:o + s] for o, s in zip(offs, sizes)]
    srv = S.Server(g, level=6, hbm_budget=64 << 30, max_batch_bytes=64 << 20, streams=3)
    ids = [None] * n
    outs = [None] * n
    err = []

    def producer():
        t0 = time.perf_counter()
        for i in range(n):
            w = arrivals[i] - (time.perf_counter() - t0)
            if w > 0:
                time.sleep(w)
            ids[i] = srv.submit(reqs[i])

    def consumer():
        try:
            for i in range(n):
                while ids[i] is None:
                    time.sleep(0.0002)
                frames, fs, res = srv.wait(ids[i])
                outs[i] = (frames, fs)
        except Exception as e:
            err.append(e)

    tp, tc = threading.Thread(target=producer), threading.Thread(target=consumer)
    tp.start(); tc.start(); tp.join(); tc.join()
    st = srv.stats()
    srv.close()
    assert not err, err
    assert st["done"] == n and st["reserved"] <= 64 << 30
    want = encode_many(reqs)
    for i, ((f, fs), (of, osz)) in enumerate(zip(outs, want)):
        assert np.array_equal(fs, osz) and f == of, f"request {i} ({sizes[i]} bytes)"


def test_serving_budget_admits_one_slot(g):
    """A budget that holds one launch slot but not two: one launch at a time (none overlaps
    another), requests wait for it, the reserved device bytes stay within the budget, and every
    result is still the oracle's."""
    from przm_amd import serve as S
    from przm_amd.gpu import GPUError
    from przm_amd import _lib as L
    probe = S.Server(g, level=6, hbm_budget=64 << 30, max_batch_bytes=4 << 20, streams=2)
    per_slot = probe.stats()["reserved"] // 2
    probe.close()
    budget = per_slot + per_slot // 2
    arena, offs, sizes = D.batch(40, 64 << 10, 2 << 20, 43)
    bufs = [arena[int(o): int(o) + int(s)] for o, s in zip(offs, sizes)]
    srv = S.Server(g, level=6, hbm_budget=budget, max_batch_bytes=4 << 20, streams=3)
    ids = [srv.submit(b) for b in bufs]
    outs = [srv.wait(i) for i in ids]
    st = srv.stats()
    with pytest.raises(GPUError) as e:  # one request above max_batch_bytes
        srv.submit(np.zeros((4 << 20) + 1, np.uint8))
    assert e.value.code == L.ERR_CAPACITY
    srv.close()
    assert st["reserved"] == per_slot and st["reserved"] <= budget
    assert st["overlapped"] == 0 and st["done"] == len(bufs)
    assert st["launches"] >= int(sizes.sum()) // (4 << 20)  # launches of at most 4 MiB, one at a time
    waited = sum(1 for _, _, r in outs if r.start_ms > r.submit_ms + 1.0)
    assert waited > len(bufs) // 2, waited
    want = encode_many(bufs)
    for i, ((f, fs, _), (of, osz)) in enumerate(zip(outs, want)):
        assert np.array_equal(fs, osz) and f == of, i
    with pytest.raises(GPUError) as e:  # not even one slot fits
        S.Server(g, level=6, hbm_budget=per_slot // 2, max_batch_bytes=4 << 20, streams=1)
    assert e.value.code == L.ERR_CAPACITY


def test_hazard_frame_is_uncompressed(g):
    """Literals that would start 'V','A',0x01 as a raw Section 4: the GPU emits the same
    uncompressed frame as the oracle (frame kind 3)."""
    from test_oracle_kats import hazard_input
    a = hazard_input()
    frames, sizes, _ = g.CompressFrames(a)
    of, osz = oracle().encode_buffer(a)
    _, info = oracle().encode_block(a)
    assert info.kind == 3
    assert frames == of and frames[4] == 0x04 and np.array_equal(sizes, osz)


def _shape(g, sb, rj):
    from przm_amd import _lib as L
    assert L.lib().przm_gpu_set_kernel_shape(g.handle, sb, rj) == 0


def test_kernel_shapes_give_identical_frames(g):
    """Every shape the autotuner may pick (chain candidates per search step 2 / 4 / 8, coder jobs
    per workgroup 16 / 32) gives the default shape's frames, on every chain-depth class."""
    arena, offs, sizes = D.batch(40, 1024, 1 << 20, 71)
    bufs = [arena[int(o): int(o) + int(s)] for o, s in zip(offs, sizes)]
    deep = np.frombuffer(b"The quick brown fox jumps over the lazy dog. " * 1500, np.uint8).copy()
    ref_batch, _ = g.EncodeBatch(bufs)
    ref_lv = {lv: g.CompressFrames(deep, lv)[0] for lv in (1, 3, 6, 10, 11)}
    try:
        for sb, rj in ((2, 16), (8, 16), (4, 32), (2, 32), (8, 32)):
            _shape(g, sb, rj)
            res, _ = g.EncodeBatch(bufs)
            for i, ((f, s), (rf, rs)) in enumerate(zip(res, ref_batch)):
                assert f == rf and np.array_equal(s, rs), (sb, rj, i)
            for lv, want in ref_lv.items():
                assert g.CompressFrames(deep, lv)[0] == want, (sb, rj, lv)
    finally:
        _shape(g, 4, 16)
    assert ref_lv[11] == oracle().encode_buffer(deep, 11)[0]


def test_autotune_picks_a_valid_shape(g):
    from przm_amd import _lib as L
    lib = L.lib()
    arena, offs, sizes = D.batch(60, 1024, 2 << 20, 72)
    d = ctypes.c_void_p()
    assert lib.przm_gpu_malloc(g.handle, arena.size + 64, ctypes.byref(d)) == 0
    try:
        lib.przm_gpu_h2d(g.handle, d, ctypes.c_void_p(arena.ctypes.data), arena.size)
        spans = (L.Span * len(sizes))()
        for i, (o, s) in enumerate(zip(offs, sizes)):
            spans[i].offset, spans[i].len = int(o), int(s)
        tuned = (ctypes.c_int32 * 4)()
        assert lib.przm_gpu_autotune(g.handle, d, spans, len(sizes), 6, tuned) == 0
        assert tuned[0] in (g.SMCount, 2 * g.SMCount) and tuned[1] in (16384, 32768, 65536)
        assert tuned[2] in (16, 32) and tuned[3] in (2, 4, 8)
        got = (ctypes.c_int32 * 4)()
        lib.przm_gpu_get_tuning(g.handle, got)
        assert list(got) == list(tuned)
        bufs = [arena[int(o): int(o) + int(s)] for o, s in zip(offs, sizes)]
        res, _ = g.EncodeBatch(bufs)
        want = encode_many(bufs)
        assert all(f == of for (f, _), (of, _) in zip(res, want))
    finally:
        lib.przm_gpu_free(g.handle, d)
        _tuning(g, g.SMCount, 8192)
        _shape(g, 4, 16)
