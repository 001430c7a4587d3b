"""The C ABI boundary without a GPU: the library loads, exports every entry
point include/przm_gpu.h declares, and the ctypes table binds each of them.
No compute call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "przm_gpu.h")
LIB = os.path.join(ROOT, "przm_amd", "libprzm_gpu.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(przm_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 30, names
    for must in ("przm_gpu_create", "przm_gpu_encode", "przm_gpu_encode_batch_device", "przm_gpu_match",
                 "przm_gpu_tokenize", "przm_gpu_histogram", "przm_gpu_vans_encode", "przm_serve_submit"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libprzm_gpu.so not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libprzm_gpu.so not built (run __graft_entry__.build())")
def test_ctypes_table_covers_header():
    from przm_amd import _lib
    assert set(declared()) == set(_lib.SIGNATURES), set(declared()) ^ set(_lib.SIGNATURES)
    _lib.lib()  # binds every signature


@pytest.mark.skipif(not os.path.exists(LIB), reason="libprzm_gpu.so not built (run __graft_entry__.build())")
def test_host_only_entry_points():
    """Frequency normalisation, CDF and matrix serialisation run on the host side of the ABI."""
    import numpy as np
    from przm_amd import gpu as G
    counts = np.bincount(np.frombuffer(b"abracadabra" * 50, np.uint8), minlength=256).astype(np.uint32)
    f = G.NormalizeFreq(counts)
    assert int(f.sum()) == 4096 and all((f > 0) == (counts > 0))
    c = G.BuildCDF(f)
    assert c[0] == 0 and c[-1] == 4096
    assert len(G.SerializeMatrix(f)) <= 546
    from oracle_lib import oracle
    rng = np.random.default_rng(5)
    for _ in range(20):
        cnt = (rng.geometric(0.02, 256) * (rng.random(256) < 0.6)).astype(np.uint32)
        cnt[rng.integers(0, 256)] += 1
        fo = oracle().normalize(cnt)
        assert np.array_equal(G.NormalizeFreq(cnt), fo)
        assert G.SerializeMatrix(fo) == oracle().matrix(fo)


def test_pack_match32_reference_kats():
    """TestPackUnpackMatch, gpu/gpu_test.go:210-239."""
    from przm_amd import gpu as G
    for length, offset in [(4, 1), (10, 50), (258, 65535), (4095, 1048575)]:
        packed = G.PackMatch32(length, offset)
        assert G.IsMatch32(packed)
        assert G.UnpackMatch32(packed) == (length, offset)
    assert G.PackMatch32(3, 10) == 0
    assert G.PackMatch32(10, 0) == 0


DEC_HEADER = os.path.join(ROOT, "include", "przm_decode.h")
DEC_LIB = os.path.join(ROOT, "przm_amd", "libprzm_decode.so")


@pytest.mark.skipif(not os.path.exists(DEC_LIB), reason="libprzm_decode.so not built (run __graft_entry__.build())")
def test_decoder_library_exports_its_header():
    text = re.sub(r"/\*.*?\*/", "", open(DEC_HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(przm_decode_\w+)\s*\(", text)))
    assert names == ["przm_decode_frame", "przm_decode_frames", "przm_decode_size", "przm_decode_stream",
                     "przm_decode_stream_size"], names
    lib = ctypes.CDLL(DEC_LIB)
    assert all(hasattr(lib, n) for n in names)
