"""The oracle's Section-3 prefix option (PackageOffsetsSection, core/distance/distance.go:
165-376, 692-849) against the reference's tests and decoder, on the CPU.

The known answers live in tests/offsets_cases.py (shared with the mutation tests, which
show that each rule of the option is pinned by at least one of them)."""
import json

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec
from offsets_cases import (BITS_DISTANCES, BITS_EXPECTED, TIES_PATH, copy_block, dist_freq_table, header, leb,
                           m12_decode_strict, offsets_failures)

from przm_amd import datagen as D

needs_ref = pytest.mark.skipif(not have_refdec(), reason="reference decoder not built")


def test_every_known_answer():
    assert offsets_failures(oracle(), refdec() if have_refdec() else None) == []


@needs_ref
def test_dist_code_equals_reference_function():
    """DistanceToCode (distance.go:165-180) against src/przm_dec.c's przm_distance_to_code for
    every distance a 64 KiB window can hold and a sample above it; each (code, extra) maps
    back through the decoder's przm_code_to_distance."""
    import ctypes
    O, R = oracle(), refdec()
    R.l.przm_code_to_distance.restype = ctypes.c_int32
    R.l.przm_code_to_distance.argtypes = [ctypes.c_uint8, ctypes.c_uint32, ctypes.c_void_p]
    rng = np.random.default_rng(3)
    ds = list(range(1, 65537)) + [int(v) for v in rng.integers(65537, (1 << 24) + 2, 2000)]
    for d in ds:
        c, e, nb = O.dist_code(d)
        assert (c, e, nb) == R.distance_to_code(d), d
        assert e < (1 << nb) or nb == 0 and e == 0, d
        back = ctypes.c_uint32()
        assert R.l.przm_code_to_distance(c, e, ctypes.byref(back)) == 0 and back.value == d, d


def test_bit_writer_known_answer():
    """TestBitstreamReaderWriter's values (distance_test.go:53-89) as extra bits."""
    sec, info = oracle().pack_distance_prefix(BITS_DISTANCES)
    mode, orig, n, cl, h = header(sec)
    assert (mode, n, cl) == (0x0B, 5, info[1])
    assert sec[h + cl:] == BITS_EXPECTED


@needs_ref
def test_m12_streams_decode_exactly():
    """Mode-12 code streams (distance.go:323-376): the reference decoder reads the codes back,
    and a strict decode ends on the encoder's initial state with every byte used."""
    O, R = oracle(), refdec()
    freq = dist_freq_table()
    rng = np.random.default_rng(4)
    cases = [bytes([c]) for c in range(49)] + [bytes(range(49)) * 3, bytes([0] * 300), bytes([48] * 300)]
    for n in (2, 7, 100, 1000, 16384):
        cases.append(bytes(rng.integers(0, 49, n).astype(np.uint8)))
        cases.append(bytes(np.minimum(rng.geometric(0.15, n) + 10, 48).astype(np.uint8)))
    for codes in cases:
        st = O.dist_m12(codes)
        assert R.dist4_decode(st, len(codes)) == codes
        assert m12_decode_strict(st, len(codes), freq) == codes


@needs_ref
def test_reference_package_offsets_cases_decode():
    """TestPackageOffsetsSectionModes' inputs: the containers decode to the varints."""
    O, R = oracle(), refdec()
    for n in (70, 200, 1000, 70000):
        ds = [1000] * n
        sec, m, ck = O.package_offsets(leb(ds))
        assert R.prefix_decode(sec, 5 * n + 64) == leb(ds)


@needs_ref
def test_random_offset_streams_roundtrip_and_shrink():
    """Many offset streams: the section is the container exactly when it is shorter than the
    varints, decodes to them, and its header states the varint bytes, count and code bytes."""
    O, R = oracle(), refdec()
    rng = np.random.default_rng(5)
    kinds = {0x09: 0, 0x0A: 0, 0x0B: 0}
    for it in range(600):
        n = int(rng.integers(1, 3000))
        if it % 3 == 0:
            ds = rng.integers(1, 65536, n)
        elif it % 3 == 1:
            ds = np.minimum(np.exp(rng.random(n) * np.log(65536)).astype(np.int64) + 1, 65535)
        else:
            ds = rng.choice(rng.integers(1, 65536, 8), n)
        v = leb(ds)
        sec, m, ck = O.package_offsets(v)
        kinds[m] = kinds.get(m, 0) + 1
        if m == 0x09:
            assert sec == b"\x09" + v
            cont, info = O.pack_distance_prefix(ds)
            assert len(cont) >= len(v) or info[1] == n
        else:
            assert len(sec) < len(v)
            mode, orig, cnt, cl, h = header(sec)
            assert (orig, cnt) == (len(v), n) and cl != n
            assert sec[h] in (0x0C, ord("V"))
            assert R.prefix_decode(sec, len(v) + 64) == v
    assert all(kinds[k] > 0 for k in (0x09, 0x0A, 0x0B)), kinds


def test_tie_fixtures_hold():
    """The recorded lengths still hold for the oracle (the fixtures stay meaningful)."""
    O = oracle()
    T = json.load(open(TIES_PATH))
    ds = T["vans_eq_m12"]["distances"]
    codes = np.array([O.dist_code(d)[0] for d in ds], np.uint8)
    assert len(ds) >= 64 and len(O.dist_m12(codes)) == len(O.vans(codes)) == T["vans_eq_m12"]["m12_len"]
    ds = T["prefix_eq_raw"]["distances"]
    assert len(O.pack_distance_prefix(ds)[0]) == len(leb(ds))
    ds = T["prefix_lt_raw"]["distances"]
    assert len(O.pack_distance_prefix(ds)[0]) == len(leb(ds)) - 1
    ds = T["code_len_eq_count"]["distances"]
    cont, info = O.pack_distance_prefix(ds)
    assert info[1] == len(ds) and len(cont) < len(leb(ds))


@needs_ref
def test_code_stream_as_long_as_count_would_misdecode():
    """Why mode G writes such a section flat: the reference decoder reads the container's
    code stream as raw codes and does not get the distances back."""
    O, R = oracle(), refdec()
    ds = json.load(open(TIES_PATH))["code_len_eq_count"]["distances"]
    cont, info = O.pack_distance_prefix(ds)
    try:
        got = R.prefix_decode(cont, 4096)
    except ValueError:
        got = None
    assert got != leb(ds)


def test_hazard_block_tokens_and_frame():
    O = oracle()
    hb = json.load(open(TIES_PATH))["hazard_block"]
    blk = copy_block(hb["distances"], hb["seed"])
    assert O.tokenize(blk)["off"].tobytes() == leb(hb["distances"])
    fr, info = O.encode_block(blk, 6, True)
    assert info.kind == 0 and info.s3_mode == 0x09 and info.s3_code == -1
    if have_refdec():
        assert refdec().decompress(fr) == blk.tobytes()


def _inputs():
    rng = np.random.default_rng(6)
    yield "mixture", D.mixture(4 << 20, 6)
    yield "markov", D.markov(1 << 20, 7)
    yield "text", np.frombuffer(b"the quick brown fox jumps over the lazy dog; " * 5000, np.uint8).copy()
    yield "random", rng.integers(0, 256, 100000, dtype=np.uint8)
    yield "zeros", np.zeros(70000, np.uint8)
    yield "tiny", np.frombuffer(b"abcabcabcabd", np.uint8).copy()
    yield "empty", np.zeros(0, np.uint8)


@needs_ref
def test_prefix_frames_decode_and_never_grow():
    """Whole frames with the option: the reference decoder reads them, each frame is at most
    as long as the flat-Section-3 frame and differs from it only where its Section 3 does."""
    O, R = oracle(), refdec()
    used = 0
    for name, a in _inputs():
        fp, sp = O.encode_buffer(a, 6, True)
        ff, sf = O.encode_buffer(a, 6, False)
        assert R.decompress_frames(fp, sp) == a.tobytes(), name
        assert len(sp) == len(sf) and (sp <= sf).all(), name
        for b in range(len(sp)):
            blk = a[b * 65536:(b + 1) * 65536]
            f1, i1 = O.encode_block(blk, 6, True)
            f0, i0 = O.encode_block(blk, 6, False)
            if i1.s3_mode in (0, 0x09) or i1.kind != 0:
                assert f1 == f0 or i1.kind != i0.kind, (name, b)
            else:
                used += 1
                assert i1.s3 < i0.s3 and len(f1) - len(f0) == i1.s3 - i0.s3 - (2 if i0.s3 > 65535 >= i1.s3 else 0)
    assert used > 0
