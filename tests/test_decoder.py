"""The library's own CPU decoder (include/przm_decode.h), on the CPU.

  * round trips of the oracle's frames (every frame kind, section choice and
    chain depth the encoder produces), byte-equal to the reference C99
    decoder's output on the same frames;
  * the reference decoder test suite's header and frame cases
    (tests/test_przm_dec.c:14-87), restated as data;
  * hand-built frames for each error the decoder reports, and corrupted
    frames (every truncation, byte flips) that must fail cleanly.
"""
import json
import os

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec
from offsets_cases import copy_block
from sheaf_cases import lits_frame
from tokenizer_kats import KATS

from przm_amd import datagen as D
from przm_amd import decode as X

HERE = os.path.dirname(os.path.abspath(__file__))
needs_ref = pytest.mark.skipif(not have_refdec(), reason="reference decoder not built")


def _inputs():
    rng = np.random.default_rng(81)
    yield "empty", np.zeros(0, np.uint8)
    for n in (1, 3, 4, 5, 12, 13, 100, 1024, 4097, 65535, 65536, 65537):
        yield f"markov{n}", D.markov(n, n)
    yield "zeros", np.zeros(200000, np.uint8)
    yield "ab", np.frombuffer(b"ab" * 50000, np.uint8).copy()
    yield "random", rng.integers(0, 256, 70000, dtype=np.uint8)
    yield "lowalpha", rng.integers(0, 3, 70000, dtype=np.uint8)
    yield "runs", np.repeat(rng.integers(0, 256, 3000, dtype=np.uint8), rng.integers(1, 40, 3000))[:90000].copy()
    yield "mixture", D.mixture(2 << 20, 81)


def test_roundtrip_oracle_frames():
    O = oracle()
    for name, a in _inputs():
        frames, sizes = O.encode_buffer(a)
        assert X.DecompressFrames(frames, sizes) == a.tobytes(), name


def _prefix_blocks():
    """Blocks whose Section 3 is a distance-prefix container or hits one of its edge rules."""
    T = json.load(open(os.path.join(HERE, "golden", "offsets_ties.json")))
    for key in ("hazard_block", "renorm_block"):
        yield key, copy_block(T[key]["distances"], T[key]["seed"])


def test_roundtrip_prefix_offset_frames():
    """Frames whose Section 3 the oracle writes as PackageOffsetsSection chooses it."""
    O = oracle()
    used = 0
    for name, a in list(_inputs()) + list(_prefix_blocks()):
        frames, sizes = O.encode_buffer(a, 6, True)
        assert X.DecompressFrames(frames, sizes) == a.tobytes(), name
        f = np.frombuffer(frames, np.uint8)
        pos = 0
        for sz in sizes:
            fr = f[pos: pos + int(sz)]
            pos += int(sz)
            if fr.size > 12 and fr[4] & 0x04 == 0:
                d = 12 + (2 if fr[4] & 0x08 else 0)
                lc = int.from_bytes(fr[8:11].tobytes() if fr[4] & 0x08 else fr[8:10].tobytes(), "little")
                s3 = d + 8 + lc
                used += int(fr[s3] in (0x06, 0x0A, 0x0B) and lc < fr.size - s3)
    assert used > 10


@pytest.mark.parametrize("level", [1, 3, 5, 9, 11])
def test_roundtrip_levels(level):
    a = D.mixture(1 << 20, 82 + level)
    frames, sizes = oracle().encode_buffer(a, level)
    assert X.DecompressFrames(frames, sizes) == a.tobytes()


def test_roundtrip_golden_fixtures():
    man = json.load(open(os.path.join(HERE, "golden", "manifest.json")))
    for name, case in man.items():
        src = open(os.path.join(HERE, "golden", case["input"]), "rb").read()
        frames = open(os.path.join(HERE, "golden", name + ".frames"), "rb").read()
        assert X.DecompressFrames(frames, case["frame_sizes"]) == src, name


def test_roundtrip_sheaf_literal_frames():
    """Frames whose Section 4 the oracle writes as PackageLiteralsSection chooses it (sheaf
    containers of every mode among them), and containers of each mode on their own."""
    O = oracle()
    kinds = set()
    rng = np.random.default_rng(88)
    text = np.frombuffer(bytes(rng.choice(list(b'{"name": "value_1", "n": 12.5}, \n'), 400000).tolist()),
                         np.uint8).copy()
    for name, a in list(_inputs()) + [("json", text)]:
        frames, sizes = O.encode_buffer(a, 6, True, True)
        assert X.DecompressFrames(frames, sizes) == a.tobytes(), name
        for b in range(max(1, (a.size + 65535) // 65536)):
            kinds.add(O.encode_block(a[b * 65536:(b + 1) * 65536], 6, False, True)[1].s4_kind)
    assert kinds & {2, 3, 4, 5}, kinds
    for m in range(4):
        for n in (1, 7, 300, 5000):
            lits = bytes(rng.choice(list(b'{"k": "v_1", "n": 1.5}, \n abc'), n).tolist())
            assert X.DecompressFrame(lits_frame(lits, O.pack_sheaf(lits, m))) == lits, (m, n)


@needs_ref
@pytest.mark.parametrize("prefix,sheaf", [(False, False), (True, False), (True, True)])
def test_equals_reference_decoder(prefix, sheaf):
    """Frame by frame, the same bytes as the reference C99 decoder."""
    O, R = oracle(), refdec()
    for name, a in list(_inputs()) + list(_prefix_blocks()):
        frames, sizes = O.encode_buffer(a, 6, prefix, sheaf)
        f = np.frombuffer(frames, np.uint8)
        pos = 0
        for s in sizes:
            fr = f[pos: pos + int(s)]
            assert X.DecompressFrame(fr) == R.decompress(fr, 65536), name
            pos += int(s)


def test_reference_suite_header_and_frame_cases():
    # przm_get_uncompressed_size: 'P','R','Z','M', flags 0, size 0x004E20 = 20000
    hdr = bytes([0x50, 0x52, 0x5A, 0x4D, 0, 0x20, 0x4E, 0x00])
    assert X.GetUncompressedSize(hdr) == 20000
    with pytest.raises(X.DecodeError) as e:
        X.GetUncompressedSize(hdr[:7])
    assert e.value.code == X.ERR_CORRUPT
    with pytest.raises(X.DecodeError) as e:
        X.GetUncompressedSize(b"XYZW\0\0\0\0")
    assert e.value.code == X.ERR_MAGIC
    # an uncompressed frame of 25 'A'
    assert X.DecompressFrame(b"PRZM\x04\x19\x00\x00" + b"A" * 25) == b"A" * 25
    # 8 literals: directory (1, 0), Section 1 with window log 16, a raw Section 2 of one opcode
    # 0xFA, Section 4 "ABCDEFGH"
    frame = b"PRZM\x00\x08\x00\x00" + b"\x01\x00\x00\x00" + bytes([0, 16, 0, 0, 0, 0, 0, 0]) + b"\xFA" + b"ABCDEFGH"
    assert X.DecompressFrame(frame) == b"ABCDEFGH"


def _frame(u, cmd, off=b"", lit=b"", flags=0, s2_mode=b"\x09", s3_mode=b"\x09"):
    s2 = s2_mode + cmd if cmd else b""
    s3 = s3_mode + off if off else b""
    return (b"PRZM" + bytes([flags]) + u.to_bytes(3, "little") + len(s2).to_bytes(2, "little")
            + len(s3).to_bytes(2, "little") + bytes([0, 16, 0, 0, 0, 0, 0, 0]) + s2 + s3 + lit)


@pytest.mark.parametrize("k", KATS, ids=[k["name"] for k in KATS])
def test_hand_worked_streams_decode(k):
    """The tokenizer's hand-worked known answers, as flat-section frames."""
    if not k.get("decodes", True):
        pytest.skip("the case's descriptor is longer than the bytes it covers")
    lit = k["lit"]
    if lit[:3] == b"VA\x01" or (len(lit) >= 8 and int.from_bytes(lit[4:8], "little") ==
                                (int.from_bytes(lit[:4], "little") + 1) // 2):
        pytest.skip("a raw literal section that reads as a container")
    f = _frame(len(k["src"]), bytes.fromhex(k["cmd"]), bytes.fromhex(k["off"]), lit)
    assert X.DecompressFrame(f) == k["src"]


def _code(frame):
    with pytest.raises(X.DecodeError) as e:
        X.DecompressFrame(frame)
    return e.value.code


def test_errors():
    # explicit distance 5 with nothing written yet, and distance 0
    assert _code(_frame(4, b"\x80", b"\x05")) == X.ERR_RANGE
    assert _code(_frame(8, b"\xF6\x80", b"\x00", b"abcd")) == X.ERR_RANGE
    # a match running past the frame's size, and commands ending short of it
    assert _code(_frame(6, b"\xF6\x80", b"\x04", b"abcd")) == X.ERR_SIZE
    assert _code(_frame(9, b"\xF6", b"", b"abcd")) == X.ERR_SIZE
    # literals the section does not hold
    assert _code(_frame(8, b"\xFA", b"", b"abcd")) == X.ERR_CORRUPT
    # opcodes and flags outside the subset: 0xE9, repeat register 3, stride, dictionary
    assert _code(_frame(8, b"\xF6\xE9\x04", b"\x04", b"abcd")) == X.ERR_UNSUPPORTED
    assert _code(_frame(8, b"\x43", b"", b"abcd")) == X.ERR_UNSUPPORTED
    assert _code(_frame(4, b"\xF6", b"", b"abcd", flags=0x01)) == X.ERR_UNSUPPORTED
    assert _code(_frame(4, b"\xF6", b"", b"abcd", flags=0x10)) == X.ERR_UNSUPPORTED
    # an offset section mode the format does not define, and a short prefix container
    assert _code(_frame(8, b"\xF6\x80", b"\x04", b"abcd", s3_mode=b"\x0C")) == X.ERR_UNSUPPORTED
    assert _code(_frame(8, b"\xF6\x80", b"\x04", b"abcd", s3_mode=b"\x06")) == X.ERR_CORRUPT
    # prefix containers (0x0B: varint bytes, count, code bytes) holding the one distance 4
    # (code 3, no extra bits): raw codes, then a code above 48, a varint count that does not
    # match, a Mode-12 stream with a byte left over, and extra bits that run out
    ok = b"\x01\x01\x01\x03"
    assert X.DecompressFrame(_frame(8, b"\xF6\x80", ok, b"abcd", s3_mode=b"\x0B")) == b"abcdabcd"
    assert _code(_frame(8, b"\xF6\x80", b"\x01\x01\x01\x31", b"abcd", s3_mode=b"\x0B")) == X.ERR_CORRUPT
    assert _code(_frame(8, b"\xF6\x80", b"\x02\x01\x01\x03", b"abcd", s3_mode=b"\x0B")) == X.ERR_CORRUPT
    m12 = oracle().dist_m12(bytes([3]))
    assert X.DecompressFrame(_frame(8, b"\xF6\x80", bytes([1, 1, len(m12)]) + m12, b"abcd",
                                    s3_mode=b"\x0B")) == b"abcdabcd"
    assert _code(_frame(8, b"\xF6\x80", bytes([1, 1, len(m12) + 1]) + m12 + b"\x00", b"abcd",
                        s3_mode=b"\x0B")) == X.ERR_CORRUPT
    # distance 1000 = code 19 + 231 in 8 extra bits, with the extra byte missing
    assert _code(_frame(8, b"\xF6\x80", b"\x02\x01\x01\x13", b"abcd", s3_mode=b"\x0B")) == X.ERR_CORRUPT
    # a V-ANS literal section whose weights sum to 4095
    vans = bytearray(oracle().vans(np.frombuffer(b"abcd" * 10, np.uint8)))
    assert vans[8:10] == b"\x40\x00" and vans[10:12] == b"\x1E\x00"  # row 6, columns 1-4
    assert vans[12:14] == b"\x84\x00"  # weight 1024 as two bytes: make it 1023
    vans[12:14] = b"\x83\xFF"
    assert _code(_frame(40, b"\xF0\x28", b"", bytes(vans))) == X.ERR_CORRUPT
    assert _code(b"PRZM\x04\x19\x00\x00" + b"A" * 24) == X.ERR_CORRUPT  # uncompressed, short


def test_corrupted_frames_fail_cleanly():
    """Every truncation and many byte flips of real frames: an error or some output, never a
    read or write outside the buffers (the decoder bounds-checks every access)."""
    O = oracle()
    rng = np.random.default_rng(83)
    frames = []
    T = json.load(open(os.path.join(HERE, "golden", "offsets_ties.json")))
    rb = copy_block(T["renorm_block"]["distances"], T["renorm_block"]["seed"])
    js = np.frombuffer(bytes(rng.choice(list(b'{"name": "value_1", "n": 12.5}, \n'), 65536).tolist()), np.uint8)
    for a, pre in ((D.mixture(65536, 84), False), (D.markov(30000, 85), False),
                   (np.frombuffer(b"abcabcabd" * 2000, np.uint8).copy(), False), (D.mixture(65536, 86), True),
                   (rb, True), (js.copy(), True)):
        f, s = O.encode_buffer(a, 6, pre, pre)
        frames.append((np.frombuffer(f, np.uint8), a))
    for f, a in frames:
        for cut in range(0, f.size, max(1, f.size // 300)):
            try:
                X.DecompressFrame(f[:cut])
            except X.DecodeError:
                pass
        for _ in range(300):
            g = f.copy()
            for _ in range(int(rng.integers(1, 4))):
                g[int(rng.integers(0, g.size))] ^= np.uint8(1 << int(rng.integers(0, 8)))
            try:
                X.DecompressFrame(g)
            except X.DecodeError:
                pass
