"""Pins the C oracle of the GPU encode mode (CPU only).

Each check is chosen so a plausible slip in the oracle (a dropped term, a
wrong index or tie-break, a transposed operand) fails it:
  * hash4: closed form recomputed with numpy uint32 arithmetic + KATs;
  * chain search: brute-force enumeration of every earlier same-hash
    position on small inputs;
  * tokenizer / sections / frames: decoded by the reference C99 decoder
    (compiled from the reference sources into oracle/_ref), plus the
    reference's own split of the command stream;
  * normalisation: hand-worked known answers from literals.go:78-140,
    invariants (sum 4096, presence, monotonicity) and decoder acceptance;
  * V-ANS and Meta8: decoded by the reference decoder; final states and
    stream layout checked against an independent Python decoder.
"""
import os

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec

from przm_amd import datagen as D

needs_ref = pytest.mark.skipif(not have_refdec(), reason="reference decoder not built")
MUL = np.uint32(0x1E35A7BD)


def np_hash4(a, p):
    w = (a[p].astype(np.uint32) | (a[p + 1].astype(np.uint32) << 8) | (a[p + 2].astype(np.uint32) << 16)
         | (a[p + 3].astype(np.uint32) << 24))
    with np.errstate(over="ignore"):
        return (w * MUL) >> np.uint32(16)


def test_hash4_closed_form():
    O = oracle()
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, 4096, dtype=np.uint8)
    for p in range(0, 4092, 7):
        assert O.hash4(a[p:p + 4]) == int(np_hash4(a, np.array([p]))[0])
    # KATs: LE32 word w -> (w * 0x1E35A7BD mod 2^32) >> 16
    assert O.hash4(b"\x00\x00\x00\x00") == 0
    assert O.hash4(b"\x01\x00\x00\x00") == 0x1E35
    assert O.hash4(b"\xff\xff\xff\xff") == ((0xFFFFFFFF * 0x1E35A7BD) & 0xFFFFFFFF) >> 16


def test_search_depth_levels():
    O = oracle()
    want = {1: 2, 2: 4, 3: 4, 4: 8, 5: 8, 6: 16, 7: 16, 8: 32, 9: 32, 10: 64, 11: 128, 0: 16}
    for lv, d in want.items():
        assert O.l.orc_search_depth(lv) == d


def brute_desc(a, depth):
    """Every earlier position with the same hash, newest first, first `depth`."""
    n = a.size
    h = np_hash4(a, np.arange(max(n - 3, 0))) if n >= 4 else np.zeros(0, np.uint32)
    desc = np.zeros(n, np.uint32)
    for p in range(n - 3):
        cands = [q for q in range(p - 1, -1, -1) if h[q] == h[p]][:depth]
        cap = min(n - p, 1024)
        best, bd = 0, 0
        for q in cands:
            l = 0
            while l < cap and a[p + l] == a[q + l]:
                l += 1
            if l > best:
                best, bd = l, p - q
        if best >= 4:
            desc[p] = (best << 16) | bd
    return desc


@pytest.mark.parametrize("kind,n,depth", [("markov", 1500, 16), ("rep", 1200, 4), ("mixture", 2000, 2),
                                          ("random", 700, 16), ("rep", 1500, 32)])
def test_match_block_brute_force(kind, n, depth):
    O = oracle()
    if kind == "markov":
        a = D.markov(n, 11)
    elif kind == "mixture":
        a = D.mixture(65536, 3)[:n].copy()
    elif kind == "random":
        a = np.random.default_rng(5).integers(0, 4, n, dtype=np.uint8)
    else:
        a = np.frombuffer((b"abcabcabd" * 400)[:n], np.uint8).copy()
        a[::97] = 0x7A
    assert np.array_equal(O.match_block(a, depth=depth), brute_desc(a, depth))


def _blocks():
    yield D.markov(65536, 21)
    yield D.mixture(4 << 20, 7)[3 * 65536: 4 * 65536].copy()  # a table or json block
    yield np.zeros(65536, np.uint8)
    yield np.frombuffer(b"ab" * 20000, np.uint8).copy()
    yield np.frombuffer(b"The quick brown fox " * 1000, np.uint8).copy()
    yield np.random.default_rng(3).integers(0, 256, 5000, dtype=np.uint8)
    yield np.frombuffer(b"x" * 11 + b"y" * 12 + b"z" * 13 + b"abcdabcdabcdabcd" + b"\0" * 70000, np.uint8)[:65536].copy()
    for n in (1, 3, 4, 5, 11, 12, 13, 100, 1023, 1024, 1025, 4097, 16385):
        yield D.markov(n, n)


@needs_ref
def test_frames_roundtrip_reference_decoder():
    O, R = oracle(), refdec()
    for a in _blocks():
        frame, info = O.encode_block(a)
        assert R.decompress(frame, a.size) == a.tobytes(), (a.size, info.kind)


@needs_ref
def test_buffers_roundtrip_configs():
    O, R = oracle(), refdec()
    for a in (D.markov(1 << 20, 1), D.mixture(2 << 20, 2), D.markov(200000, 4)):
        frames, sizes = O.encode_buffer(a)
        assert len(sizes) == (a.size + 65535) // 65536
        assert R.decompress_frames(frames, sizes) == a.tobytes()
    frames, sizes = O.encode_buffer(np.zeros(0, np.uint8))
    assert frames == b"PRZM\x04\x00\x00\x00" and list(sizes) == [8]


def test_frame_header_fields():
    O = oracle()
    a = D.markov(65536, 8)
    f, info = O.encode_block(a)
    assert f[:4] == b"PRZM" and info.kind == 0
    flags = f[4]
    assert int.from_bytes(f[5:8], "little") == 65536 % (1 << 24) == 65536
    l2 = int.from_bytes(f[8:10], "little")
    l3 = int.from_bytes(f[10:12], "little")
    s1 = f[12:20]
    assert s1[1] == 16 and s1[0] == 0 and s1[2:] == b"\0" * 6  # compress.go:952, byte 5 must be 0
    assert l2 == info.s2 and l3 == info.s3 and len(f) == 20 + l2 + l3 + info.s4
    assert bool(flags & 0x80) == bool(info.s4v) and not flags & 0x08
    assert f[20] in (0x08, 0x09) and (l3 == 0 or f[20 + l2] == 0x09)
    # small block: window log 10 (compress.go:183-202)
    g, _ = O.encode_block(D.markov(1000, 2))
    if g[4] != 0x04:
        assert g[12 + 1] == 10


def test_uncompressed_guards():
    O = oracle()
    rnd = np.random.default_rng(9).integers(0, 256, 65536, dtype=np.uint8)
    f, info = O.encode_block(rnd)
    assert f[4] == 0x04 and f[8:] == rnd.tobytes() and info.kind in (1, 2)
    f, info = O.encode_block(np.zeros(0, np.uint8))
    assert f == b"PRZM\x04\x00\x00\x00"


def test_high_entropy_sample_rule():
    O = oracle()
    # compress.go:113-132: sample of min(n, 512) bytes, threshold 225 distinct when the sample >= 256
    a = np.arange(512, dtype=np.uint32).astype(np.uint8)  # 256 distinct > 225
    assert O.l.orc_check_high_entropy(O_ptr(a), a.size) == 1
    b = (np.arange(512) % 225).astype(np.uint8)  # exactly 225 distinct: not above the threshold
    assert O.l.orc_check_high_entropy(O_ptr(b), b.size) == 0
    c = np.arange(63, dtype=np.uint8)  # sample < 64 never counts
    assert O.l.orc_check_high_entropy(O_ptr(c), c.size) == 0
    d = np.arange(100, dtype=np.uint8)  # 100 distinct > 100*88/100 = 88
    assert O.l.orc_check_high_entropy(O_ptr(d), d.size) == 1


def O_ptr(a):
    import ctypes
    return ctypes.c_void_p(a.ctypes.data)


@needs_ref
def test_split_matches_reference_split():
    O, R = oracle(), refdec()
    for a in _blocks():
        s = O.tokenize(a)
        op, meta, var = O.split(s["cmd"])
        assert np.array_equal(op, s["op"]) and np.array_equal(meta, s["meta"]) and np.array_equal(var, s["var"])
        rop, rmeta, rvar = R.split(s["cmd"])
        assert np.array_equal(op, rop) and np.array_equal(meta, rmeta) and np.array_equal(var, rvar)


def test_split_kat_mode_g_opcodes():
    # the mode-G opcode families of core/commands/commands_test.go:17-58 (fused repeat,
    # explicit with / without LEB128, long repeat with / without LEB128, literal / RLE runs)
    O = oracle()
    cmd = bytes([0x05, 0x12, 0x43, 0x7E, 0x82, 0x8F, 0x85, 0x02, 0xC0, 0x24, 0xC2, 0x3F, 0x81, 0x01,
                 0xF0, 0x8A, 0x01, 0xF1, 0x64, 0xF3, 0xF6, 0xFA])
    op, meta, var = O.split(cmd)
    assert op.tobytes() == bytes([0x05, 0x12, 0x43, 0x7E, 0x82, 0x8F, 0xC0, 0xC2, 0xF0, 0xF1, 0xF3, 0xF6, 0xFA])
    assert meta.tobytes() == bytes([0x24, 0x3F])
    assert var.tobytes() == bytes([0x85, 0x02, 0x81, 0x01, 0x8A, 0x01, 0x64])


def test_tokenizer_streams_invariants():
    """Literal bytes + run bytes account for the input; opcode fields in range
    (docs/PRZM_SPEC.md:504-526); explicit offsets are LEB128."""
    O = oracle()
    for a in _blocks():
        s = O.tokenize(a)
        cmd = s["cmd"]
        i, covered, lits_used, offs = 0, 0, 0, 0
        while i < cmd.size:
            c = int(cmd[i]); i += 1
            if c < 0x80:
                covered += ((c >> 4) & 7) + ((c >> 2) & 3) + 4
                lits_used += (c >> 4) & 7
                assert ((c >> 2) & 3) + 4 <= 7
            elif c < 0xC0:
                L = (c & 15) + 4
                if c & 15 == 15:
                    v, sh = 0, 0
                    while True:
                        b = int(cmd[i]); i += 1; v |= (b & 0x7F) << sh; sh += 7
                        if not b & 0x80:
                            break
                    L += v
                covered += ((c >> 4) & 3) + L
                lits_used += (c >> 4) & 3
                offs += 1
            elif c <= 0xC3:
                t = int(cmd[i]); i += 1
                L = (t & 15) + 4
                if t & 15 == 15:
                    v, sh = 0, 0
                    while True:
                        b = int(cmd[i]); i += 1; v |= (b & 0x7F) << sh; sh += 7
                        if not b & 0x80:
                            break
                    L += v
                covered += (t >> 4) + L
                lits_used += t >> 4
            elif c in (0xF0, 0xF1):
                v, sh = 0, 0
                while True:
                    b = int(cmd[i]); i += 1; v |= (b & 0x7F) << sh; sh += 7
                    if not b & 0x80:
                        break
                covered += v
                lits_used += v if c == 0xF0 else 1
            else:
                assert 0xF3 <= c <= 0xFA
                covered += c - 0xF2
                lits_used += c - 0xF2
        assert covered == a.size and lits_used == s["lit"].size
        # offsets section: count of LEB128 values == explicit opcodes
        o = s["off"]
        assert int(np.sum(o < 0x80)) == offs


def test_normalize_hand_worked_answers():
    O = oracle()
    c = np.zeros(256, np.uint32)
    c[ord("a")] = c[ord("b")] = c[ord("c")] = 1
    # unique 3, remaining 4093, share 4093*1//3 = 1364 -> 1365 each, sum 4095, one point to the
    # highest count, lowest symbol first: 'a'
    f = O.normalize(c)
    assert f[ord("a")] == 1366 and f[ord("b")] == 1365 and f[ord("c")] == 1365
    c = np.zeros(256, np.uint32)
    c[7] = 5
    assert O.normalize(c)[7] == 4096 and O.normalize(c).sum() == 4096
    c = np.ones(256, np.uint32)
    assert (O.normalize(c) == 16).all()
    c = np.zeros(256, np.uint32)
    c[10], c[20], c[30] = 3, 1, 3
    # unique 3, rem 4093, total 7: 1+1754=1755, 1+584=585, 1755 -> 4095; +1 to symbol 10 (count 3, lower index)
    f = O.normalize(c)
    assert (f[10], f[20], f[30]) == (1756, 585, 1755)


def test_normalize_invariants():
    O = oracle()
    rng = np.random.default_rng(2)
    for t in range(300):
        k = int(rng.integers(1, 257))
        c = np.zeros(256, np.uint32)
        idx = rng.choice(256, k, replace=False)
        c[idx] = rng.integers(1, 5000, k)
        f = O.normalize(c)
        assert int(f.sum()) == 4096
        assert ((f > 0) == (c > 0)).all()
        order = np.argsort(c[idx])
        cs, fs = c[idx][order], f[idx][order]
        assert (np.diff(fs.astype(np.int64))[np.diff(cs.astype(np.int64)) > 0] >= 0).all()
        ideal = c.astype(np.float64) * 4096 / c.sum()
        assert (np.abs(f - ideal) < 2 + k / 100.0 + 1).all()


@needs_ref
def test_vans_roundtrip_and_layout():
    O, R = oracle(), refdec()
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 4, 5, 17, 100, 4096, 30000):
        a = rng.geometric(0.2, n).clip(0, 255).astype(np.uint8)
        v = O.vans(a)
        assert v[:4] == b"VA\x01\x00" and int.from_bytes(v[4:8], "little") == n
        assert R.vans_decode(v, n) == a.tobytes()
        assert py_vans_decode(v, n) == a.tobytes()


def py_vans_decode(v, n):
    """Independent decoder (docs/PRZM_SPEC.md:582-647) that also checks the
    initial state 65536 of every lane and the byte count."""
    pos = 8
    rows = int.from_bytes(v[pos:pos + 2], "little"); pos += 2
    freq = [0] * 256
    for h in range(16):
        if not rows >> h & 1:
            continue
        cols = int.from_bytes(v[pos:pos + 2], "little"); pos += 2
        for l in range(16):
            if not cols >> l & 1:
                continue
            b = v[pos]; pos += 1
            w = b if b < 128 else ((b & 0x7F) << 8) | v[pos]
            if b >= 128:
                pos += 1
            freq[(h << 4) | l] = w
    assert sum(freq) == 4096
    cdf = np.concatenate([[0], np.cumsum(freq)])
    sym = np.repeat(np.arange(256), freq)
    x = [int.from_bytes(v[pos + 4 * j: pos + 4 * j + 4], "little") for j in range(4)]
    pos += 16
    out = bytearray()
    for i in range(n):
        j = i & 3
        s = int(sym[x[j] & 4095])
        out.append(s)
        x[j] = freq[s] * (x[j] >> 12) + (x[j] & 4095) - int(cdf[s])
        while x[j] < 65536:
            x[j] = (x[j] << 8) | v[pos]; pos += 1
    assert pos == len(v) and x == [65536] * 4
    return bytes(out)


@needs_ref
def test_meta8_roundtrip():
    O, R = oracle(), refdec()
    for a in _blocks():
        s = O.tokenize(a)
        for k in ("op", "meta", "var"):
            if s[k].size >= 24:
                m = O.meta8(s[k])
                assert m[:4] == b"VA\x01\x41"
                assert R.meta8_decode(m, s[k].size) == s[k].tobytes()


def test_pack_substream_choice_order():
    O = oracle()
    # <16 bytes: flat only; commands.go:371-406 keeps the first strictly smaller candidate
    p, ch = O.pack_substream(b"\x01\x02")
    assert p == b"\x09\x01\x02" and ch == 0
    p, ch = O.pack_substream(b"")
    assert p == b"\x09" and ch == 0
    a = np.zeros(4000, np.uint8)
    p, ch = O.pack_substream(a)
    assert ch in (1, 2) and len(p) < 4001
