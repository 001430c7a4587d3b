"""Known answers for Section 3 as a distance-prefix container (the encoder option
that brings PackageOffsetsSection, core/distance/distance.go:825-849, to the GPU
frames), shared by the oracle tests, the oracle's mutation tests and the GPU tests.

Where the answers come from:
  * the reference's own tests: TestBitstreamReaderWriter (distance_test.go:53-89),
    TestPackageOffsetsSectionModes / ...Fallback (distance_test.go:404-495);
  * the reference C99 decoder compiled from its sources (oracle/_ref):
    przm_distance_to_code, przm_vans_decode_dist4 and przm_distance_prefix_decode
    (src/przm_dec.c:942-1137) decode what the oracle writes;
  * the rules of PackDistancePrefix / SelectPrefixMode (distance.go:692-758) at
    their boundaries, on inputs worked out by hand (all distances 1000: one
    2-byte varint and code 19 each; 64 equal codes make a 30-byte V-ANS container);
  * seeded ties (tests/golden/offsets_ties.json, made by
    tests/golden/make_offsets_fixtures.py) for the strict comparisons.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TIES_PATH = os.path.join(HERE, "golden", "offsets_ties.json")


def leb(ds):
    """LEB128 varints of the distances, back to back (the offset stream)."""
    out = bytearray()
    for v in ds:
        v = int(v)
        while v >= 0x80:
            out.append((v & 0x7F) | 0x80)
            v >>= 7
        out.append(v)
    return bytes(out)


def copy_positions(n):
    """Where copy_block puts its n copies: spread over the second half of the block."""
    return [33000 + k * ((65536 - 33100) // n) for k in range(n)]


def copy_block(ds, seed):
    """A 64 KiB block whose explicit match distances are exactly ds: 512 zero bytes (one run
    token, and a first 512 bytes that the high-entropy sample does not flag), uniform random
    bytes from `seed`, and at copy_positions(len(ds))[k] an 8-byte forward copy from ds[k]
    bytes back.  The caller checks the tokenizer's offset stream against leb(ds)."""
    rng = np.random.default_rng(seed)
    b = rng.integers(0, 256, 65536, dtype=np.uint8)
    b[:512] = 0
    for p, d in zip(copy_positions(len(ds)), ds):
        for i in range(8):
            b[p + i] = b[p - d + i]
    return b


def header(section):
    """(mode, orig varint bytes, match count, code stream bytes, header bytes) of a container."""
    m = section[0]
    if m == 0x0B:
        return m, section[1], section[2], section[3], 4
    if m == 0x0A:
        r = lambda i: int.from_bytes(section[i:i + 2], "little")  # noqa: E731
        return m, r(1), r(3), r(5), 7
    if m == 0x06:
        r = lambda i: int.from_bytes(section[i:i + 4], "little")  # noqa: E731
        return m, r(1), r(5), r(9), 13
    raise ValueError(f"not a prefix container: {m:#x}")


def m12_decode_strict(stream, n, freq):
    """Mode-12 decode (distance.go:287-321) that also requires what the encoder guarantees:
    every byte used and the final state back at 65536 (the state the encoder starts from)."""
    assert stream[0] == 0x0C
    cdf = np.zeros((4, 65), np.int64)
    cdf[:, 1:] = np.cumsum(freq, axis=1)
    x = int.from_bytes(stream[1:5], "little")
    pos, prev, out = 5, 0, []
    for _ in range(n):
        ctx = 0 if prev < 6 else 1 if prev < 14 else 2 if prev < 26 else 3
        slot = x & 4095
        s = int(np.searchsorted(cdf[ctx], slot, side="right")) - 1
        out.append(s)
        prev = s
        x = int(freq[ctx][s]) * (x >> 12) + slot - int(cdf[ctx][s])
        while x < 65536:
            assert pos < len(stream), "stream ends early"
            x = (x << 8) | stream[pos]
            pos += 1
    assert x == 65536 and pos == len(stream), (x, pos, len(stream))
    return bytes(out)


def dist_freq_table():
    """The 4 x 64 static table from przm_amd/csrc/przm_dist_freq.h (the header the oracle and
    the kernels compile); the reference decoder's decode of the oracle's streams pins it."""
    import re
    text = open(os.path.join(os.path.dirname(HERE), "przm_amd", "csrc", "przm_dist_freq.h")).read()
    body = text[text.index("= {") + 3:]
    vals = [int(v) for v in re.findall(r"\b\d+\b", re.sub(r"/\*.*?\*/", "", body, flags=re.S))]
    t = np.array(vals[:256], np.int64).reshape(4, 64)
    assert (t.sum(axis=1) == 4096).all()
    return t


# TestBitstreamReaderWriter's five values as extra bits: each distance is its bracket base plus the
# value, in a bracket with that many extra bits; worked by hand, the 38 bits MSB first are
# 101 00110 111100001010 1 10101010101010101 + 2 zero bits = A6 F0 AD 55 54.
BITS_DISTANCES = [17 + 0b101, 65 + 0b00110, 8193 + 0b111100001010, 5 + 0b1, 262145 + 0b10101010101010101]
BITS_EXPECTED = bytes.fromhex("A6F0AD5554")


def offsets_failures(O, R=None):
    """Names of the known-answer checks of the prefix option that oracle O fails (R: the
    reference C99 decoder, or None when it is not built)."""
    bad = []
    # DistanceToCode against the reference decoder's own function
    if R is not None:
        for d in list(range(1, 300)) + [511, 512, 513, 4096, 4097, 65535, 65536, 65537, 1 << 24, (1 << 24) + 1]:
            if O.dist_code(d) != R.distance_to_code(d):
                bad.append("dist_code")
                break
    # the extra-bit writer
    sec, info = O.pack_distance_prefix(BITS_DISTANCES)
    if info[3] != 0x0B or sec[4 + info[1]:] != BITS_EXPECTED:
        bad.append("bit_writer")
    # Mode-12 code streams: the reference decoder reads them back, and the stream is exact
    freq = dist_freq_table()
    rng = np.random.default_rng(7)
    # (the renorm block's codes meet the renormalisation bound 4096 f - 1 exactly, at the
    # recorded step: a coder that emits at the bound desynchronises the decode there)
    T = json.load(open(TIES_PATH))
    rb = T["renorm_block"]
    for codes in (bytes(range(49)), bytes([19] * 70), bytes(rng.integers(0, 49, 500).astype(np.uint8)),
                  bytes([0, 0, 48, 5, 6, 13, 14, 25, 26, 47]), bytes(O.dist_code(d)[0] for d in rb["distances"])):
        st = O.dist_m12(codes)
        try:
            ok = m12_decode_strict(st, len(codes), freq) == codes
        except AssertionError:
            ok = False
        if R is not None and ok:
            try:
                ok = R.dist4_decode(st, len(codes)) == codes
            except ValueError:
                ok = False
        if not ok:
            bad.append("m12_stream")
            break
    # TestPackageOffsetsSectionModes / Fallback (distance_test.go:404-495)
    for ds, mode in (([1000] * 70, 0x0B), ([1000] * 1000, 0x0A), ([1000] * 200, 0x0A), ([1000] * 70000, 0x06),
                     ([(i % 4) + 1 for i in range(10)], 0x09)):
        sec, m, ck = O.package_offsets(leb(ds))
        if m != mode or sec[0] != mode:
            bad.append(f"reference_modes_{len(ds)}")
        elif R is not None and m != 0x09:
            try:
                ok = R.prefix_decode(sec, 5 * len(ds) + 64) == leb(ds)
            except ValueError:
                ok = False
            if not ok:
                bad.append(f"reference_modes_{len(ds)}_decode")
    # the code stream: Mode 12 below 64 matches, adaptive V-ANS from 64 on when shorter
    for n, first in ((63, 0x0C), (64, ord("V"))):
        sec, m, ck = O.package_offsets(leb([1000] * n))
        if m != 0x0B or sec[4] != first:
            bad.append(f"code_kind_{n}")
    # SelectPrefixMode at its boundaries: the varint bytes (255 / 256, 65535 / 65536) and the count
    for ds, mode in (([1000] * 127, 0x0B), ([1000] * 128, 0x0A), ([100] * 255, 0x0B), ([100] * 256, 0x0A),
                     ([1000] * 32767 + [100], 0x0A), ([1000] * 32768, 0x06)):
        sec, m, ck = O.package_offsets(leb(ds))
        if m != mode:
            bad.append(f"select_mode_{len(ds)}_{ds[0]}")
    # the strict comparisons, on seeded ties
    sec, info = O.pack_distance_prefix(T["vans_eq_m12"]["distances"])
    if info[0] != 0 or info[1] != T["vans_eq_m12"]["m12_len"]:
        bad.append("vans_eq_m12")
    for key, mode_flat in (("prefix_eq_raw", True), ("prefix_lt_raw", False)):
        ds = T[key]["distances"]
        sec, m, ck = O.package_offsets(leb(ds))
        if (m == 0x09) != mode_flat or len(sec) != T[key]["section_len"]:
            bad.append(key)
    # a code stream as long as the match count is written flat
    ds = T["code_len_eq_count"]["distances"]
    sec, m, ck = O.package_offsets(leb(ds))
    if m != 0x09 or sec != b"\x09" + leb(ds):
        bad.append("code_len_eq_count")
    hb = T["hazard_block"]
    blk = copy_block(hb["distances"], hb["seed"])
    fr, fi = O.encode_block(blk, 6, True)
    if fi.kind != 0 or fi.s3_mode != 0x09 or fr != O.encode_block(blk, 6, False)[0]:
        bad.append("hazard_block")
    blk = copy_block(rb["distances"], rb["seed"])
    fr, fi = O.encode_block(blk, 6, True)
    try:
        ok = fi.kind == 0 and fi.s3_code == 0 and (R is None or R.decompress(fr) == blk.tobytes())
    except ValueError:  # the decoder refused the frame
        ok = False
    if not ok:
        bad.append("renorm_block")
    return bad
