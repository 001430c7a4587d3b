"""Known-answer tests of the tokenizer's decision rules, worked by hand from
gpu/kernel/przm_tokenizer.cuh:115-466 (one tile == the block, ring registers
{1, 4, 8} at the start, :162-164).

Each case gives an input, hand-built match descriptors (position ->
(length, distance); every other position has none) and the six streams the
reference rules emit for them, derived in the comment above the case.  The
expected bytes are written out, not computed, so a slip in the oracle's
rules (lookahead margin, RLE thresholds and caps, the 4-byte validation,
the extension cap, the ring-register choice and update, the literal splits,
the opcode layouts) fails a case.  Test infrastructure only.
"""
import numpy as np


def copy_from(buf: bytearray, d: int, n: int) -> None:
    """Append n bytes, each a copy of the byte d back (overlapping copies allowed)."""
    for _ in range(n):
        buf.append(buf[-d])


def _ring_src():
    b = bytearray(b"abcdefgh")   # 0..7
    copy_from(b, 4, 4)           # 8..11  "efgh"       T1 d=4
    b += b"X"                    # 12
    copy_from(b, 8, 4)           # 13..16 "fghe"       T2 d=8
    b += b"Y"                    # 17
    copy_from(b, 8, 4)           # 18..21 "ghXf"       T3 d=8
    b += b"Z"                    # 22
    copy_from(b, 5, 5)           # 23..27 "ghXfZ"      T4 d=5
    b += b"WV"                   # 28, 29
    copy_from(b, 1, 4)           # 30..33 "VVVV"       T5 d=1
    b += b"U"                    # 34
    copy_from(b, 8, 8)           # 35..42 "ZWVVVVVU"   T6 d=8
    b += b"TS"                   # 43, 44
    assert len(b) == 45
    return bytes(b)


def _rep_src(lits: bytes, d: int, n: int, tail: bytes) -> bytes:
    b = bytearray(lits)
    copy_from(b, d, n)
    return bytes(b + tail)


P16 = b"ABCDEFGHIJKLMNOP"
X14 = b"a" * 12 + b"bc"
X13 = b"a" * 11 + b"bc"

KATS = [
    # Lookahead (:196-216): desc[17] is 7 > 5 + 1, so mp = 17.  Token at 17, d = 16: the 4-byte
    # check passes (src[17..21) == src[1..5)); extension runs to the end: 17 + 15 = 32, len 15.
    # Explicit offset (16 is not a register): 17 pending literals > 3, so 14 go first
    # (F0, LEB 14 = 0E) and 3 stay in the opcode: 0x80 | 3 << 4 | (15 - 4) = 0xBB; offset 0x10.
    dict(name="lookahead_skip", src=P16 + P16, desc={16: (5, 16), 17: (7, 16)},
         cmd="F0 0E BB", op="F0 BB", meta="", var="0E", off="10", lit=(P16 + P16)[:17], ntok=1),
    # Lookahead margin: desc[17] is 6 == 5 + 1, not greater, so the token starts at 16 and
    # extends to 16 bytes.  16 literals: 13 first (F0 0D), 3 in 0x80 | 0x30 | 12 = 0xBC.
    dict(name="lookahead_equal_no_skip", src=P16 + P16, desc={16: (5, 16), 17: (6, 16)},
         cmd="F0 0D BC", op="F0 BC", meta="", var="0D", off="10", lit=P16, ntok=1),
    # Two skips in a row: 6 > 4 + 1 at 17, then 8 > 6 + 1 at 18; 19 has none.  Token at 18,
    # len 14 to the end; 18 literals: 15 first (F0 0F), 3 in 0x80 | 0x30 | 10 = 0xBA.
    dict(name="lookahead_chain", src=P16 + P16, desc={16: (4, 16), 17: (6, 16), 18: (8, 16)},
         cmd="F0 0F BA", op="F0 BA", meta="", var="0F", off="10", lit=(P16 + P16)[:18], ntok=1),
    # RLE (:218-286) at 14: 14 + 12 <= 28, src[14..18) = "aaaa", the run is 12 bytes (src[26]
    # is 'b').  12 >= 12 and 12 >= the descriptor length 12: a run.  14 pending literals go
    # first (F0 0E), then F1 + LEB 12 (0C) and the run byte 'a' to the literals.  The trailing
    # "bc" is F4.
    dict(name="rle_12_equals_len", src=X14 + X14, desc={14: (12, 14)},
         cmd="F0 0E F1 0C F4", op="F0 F1 F4", meta="", var="0E 0C", off="", lit=X14 + b"a" + b"bc", ntok=1),
    # The same run against a 14-byte descriptor: 12 < 14, so it is a match (4-byte check
    # passes, len 14 reaches the end).  14 literals: 11 first (F0 0B), 3 in 0x80|0x30|10 = 0xBA,
    # offset 14 = 0E.
    dict(name="rle_12_below_len", src=X14 + X14, desc={14: (14, 14)},
         cmd="F0 0B BA", op="F0 BA", meta="", var="0B", off="0E", lit=X14, ntok=1),
    # A run of 11 is no run even where its descriptor is 11: the match extends over "bc" to
    # 13 bytes.  13 literals: 10 first (F0 0A), 3 in 0x80 | 0x30 | 9 = 0xB9, offset 13 = 0D.
    dict(name="rle_11_is_match", src=X13 + X13, desc={13: (11, 13)},
         cmd="F0 0A B9", op="F0 B9", meta="", var="0A", off="0D", lit=X13, ntok=1),
    # A run of 13: F3 'Q', F1 + LEB 13 (0D) with 'a', then 4 trailing 'Z' (F6).
    dict(name="rle_13", src=b"Q" + b"a" * 13 + b"ZZZZ", desc={1: (13, 1)},
         cmd="F3 F1 0D F6", op="F3 F1 F6", meta="", var="0D", off="", lit=b"QaZZZZ", ntok=1),
    # The run test comes before the distance check, and the run is capped at 65535: a
    # descriptor at 0 (distance 1 > position 0, an invalid match) over 65536 zero bytes is a
    # run of 65535 (LEB FF FF 03), then one trailing literal (F3).
    dict(name="rle_cap_65535", src=bytes(65536), desc={0: (1024, 1)},
         cmd="F1 FF FF 03 F3", op="F1 F3", meta="", var="FF FF 03", off="", lit=b"\0\0", ntok=1),
    # With 12 bytes from position 1 to the end, 1 + 12 > n = 12, so no run test; the match
    # (d = 1, the register r0) is 11 long: repeat k = 0, len > 7, 1 literal:
    # 0xC0, token 1 << 4 | 7 = 0x17 (to meta).
    dict(name="rle_needs_12_bytes", src=b"a" * 12, desc={1: (11, 1)},
         cmd="C0 17", op="C0", meta="17", var="", off="", lit=b"a", ntok=1),
    # 4-byte validation (:288-292): src[8..12) = "abcX" != src[0..4) = "abcd", so no token;
    # 16 literals: F0 10.
    dict(name="validation_first4_mismatch", src=b"abcdefghabcXefgh", desc={8: (8, 8)},
         cmd="F0 10", op="F0", meta="", var="10", off="", lit=b"abcdefghabcXefgh", ntok=0),
    # Distance beyond the position and a length below 4 are no matches: 8 literals, FA.
    dict(name="validation_dist_beyond_pos", src=b"abcdabcd", desc={4: (4, 5)},
         cmd="FA", op="FA", meta="", var="", off="", lit=b"abcdabcd", ntok=0),
    dict(name="validation_len_below_4", src=b"abcdabcd", desc={4: (3, 4)},
         cmd="FA", op="FA", meta="", var="", off="", lit=b"abcdabcd", ntok=0),
    # d = 4 is register r1: repeat k = 1, len 4, 4 literals: 4 << 4 | 0 << 2 | 1 = 0x41.
    dict(name="repeat_r1_short", src=b"abcdabcd", desc={4: (4, 4)},
         cmd="41", op="41", meta="", var="", off="", lit=b"abcd", ntok=1),
    # Only the first 4 bytes are checked: a descriptor of 8 over "abcdXXXX" vs "abcdefgh" is
    # taken at its length.  d = 8 is r2: k = 2, len 8 > 7: 0xC2, token 8 << 4 | 4 = 0x84.
    # (The descriptor claims more than the bytes hold, so this parse does not rebuild the input.)
    dict(name="descriptor_length_trusted", src=b"abcdefghabcdXXXX", desc={8: (8, 8)},
         cmd="C2 84", op="C2", meta="84", var="", off="", lit=b"abcdefgh", ntok=1, decodes=False),
    # Extension cap 65530 (:296-320): "ab" * 32768, descriptor (1024, 2) at 2.  The match
    # extends to 65530 (2 + 65530 < n).  Explicit (2 is no register), 2 literals, base 15:
    # 0x80 | 2 << 4 | 15 = 0xAF, LEB(65530 - 19 = 65511) = E7 FF 03, offset 02.  Then 4
    # trailing literals "abab": F6.
    dict(name="extension_cap_65530", src=b"ab" * 32768, desc={2: (1024, 2)},
         cmd="AF E7 FF 03 F6", op="AF F6", meta="", var="E7 FF 03", off="02", lit=b"ab" + b"abab", ntok=1),
    # Ring registers (:328-427), start {1, 4, 8}:
    #  T1 @8  d=4 = r1 -> k=1, len 4; 8 literals > 7: F3 'a' first, then 7 << 4 | 0 | 1 = 0x71;
    #         ring {4, 1, 8}
    #  T2 @13 d=8 = r2 -> k=2, 1 literal: 0x12; ring {8, 4, 1}
    #  T3 @18 d=8 = r0 -> k=0, 1 literal: 0x10; ring unchanged
    #  T4 @23 d=5 explicit, len 5, 1 literal: 0x80 | 1 << 4 | 1 = 0x91, offset 05; ring {5, 8, 4}
    #  T5 @30 d=1 explicit (1 left the ring), len 4, 2 literals: 0xA0, offset 01; ring {1, 5, 8}
    #         (a run of only 4 bytes: no RLE)
    #  T6 @35 d=8 = r2 -> k=2, len 8 > 7, 1 literal: 0xC2, token 0x14; ring {8, 1, 5}
    #  trailing "TS": F4
    dict(name="ring_registers", src=_ring_src(),
         desc={8: (4, 4), 13: (4, 8), 18: (4, 8), 23: (5, 5), 30: (4, 1), 35: (8, 8)},
         cmd="F3 71 12 10 91 A0 C2 14 F4", op="F3 71 12 10 91 A0 C2 F4", meta="14", var="", off="05 01",
         lit=b"abcdefghXYZWVUTS", ntok=6),
    # A repeat of r1 swaps r0 and r1: "abcd" then d = 4 (r1, k = 1, 4 literals: 0x41) gives
    # {4, 1, 8}, so the following d = 1 is r1 again (k = 1, 1 literal 'Y': 0x11), not an explicit
    # offset.  The d = 1 match covers "YYYY" after 'Y' (a run of 5 < 12: no RLE); 'Z' trails (F3).
    dict(name="ring_r1_swap", src=_rep_src(_rep_src(b"abcd", 4, 4, b"Y"), 1, 4, b"Z"), desc={4: (4, 4), 9: (4, 1)},
         cmd="41 11 F3", op="41 11 F3", meta="", var="", off="", lit=b"abcdYZ", ntok=2),
    # Literal split of a short repeat: 7 pending literals stay in the opcode (0x71), 8 split.
    dict(name="short_repeat_7_literals", src=_rep_src(b"bcdefgh", 4, 4, b"X"), desc={7: (4, 4)},
         cmd="71 F3", op="71 F3", meta="", var="", off="", lit=b"bcdefghX", ntok=1),
    # Long repeat (len 8, k = 1): 15 literals stay in the token (15 << 4 | 4 = 0xF4), 16 split.
    dict(name="long_repeat_15_literals", src=_rep_src(P16[1:], 4, 8, b"Z"), desc={15: (8, 4)},
         cmd="C1 F4 F3", op="C1 F3", meta="F4", var="", off="", lit=P16[1:] + b"Z", ntok=1),
    dict(name="long_repeat_16_literals", src=_rep_src(P16, 4, 8, b"Z"), desc={16: (8, 4)},
         cmd="F3 C1 F4 F3", op="F3 C1 F3", meta="F4", var="", off="", lit=P16 + b"Z", ntok=1),
    # Explicit offset: 3 literals stay (0x80 | 3 << 4 | 0 = 0xB0), 4 split.  d = 3, len 4.
    dict(name="explicit_3_literals", src=_rep_src(b"abc", 3, 4, b"Z"), desc={3: (4, 3)},
         cmd="B0 F3", op="B0 F3", meta="", var="", off="03", lit=b"abcZ", ntok=1),
    dict(name="explicit_4_literals", src=_rep_src(b"xabc", 3, 4, b"Z"), desc={4: (4, 3)},
         cmd="F3 B0 F3", op="F3 B0 F3", meta="", var="", off="03", lit=b"xabcZ", ntok=1),
    # Repeat length classes with d = 4 (r1), 4 literals: len 7 short (4 << 4 | 3 << 2 | 1 =
    # 0x4D), len 8 long (0xC1, 0x44), len 18 long base 14 (0x4E), len 19 base 15 + LEB 0.
    dict(name="repeat_len_7", src=_rep_src(b"ABCD", 4, 7, b"Z"), desc={4: (7, 4)},
         cmd="4D F3", op="4D F3", meta="", var="", off="", lit=b"ABCDZ", ntok=1),
    dict(name="repeat_len_8", src=_rep_src(b"ABCD", 4, 8, b"Z"), desc={4: (8, 4)},
         cmd="C1 44 F3", op="C1 F3", meta="44", var="", off="", lit=b"ABCDZ", ntok=1),
    dict(name="repeat_len_18", src=_rep_src(b"ABCD", 4, 18, b"Z"), desc={4: (18, 4)},
         cmd="C1 4E F3", op="C1 F3", meta="4E", var="", off="", lit=b"ABCDZ", ntok=1),
    dict(name="repeat_len_19", src=_rep_src(b"ABCD", 4, 19, b"Z"), desc={4: (19, 4)},
         cmd="C1 4F 00 F3", op="C1 F3", meta="4F", var="00", off="", lit=b"ABCDZ", ntok=1),
    # Explicit length classes, d = 3, 3 literals: len 18 -> 0xB0 | 14 = 0xBE; len 19 -> 0xBF,
    # LEB 0; len 200 -> 0xBF, LEB 181 = B5 01.
    dict(name="explicit_len_18", src=_rep_src(b"abc", 3, 18, b"Z"), desc={3: (18, 3)},
         cmd="BE F3", op="BE F3", meta="", var="", off="03", lit=b"abcZ", ntok=1),
    dict(name="explicit_len_19", src=_rep_src(b"abc", 3, 19, b"Z"), desc={3: (19, 3)},
         cmd="BF 00 F3", op="BF F3", meta="", var="00", off="03", lit=b"abcZ", ntok=1),
    dict(name="explicit_len_200", src=_rep_src(b"abc", 3, 200, b"Z"), desc={3: (200, 3)},
         cmd="BF B5 01 F3", op="BF F3", meta="", var="B5 01", off="03", lit=b"abcZ", ntok=1),
    # Literal runs (:57-76): 8 -> FA, 9 -> F0 09, 128 -> F0 80 01.
    dict(name="literals_8", src=b"ABCDEFGH", desc={}, cmd="FA", op="FA", meta="", var="", off="",
         lit=b"ABCDEFGH", ntok=0),
    dict(name="literals_9", src=b"ABCDEFGHI", desc={}, cmd="F0 09", op="F0", meta="", var="09", off="",
         lit=b"ABCDEFGHI", ntok=0),
    dict(name="literals_128", src=bytes(range(128)), desc={}, cmd="F0 80 01", op="F0", meta="", var="80 01",
         off="", lit=bytes(range(128)), ntok=0),
]

STREAMS = ("cmd", "op", "meta", "var", "off")


def desc_array(n, desc):
    d = np.zeros(max(n, 1), np.uint32)
    for p, (length, dist) in desc.items():
        d[p] = (length << 16) | dist
    return d[:n]


def run_kat(O, k):
    """Mismatches of one case against an Oracle-like object (empty list when it passes)."""
    src = np.frombuffer(k["src"], np.uint8).copy()
    got = O.tokenize(src, desc=desc_array(src.size, k["desc"]))
    bad = []
    for s in STREAMS:
        want = bytes.fromhex(k[s])
        if got[s].tobytes() != want:
            bad.append(f"{k['name']} {s}: got {got[s].tobytes().hex(' ')} want {want.hex(' ')}")
    if got["lit"].tobytes() != k["lit"]:
        bad.append(f"{k['name']} lit: got {got['lit'].tobytes()[:40]!r} want {k['lit'][:40]!r}")
    if got["ntok"] != k["ntok"]:
        bad.append(f"{k['name']} ntok: got {got['ntok']} want {k['ntok']}")
    return bad
