"""Known answers for Section 4 with the sheaf trials (the encoder option that brings
PackageLiteralsSection, core/literals/literals.go:667-693, to the GPU frames), shared by
the oracle tests, the oracle's mutation tests and the GPU tests.

Where the answers come from:
  * the reference C99 decoder compiled from its sources (oracle/_ref): its own
    przm_context_classify / przm_context_classify_order2 (src/przm_dec.c:334-352) for every
    context, and its frame decoder (stage_unpack_sheaf, src/przm_dec.c:805-851) on frames
    that carry sheaf containers;
  * containers worked out by hand from DecomposeFibers (sheaf.go:125-152) and
    assembleSheafContainer (literals.go:645-664) for one short input per mode;
  * seeded ties (tests/golden/sheaf_ties.json, made by tests/golden/make_sheaf_fixtures.py)
    for the strict comparisons.
"""
import ctypes
import json
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TIES_PATH = os.path.join(HERE, "golden", "sheaf_ties.json")


def container(mode, fibers):
    """A sheaf container of raw fibers: 'V','A',1, flags 0x01 | 0x04 | mode << 4, the
    literal count, the four fiber lengths twice (raw: coded length = length), the fibers."""
    lens = [len(f) for f in fibers]
    return (b"VA\x01" + bytes([0x05 | mode << 4]) + struct.pack("<I", sum(lens)) + struct.pack("<4I", *lens)
            + struct.pack("<4I", *lens) + b"".join(fibers))


# one short input per mode, split by hand: (mode, literals, fibers 0..3)
#   mode 0, "a b": 'a' after nothing (class of 0: other) -> 3, ' ' after 'a' (ident) -> 2,
#           'b' after ' ' (white space) -> 0
#   mode 1, '"k":1': '"' -> 3 (after nothing), 'k' after '"' -> 0, '"' after 'k' -> 1,
#           ':' after '"' -> 0, '1' after ':' -> 0
#   mode 2, "a{b": 'a' -> 1, '{' after 'a' -> 1, 'b' after '{' (syntax) -> 0
#   mode 3, "12a.": '1' (p1 = 0: other) -> 1, '2' after '1' with p2 = 0 -> 1, 'a' after
#           '2' with p2 = '1' (digit) -> 0, '.' after 'a' with p2 = '2' (not ident) -> 3
HAND = [
    (0, b"a b", [b"b", b"", b" ", b"a"]),
    (1, b'"k":1', [b"k:1", b'"', b"", b'"']),
    (2, b"a{b", [b"b", b"a{", b"", b""]),
    (3, b"12a.", [b"a", b"12", b"", b"."]),
]


def lits_frame(lits, s4):
    """A frame whose commands are one literal run of len(lits) and whose Section 4 is s4
    (flag 0x80 when s4 is a container)."""
    n = len(lits)
    if n <= 8:
        cmd = bytes([0xF2 + n])
    else:
        v, b = n, bytearray()
        while v >= 0x80:
            b.append((v & 0x7F) | 0x80)
            v >>= 7
        b.append(v)
        cmd = b"\xF0" + bytes(b)
    s2 = b"\x09" + cmd
    flags = 0x80 if s4[:3] == b"VA\x01" else 0
    return (b"PRZM" + bytes([flags]) + n.to_bytes(3, "little") + struct.pack("<HH", len(s2), 0)
            + bytes([0, 16, 0, 0, 0, 0, 0, 0]) + s2 + s4)


def _ref_classifiers(R):
    R.l.przm_context_classify.restype = ctypes.c_uint8
    R.l.przm_context_classify.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
    R.l.przm_context_classify_order2.restype = ctypes.c_uint8
    R.l.przm_context_classify_order2.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
    return R.l.przm_context_classify, R.l.przm_context_classify_order2


def sheaf_failures(O, R=None):
    """Names of the known-answer checks of the sheaf option that oracle O fails (R: the
    reference C99 decoder, or None when it is not built)."""
    bad = []
    if R is not None:
        c1, c2 = _ref_classifiers(R)
        if any(O.sheaf_class(0, p1, m) != c1(p1, m) for m in range(3) for p1 in range(256)):
            bad.append("classify_modes_0_2")
        if any(O.sheaf_class(p2, p1, 3) != c2(p2, p1) for p2 in range(256) for p1 in range(256)):
            bad.append("classify_order2")
    for mode, lits, fibers in HAND:
        if O.pack_sheaf(lits, mode) != container(mode, fibers):
            bad.append(f"hand_mode_{mode}")
    T = json.load(open(TIES_PATH))
    # a fiber whose coded form is exactly as long stays raw
    f = bytes.fromhex(T["fiber_eq_len"]["lits"])
    if O.pack_sheaf(f, 2) != container(2, [b"", f, b"", b""]):
        bad.append("fiber_eq_len")
    # the section choices: raw kept on a V-ANS tie, V-ANS kept on a sheaf tie, the lower mode
    # kept on a tie between modes
    for key, kind in (("vans_eq_raw", 0), ("vans_eq_sheaf", 1), ("sheaf_eq_sheaf", 2)):
        sec, k = O.package_literals(bytes.fromhex(T[key]["lits"]))
        if k != kind:
            bad.append(key)
    # frames of structured content decode through the reference decoder (where sheaf
    # containers win), and a container of each mode decodes inside a frame
    if R is not None:
        from przm_amd import datagen as D
        a = D.mixture(2 << 20, 21)
        frames, sizes = O.encode_buffer(a, 6, False, True)
        try:
            ok = R.decompress_frames(frames, sizes) == a.tobytes()
        except ValueError:
            ok = False
        if not ok:
            bad.append("mixture_frames")
        rng = np.random.default_rng(22)
        text = bytes(rng.choice(list(b'{"name": "value_1", "n": 12.5}, \n'), 3000).tolist())
        for m in range(4):
            try:
                ok = R.decompress(lits_frame(text, O.pack_sheaf(text, m))) == text
            except ValueError:
                ok = False
            if not ok:
                bad.append(f"frame_mode_{m}")
    return bad


_LOW = b"abcdefghijklmnopqrstuvwxyz"
_DIG = b"0123456789"
_UP = b"ABCDEFGHIJKLMNOPQRSTUVWXYZ"
_SYN = b' \t\r\n{}[]()<>:;,."\'`=+-*/%&|^!~?\\#'


def _next_mode1(rng, p2, p1):
    # the next byte's class follows mode 1's classes of the previous one: '"' / ':' -> digit,
    # identifier -> upper case or space, digit -> identifier, anything else -> '"' / ':'
    if p1 in b'":':
        return _DIG[rng.integers(0, 10)]
    if p1 in _LOW or p1 == ord("_"):
        return (_UP + b" ")[rng.integers(0, 27)]
    if p1 in _DIG:
        return _LOW[rng.integers(0, 26)]
    return b'":'[rng.integers(0, 2)]


def _next_mode2(rng, p2, p1):
    # white space / syntax -> content, content -> white space / syntax
    if p1 in _SYN:
        return (_LOW + _DIG + _UP)[rng.integers(0, 62)]
    return b" {}[](),;.:"[rng.integers(0, 11)]


def _next_mode3(rng, p2, p1):
    # the class of the two previous bytes (digit after digit, identifier after identifier, ...)
    if p1 in _DIG:
        return _LOW[rng.integers(0, 26)] if p2 in _DIG else _DIG[rng.integers(0, 10)]
    if p1 in _LOW or p1 == ord("_"):
        return _DIG[rng.integers(0, 10)] if (p2 in _LOW or p2 == ord("_")) else _LOW[rng.integers(0, 26)]
    return _LOW[rng.integers(0, 26)]


def modal_text(mode, n, seed):
    """n bytes whose literal streams the option codes as sheaf containers of `mode` (1, 2 or 3):
    each next byte is drawn from a byte class that the mode's context of the byte(s) before it
    predicts; the mixture content takes mode 0."""
    rng = np.random.default_rng(seed)
    step = {1: _next_mode1, 2: _next_mode2, 3: _next_mode3}[mode]
    out = bytearray(b"a")
    while len(out) < n:
        out.append(step(rng, out[-2] if len(out) > 1 else 0, out[-1]))
    return np.frombuffer(bytes(out[:n]), np.uint8).copy()


def _next_mode0(rng, p2, p1):
    """a byte class after the class of p1 (white space, '"' ':', other syntax, identifier,
    digit), as mode 0's contexts tell them apart"""
    groups = (b" \n", b'":', b"{},.-", b"abcdefghijk", b"0123456789")
    k = 0 if p1 in groups[0] else 1 if p1 in groups[1] else 2 if p1 in groups[2] else \
        3 if p1 in groups[3] else 4 if p1 in groups[4] else 3
    pr = ((0, .1, 0, .8, .1), (.2, 0, 0, .5, .3), (.4, .3, 0, 0, .3), (.15, 0, .15, .7, 0), (.1, 0, .3, 0, .6))[k]
    g = groups[rng.choice(5, p=pr)]
    return g[rng.integers(len(g))]


def unique_text(mode, n, seed):
    """n bytes drawn like modal_text (mode 0: by the class of the byte before) in which no
    4-byte string occurs twice (a draw that would repeat one is drawn again, and after 32
    draws the first byte that does not is taken), so a block has no match of its own: a block
    that repeats its first T bytes has exactly T literals, in contexts that make the sheaf
    containers of `mode` win."""
    rng = np.random.default_rng(seed)
    step = {0: _next_mode0, 1: _next_mode1, 2: _next_mode2, 3: _next_mode3}[mode]
    out = bytearray(b"a")
    seen = set()
    while len(out) < n:
        tail = bytes(out[-3:])
        c = None
        for _ in range(32):
            d = step(rng, out[-2] if len(out) > 1 else 0, out[-1])
            if len(out) < 3 or tail + bytes([d]) not in seen:
                c = d
                break
        if c is None:
            c = next(b for b in range(256) if tail + bytes([b]) not in seen)
        out.append(c)
        if len(out) >= 4:
            seen.add(bytes(out[-4:]))
    return np.frombuffer(bytes(out[:n]), np.uint8).copy()
