"""Known answers for the oracle's parse and packing decisions (CPU only).

The tokenizer cases (tests/tokenizer_kats.py) are worked by hand from
gpu/kernel/przm_tokenizer.cuh; the Section-2 choice cases use inputs on which
two candidate codings tie exactly (tests/golden/choice_ties.json), so the
order and the strictness of the comparisons in commands.go:371-406 and :558
decide the output.  test_oracle_mutations.py checks that each of these rules,
mutated, fails here.
"""
import json
import os

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec
from tokenizer_kats import KATS, run_kat

HERE = os.path.dirname(os.path.abspath(__file__))
TIES = json.load(open(os.path.join(HERE, "golden", "choice_ties.json")))
needs_ref = pytest.mark.skipif(not have_refdec(), reason="reference decoder not built")


@pytest.mark.parametrize("k", KATS, ids=[k["name"] for k in KATS])
def test_tokenizer_kat(k):
    bad = run_kat(oracle(), k)
    assert not bad, "\n".join(bad)


@needs_ref
@pytest.mark.parametrize("k", KATS, ids=[k["name"] for k in KATS])
def test_tokenizer_kat_streams_decode(k):
    """The hand-worked streams are also a valid parse: wrapped in a frame (flat sections) the
    reference decoder rebuilds the input."""
    src = k["src"]
    if not k.get("decodes", True):
        pytest.skip("the case's descriptor is longer than the bytes it covers")
    cmd, off, lit = bytes.fromhex(k["cmd"]), bytes.fromhex(k["off"]), k["lit"]
    s2 = b"\x09" + cmd if cmd else b""
    s3 = b"\x09" + off if off else b""
    if len(s2) > 65535 or len(s3) > 65535 or len(src) >= 1 << 24:
        pytest.skip("directory too large for the short form")
    s4 = lit
    hdr = b"PRZM" + bytes([0]) + len(src).to_bytes(3, "little")
    frame = hdr + len(s2).to_bytes(2, "little") + len(s3).to_bytes(2, "little") + bytes([0, 16, 0, 0, 0, 0, 0, 0])
    frame += s2 + s3 + s4
    if s4[:3] == b"VA\x01" or (len(s4) >= 8 and int.from_bytes(s4[4:8], "little") ==
                               (int.from_bytes(s4[:4], "little") + 1) // 2):
        pytest.skip("raw literal section would be sniffed as a container")
    assert refdec().decompress(frame, len(src)) == src


def _sub(key):
    return np.frombuffer(bytes.fromhex(TIES[key]["sub"]), np.uint8).copy()


def test_choice_meta8_kept_on_tie_with_vans():
    O = oracle()
    s = _sub("m8_eq_vans")
    m8, v = O.meta8(s), O.vans(s)
    assert len(m8) == len(v) < 1 + s.size  # the fixture still ties
    p, ch = O.pack_substream(s)
    assert ch == 1 and p == m8  # Meta8 is tried first; V-ANS replaces it only when strictly shorter


def test_choice_flat_kept_on_tie_with_meta8():
    O = oracle()
    s = _sub("m8_eq_flat")
    assert len(O.meta8(s)) == 1 + s.size and len(O.vans(s)) >= 1 + s.size
    p, ch = O.pack_substream(s)
    assert ch == 0 and p == b"\x09" + s.tobytes()


def test_choice_flat_kept_on_tie_with_vans():
    O = oracle()
    s = _sub("vans_eq_flat")
    assert len(O.vans(s)) == 1 + s.size and len(O.meta8(s)) >= 1 + s.size
    p, ch = O.pack_substream(s)
    assert ch == 0 and p == b"\x09" + s.tobytes()


def test_choice_vans_replaces_longer_meta8():
    """Thirty zero bytes: V-ANS (8 + 6 matrix + 16 states = 30 bytes) is shorter than flat (31)
    and than Meta8 (24 + its stream), so it replaces the Meta8 candidate.  (The size thresholds
    n >= 24 and n >= 16 of commands.go:377, 386 cannot change an output: a Meta8 container is at
    least 24 bytes and a V-ANS one at least 29, never shorter than 1 + n below them.)"""
    O = oracle()
    s = np.zeros(30, np.uint8)
    assert len(O.vans(s)) == 30 and len(O.meta8(s)) > 30
    p, ch = O.pack_substream(s)
    assert ch == 2 and p == O.vans(s)


def _mode8(cmd):
    O = oracle()
    op, meta, var = O.split(cmd)
    subs = [O.pack_substream(x)[0] for x in (op, meta, var)]
    return 13 + sum(len(x) for x in subs), subs, (op, meta, var)


def test_mode8_rejected_when_not_shorter():
    cmd = bytes.fromhex(TIES["mode8_eq_cmd"]["cmd"])
    n8, _, _ = _mode8(cmd)
    assert n8 == len(cmd)  # the fixture's Mode-8 packing is exactly as long as the stream
    sec, ch = oracle().package_cmds(cmd)
    assert sec == b"\x09" + cmd and ch == [-1, -1, -1]


def test_mode8_kept_when_shorter():
    cmd = bytes.fromhex(TIES["mode8_lt_cmd"]["cmd"])
    n8, subs, (op, meta, var) = _mode8(cmd)
    assert n8 == len(cmd) - 1
    sec, ch = oracle().package_cmds(cmd)
    # 0x08, u32 original length, u32 op section length, u32 meta section length (commands.go:437-440)
    want = (b"\x08" + len(cmd).to_bytes(4, "little") + len(subs[0]).to_bytes(4, "little")
            + len(subs[1]).to_bytes(4, "little") + subs[0] + subs[1] + subs[2])
    assert sec == want and len(sec) == n8


@needs_ref
def test_mode8_fixture_decodes():
    """The kept Mode-8 section goes through the reference decoder's Section-2 path inside a frame."""
    O, R = oracle(), refdec()
    cmd = bytes.fromhex(TIES["mode8_lt_cmd"]["cmd"])
    sec, _ = O.package_cmds(cmd)
    assert sec[0] == 0x08
    op, meta, var = R.split(cmd)
    assert sec[13:13 + int.from_bytes(sec[5:9], "little")] == O.pack_substream(op)[0]


def test_hazard_raw_section_becomes_uncompressed():
    """A raw literal section that starts 'V','A',0x01 would be misread by the decoders
    (literals.go:703-708, src/przm_dec.c:854-878): mode G emits an uncompressed frame (kind 3).
    Random bytes from 200 symbols keep the high-entropy guard (more than 225 distinct bytes in
    the first 512) from firing first, and V-ANS of them is not shorter than raw."""
    O = oracle()
    rng = np.random.default_rng(12)
    lit = np.concatenate([np.frombuffer(b"VA\x01", np.uint8), rng.integers(0, 200, 600).astype(np.uint8)])
    a = np.concatenate([lit, np.frombuffer(b"q" * 5000, np.uint8)])
    assert O.l.orc_check_high_entropy(ctypes_ptr(a), a.size) == 0
    f, info = O.encode_block(a)
    assert info.kind == 3 and info.s4v == 0 and info.nlit >= 8
    assert f == b"PRZM\x04" + a.size.to_bytes(3, "little") + a.tobytes()


def hazard_input():
    rng = np.random.default_rng(12)
    lit = np.concatenate([np.frombuffer(b"VA\x01", np.uint8), rng.integers(0, 200, 600).astype(np.uint8)])
    return np.concatenate([lit, np.frombuffer(b"q" * 5000, np.uint8)])


def test_hazard_nibble_sniff():
    """The other misreading: a raw section whose second u32 equals (first u32 + 1) / 2 passes
    the decoder's nibble-plane sniff (literals.go:723-735); mode G emits it uncompressed."""
    O = oracle()
    rng = np.random.default_rng(13)
    w0 = 0x30313233
    head = w0.to_bytes(4, "little") + ((w0 + 1) // 2).to_bytes(4, "little")
    lit = np.concatenate([np.frombuffer(head, np.uint8), rng.integers(0, 200, 600).astype(np.uint8)])
    a = np.concatenate([lit, np.frombuffer(b"q" * 5000, np.uint8)])
    f, info = O.encode_block(a)
    assert info.kind == 3 and f[4] == 0x04


def ctypes_ptr(a):
    import ctypes
    return ctypes.c_void_p(a.ctypes.data)
