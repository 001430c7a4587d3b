"""The oracle's Section-4 sheaf option (PackageLiteralsSection with PackEtaleSheaf,
core/literals/literals.go:601-693, core/literals/sheaf.go:20-152) against the reference
decoder and hand-worked containers, on the CPU.

The known answers live in tests/sheaf_cases.py (shared with the mutation tests, which show
that each rule of the option is pinned by at least one of them)."""
import json

import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec
from sheaf_cases import HAND, TIES_PATH, container, lits_frame, modal_text, sheaf_failures

from przm_amd import datagen as D

needs_ref = pytest.mark.skipif(not have_refdec(), reason="reference decoder not built")


def test_every_known_answer():
    assert sheaf_failures(oracle(), refdec() if have_refdec() else None) == []


@pytest.mark.parametrize("mode,lits,fibers", HAND, ids=[f"mode{m}" for m, _, _ in HAND])
def test_hand_worked_containers(mode, lits, fibers):
    assert oracle().pack_sheaf(lits, mode) == container(mode, fibers)


def test_tie_fixtures_hold():
    O = oracle()
    T = json.load(open(TIES_PATH))
    for key in ("vans_eq_sheaf", "sheaf_eq_sheaf"):
        lits = bytes.fromhex(T[key]["lits"])
        v = len(O.vans(np.frombuffer(lits, np.uint8)))
        sh = [len(O.pack_sheaf(lits, m)) for m in range(4)]
        assert (v, sh) == (T[key]["vans_len"], T[key]["sheaf_lens"])
    v = bytes.fromhex(T["vans_eq_raw"]["lits"])
    assert len(O.vans(np.frombuffer(v, np.uint8))) == len(v)
    f = bytes.fromhex(T["fiber_eq_len"]["lits"])
    assert len(O.vans(np.frombuffer(f, np.uint8))) - 8 == len(f)


@needs_ref
def test_every_mode_decodes_on_many_streams():
    """Containers of every mode, with raw and coded fibers, decode inside a frame."""
    O, R = oracle(), refdec()
    rng = np.random.default_rng(31)
    alphabets = [b'{"name": "value_1", "n": 12.5}, \n', b"the quick brown fox 0123456789.-",
                 bytes(range(256)), b"ab", b"x"]
    for it in range(120):
        al = alphabets[it % len(alphabets)]
        n = int(rng.integers(1, 5000))
        lits = bytes(rng.choice(list(al), n).tolist())
        for m in range(4):
            assert R.decompress(lits_frame(lits, O.pack_sheaf(lits, m))) == lits, (it, m)
        sec, k = O.package_literals(lits)
        assert R.decompress(lits_frame(lits, sec)) == lits


def _inputs():
    rng = np.random.default_rng(32)
    yield "mixture", D.mixture(4 << 20, 33)
    yield "markov", D.markov(1 << 20, 34)
    yield "json", np.frombuffer(bytes(rng.choice(list(b'{"key_a": "val", "n": 12345, "f": -0.5}\n'),
                                                  300000).tolist()), np.uint8).copy()
    yield "random", rng.integers(0, 256, 100000, dtype=np.uint8)
    yield "empty", np.zeros(0, np.uint8)


@needs_ref
def test_sheaf_frames_decode_and_never_grow():
    """Frames with the option: the reference decoder reads them, and each is at most as long as
    the frame without it (the options combine with the offsets option too)."""
    O, R = oracle(), refdec()
    kinds = set()
    for name, a in _inputs():
        for pre in (False, True):
            f1, s1 = O.encode_buffer(a, 6, pre, True)
            f0, s0 = O.encode_buffer(a, 6, pre, False)
            assert R.decompress_frames(f1, s1) == a.tobytes(), (name, pre)
            assert (s1 <= s0).all(), (name, pre)
        for b in range(0, max(1, (a.size + 65535) // 65536)):
            kinds.add(O.encode_block(a[b * 65536:(b + 1) * 65536], 6, False, True)[1].s4_kind)
    assert {0, 1} <= kinds and kinds & {2, 3, 4, 5}, kinds


def test_modal_inputs_take_their_mode():
    """The inputs the GPU tests use to reach sheaf modes 1-3 on the device (the mixture reaches 0)."""
    O = oracle()
    for mode in (1, 2, 3):
        a = modal_text(mode, 200000, 40 + mode)
        kinds = {O.encode_block(a[b * 65536:(b + 1) * 65536], 6, False, True)[1].s4_kind for b in range(4)}
        assert kinds == {2 + mode}, (mode, kinds)
    a = D.mixture(4 << 20, 44)
    kinds = {O.encode_block(a[b * 65536:(b + 1) * 65536], 6, False, True)[1].s4_kind for b in range(64)}
    assert 2 in kinds, kinds
