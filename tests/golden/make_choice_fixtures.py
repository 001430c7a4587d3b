"""Writes tests/golden/choice_ties.json: inputs on which the Section-2 choices
of core/commands/commands.go tie, found by a seeded search.

  m8_eq_vans   a substream whose Meta8 and adaptive V-ANS containers have the
               same length, shorter than flat: packSubstream (commands.go:371-406)
               tries Meta8 first and replaces it only on a strictly shorter
               V-ANS, so Meta8 must be kept.
  m8_eq_flat   Meta8 length == 1 + n (flat), V-ANS not shorter: flat is kept.
  vans_eq_flat V-ANS length == 1 + n, Meta8 not shorter: flat is kept.
  mode8_eq_cmd a command stream whose Mode-8 packing (13 + three substreams)
               is exactly as long as the stream: PackageCommandsSection keeps
               Mode 8 only when nPack < len(cmds) (commands.go:558), so the
               section must be flat 0x09.
  mode8_lt_cmd the same with nPack == len(cmds) - 1: Mode 8 is kept.

The lengths themselves come from the oracle's Meta8 and V-ANS coders, which
the reference C99 decoder pins; the tests check the lengths again before
using a fixture.  Run from the repository root:
    python tests/golden/make_choice_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

from oracle_lib import oracle  # noqa: E402
from przm_amd import datagen as D  # noqa: E402


def main():
    O = oracle()
    out = {}
    rng = np.random.default_rng(0)
    need = {"m8_eq_vans", "m8_eq_flat", "vans_eq_flat"}
    while need:
        n = int(rng.integers(24, 120))
        s = rng.integers(0, int(rng.integers(2, 40)), n).astype(np.uint8)
        if rng.random() < 0.5:
            s = (s * int(rng.integers(1, 7))).astype(np.uint8)
        m8, v, fl = len(O.meta8(s)), len(O.vans(s)), 1 + n
        key = None
        if m8 == v and m8 < fl:
            key = "m8_eq_vans"
        elif m8 == fl and v >= fl:
            key = "m8_eq_flat"
        elif v == fl and m8 >= fl:
            key = "vans_eq_flat"
        if key in need:
            out[key] = {"sub": s.tobytes().hex(), "meta8_len": m8, "vans_len": v}
            need.discard(key)
    rng = np.random.default_rng(1)
    texts = [D.mixture(4 << 20, 3), D.markov(1 << 20, 4)]
    need = {0: "mode8_eq_cmd", -1: "mode8_lt_cmd"}
    while need:
        src = texts[0] if rng.random() < 0.5 else texts[1]
        n = int(rng.integers(200, 3000))
        o = int(rng.integers(0, src.size - n))
        s = O.tokenize(src[o:o + n])
        if s["cmd"].size == 0:
            continue
        npack = 13 + sum(len(O.pack_substream(s[k])[0]) for k in ("op", "meta", "var"))
        d = npack - s["cmd"].size
        if d in need:
            out[need.pop(d)] = {"cmd": s["cmd"].tobytes().hex(), "mode8_len": npack}
    with open(os.path.join(HERE, "choice_ties.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(sorted(out))


if __name__ == "__main__":
    main()
