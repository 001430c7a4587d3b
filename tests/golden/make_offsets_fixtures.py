"""Writes tests/golden/offsets_ties.json: distance lists on which the comparisons of the
Section-3 prefix option tie, found by seeded searches over the oracle.

  vans_eq_m12        64+ distances whose adaptive V-ANS code stream is exactly as long as
                     the Mode-12 one: PackDistancePrefix replaces Mode 12 only by a strictly
                     shorter V-ANS stream (distance.go:725-731), so Mode 12 is kept.
  prefix_eq_raw      a container exactly as long as the varints: PackageOffsetsSection takes
                     the container only when strictly shorter (distance.go:837), so the
                     section is flat 0x09.
  prefix_lt_raw      a container one byte shorter than the varints: the container is taken.
  code_len_eq_count  a container shorter than the varints whose code stream is exactly as
                     long as the match count, which both reference decoders read as raw
                     codes (distance.go:601-604, src/przm_dec.c:1097-1098): mode G writes
                     such a section flat.
  hazard_block       the same inside a real 64 KiB block (tests/offsets_cases.copy_block):
                     the block's explicit distances and the seed of its filler bytes.
  renorm_block       a block whose Section 3 is a container with a Mode-12 code stream in
                     which the coder's state, shifted down, lands exactly on the bound
                     4096 f - 1 (distance.go:357: a byte is emitted only while x > 4096 f - 1).
                     Decoding catches a coder that emits at the bound too, but only on a
                     stream where the case occurs (about once in 90,000 steps over codes
                     0-31); the search finds one by stepping a model of the coder's state.

The container lengths come from the oracle's coders, which the reference C99 decoder pins;
the tests check the recorded lengths again before relying on a fixture.  Run from the
repository root:
    python tests/golden/make_offsets_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle_lib import oracle  # noqa: E402
from offsets_cases import copy_block, copy_positions, leb  # noqa: E402


def random_distances(rng, n):
    k = rng.integers(0, 3)
    if k == 0:
        return [int(v) for v in rng.integers(1, int(rng.choice([5, 20, 200, 2000, 65536])), n)]
    if k == 1:
        return [int(min(np.exp(rng.random() * np.log(65536)) + 1, 65535)) for _ in range(n)]
    return [int(v) for v in rng.choice(rng.integers(1, 65536, int(rng.integers(1, 6))), n)]


def find_ties(O):
    out = {}
    rng = np.random.default_rng(0)
    while len(out) < 3:
        n = int(rng.integers(1, 200))
        ds = random_distances(rng, n)
        sec, info = O.pack_distance_prefix(ds)
        raw = len(leb(ds))
        if n >= 64 and "vans_eq_m12" not in out:
            codes = np.array([O.dist_code(d)[0] for d in ds], np.uint8)
            m12 = len(O.dist_m12(codes))
            if len(O.vans(codes)) == m12:
                out["vans_eq_m12"] = {"distances": ds, "m12_len": m12}
        if len(sec) == raw and "prefix_eq_raw" not in out:
            out["prefix_eq_raw"] = {"distances": ds, "section_len": 1 + raw}
        if len(sec) == raw - 1 and "prefix_lt_raw" not in out:
            out["prefix_lt_raw"] = {"distances": ds, "section_len": raw - 1}
    return out


def collision_score(O, ds):
    """0 when the container is shorter than the varints and its code stream is as long as the
    match count; grows with the distance from that."""
    sec, info = O.pack_distance_prefix(ds)
    raw = len(leb(ds))
    pen = 0 if len(sec) < raw else 3 * (len(sec) - raw + 1)
    return abs(info[1] - len(ds)) + pen


def ring_ok(ds):
    """No distance equals one of the three last explicit ones (registers start {1, 4, 8}), so
    the tokenizer codes every copy with an explicit offset."""
    ring = [1, 4, 8]
    for d in ds:
        if d in ring:
            return False
        ring = [d, ring[0], ring[1]]
    return True


def climb(O, rng, n, draw, check=lambda ds: True, steps=4000):
    ds = [draw(k) for k in range(n)]
    s = collision_score(O, ds)
    for _ in range(steps):
        k = int(rng.integers(0, n))
        old = ds[k]
        ds[k] = draw(k)
        if not check(ds):
            ds[k] = old
            continue
        s2 = collision_score(O, ds)
        if s2 <= s:
            s = s2
        else:
            ds[k] = old
        if s == 0:
            return ds
    return None


def find_collision(O):
    rng = np.random.default_rng(1)
    pool = [1, 2, 3, 4, 7, 8, 12, 20, 40, 100, 200, 500, 1000, 3000, 9000, 17000, 30000, 40000, 60000, 65535]
    while True:
        n = int(rng.integers(16, 64))
        ds = climb(O, rng, n, lambda k: int(rng.choice(pool)) if rng.random() < 0.7 else int(rng.integers(1, 65536)))
        if ds:
            return {"distances": ds}


def find_block(O):
    rng = np.random.default_rng(5)
    for trial in range(10000):
        n = int(rng.integers(24, 64))
        hi = [p - 600 for p in copy_positions(n)]

        def draw(k):
            if rng.random() < 0.6:
                return int(min(hi[k], np.exp(rng.random() * np.log(hi[k] / 2)) * 2))
            return int(rng.integers(2, hi[k]))
        ds = climb(O, rng, n, draw, ring_ok)
        if not ds:
            continue
        blk = copy_block(ds, trial)
        if O.tokenize(blk)["off"].tobytes() != leb(ds):
            continue
        fr, info = O.encode_block(blk, 6, True)
        if info.kind == 0:
            return {"distances": ds, "seed": trial}
    raise RuntimeError("no block found")


def bound_steps(codes, freq):
    """Steps (symbol indices) at which the Mode-12 coder's state, after its renormalising
    shifts, equals 4096 f - 1: a model of the coder's state used only to pick the input."""
    cdf = np.zeros((4, 64), np.int64)
    cdf[:, 1:] = np.cumsum(freq, axis=1)[:, :-1]
    x, hits = 65536, []
    for i in range(len(codes) - 1, -1, -1):
        p = codes[i - 1] if i else 0
        c = 0 if p < 6 else 1 if p < 14 else 2 if p < 26 else 3
        f, cd = int(freq[c][codes[i]]), int(cdf[c][codes[i]])
        while x > 4096 * f - 1:
            x >>= 8
        if x == 4096 * f - 1:
            hits.append(i)
        x = (x // f) * 4096 + cd + (x % f)
    return hits


def find_renorm_block(O):
    """The bound is met right after a code high in its table (28-31, distances of 16385 and
    more) when the next code coded has frequency 1-3 (codes 1-5 after codes 0-5 or 26+):
    large distances with a few small ones between them (2, 3, 5, 6, 7 in turn, so no small
    distance repeats one of the three registers)."""
    from offsets_cases import dist_freq_table
    freq = dist_freq_table()
    rng = np.random.default_rng(9)
    smalls = [2, 3, 5, 6, 7]
    for trial in range(100000):
        n = int(rng.integers(24, 64))
        hi = [p - 600 for p in copy_positions(n)]
        ds, si = [], int(rng.integers(0, 5))
        for k in range(n):
            if rng.random() < 0.15:
                ds.append(smalls[si % 5])
                si += int(rng.integers(1, 3))
            else:
                ds.append(int(rng.integers(16385, hi[k] + 1)))
        if not ring_ok(ds):
            continue
        steps = bound_steps([O.dist_code(d)[0] for d in ds], freq)
        if not steps:
            continue
        sec, info = O.pack_distance_prefix(ds)
        if len(sec) >= len(leb(ds)) or info[1] == n or info[0] != 0:
            continue
        blk = copy_block(ds, trial)
        if O.tokenize(blk)["off"].tobytes() != leb(ds):
            continue
        fr, fi = O.encode_block(blk, 6, True)
        if fi.kind == 0 and fi.s3_mode in (0x06, 0x0A, 0x0B) and fi.s3_code == 0:
            return {"distances": ds, "seed": trial, "bound_steps": steps}
    raise RuntimeError("no block found")


def main():
    O = oracle()
    out = find_ties(O)
    out["code_len_eq_count"] = find_collision(O)
    out["hazard_block"] = find_block(O)
    out["renorm_block"] = find_renorm_block(O)
    with open(os.path.join(HERE, "offsets_ties.json"), "w") as f:
        json.dump(out, f, indent=1)
    print({k: len(v["distances"]) for k, v in out.items()})


if __name__ == "__main__":
    main()
