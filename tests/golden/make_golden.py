"""Writes the golden fixtures of this directory: seeded inputs and the frames
the oracle (oracle/liboracle.so) encodes them to, each frame checked to decode
back with the reference C99 decoder (oracle/_ref) before it is written.

    python3 tests/golden/make_golden.py

Fixtures: <input>.in (input bytes, shared by the cases that encode it at
several levels), <case>.frames (frames back to back) and manifest.json
(input file, level, frame sizes, SHA-256 of both files).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle_lib import oracle, refdec, have_refdec  # noqa: E402
from przm_amd import datagen as D  # noqa: E402


def cases():
    rng = np.random.default_rng(20261203)
    markov = D.markov(100000, 101)
    yield "empty", "empty", np.zeros(0, np.uint8), 6
    yield "tiny13", "tiny13", np.frombuffer(b"abcabcabcabca", np.uint8).copy(), 6
    for lv in (1, 6, 9, 11):
        yield f"markov_100k_l{lv}", "markov_100k", markov, lv
    yield "mixture_160k", "mixture_160k", D.mixture(163840, 102), 6
    yield "zeros_70000", "zeros_70000", np.zeros(70000, np.uint8), 6
    yield "random_65536", "random_65536", rng.integers(0, 256, 65536, dtype=np.uint8), 6
    yield "lowalpha_70000", "lowalpha_70000", rng.integers(0, 3, 70000, dtype=np.uint8), 6
    yield "text_20k", "text_20k", np.frombuffer((b"The quick brown fox jumps over the lazy dog. " * 500)[:20000],
                                                np.uint8).copy(), 6


def main():
    if not have_refdec():
        raise SystemExit("the reference decoder (oracle/_ref) is needed to check the fixtures: run make -C oracle")
    O, R = oracle(), refdec()
    manifest = {}
    for name, inp, a, level in cases():
        frames, sizes = O.encode_buffer(a, level)
        assert R.decompress_frames(frames, sizes) == a.tobytes(), name
        with open(os.path.join(HERE, inp + ".in"), "wb") as f:
            f.write(a.tobytes())
        with open(os.path.join(HERE, name + ".frames"), "wb") as f:
            f.write(frames)
        manifest[name] = {"input": inp + ".in", "level": level, "input_bytes": int(a.size),
                          "frame_sizes": [int(x) for x in sizes],
                          "input_sha256": hashlib.sha256(a.tobytes()).hexdigest(),
                          "frames_sha256": hashlib.sha256(frames).hexdigest()}
        print(name, a.size, "->", len(frames))
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
