"""The committed golden fixtures (tests/golden/, written by make_golden.py and
checked there with the reference C99 decoder): the oracle must still produce
them (CPU), and the HIP path through the C ABI must produce them byte for
byte (GPU)."""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(HERE, "manifest.json")))


def load(name):
    m = MANIFEST[name]
    a = np.fromfile(os.path.join(HERE, m["input"]), dtype=np.uint8)
    frames = open(os.path.join(HERE, name + ".frames"), "rb").read()
    assert hashlib.sha256(a.tobytes()).hexdigest() == m["input_sha256"]
    assert hashlib.sha256(frames).hexdigest() == m["frames_sha256"]
    return a, frames, np.array(m["frame_sizes"], np.uint32), m["level"]


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_oracle_reproduces_golden(name):
    from oracle_lib import oracle
    a, frames, sizes, level = load(name)
    got, gsz = oracle().encode_buffer(a, level)
    assert got == frames and np.array_equal(gsz, sizes)


@pytest.mark.gpu
def test_gpu_reproduces_golden():
    from przm_amd import gpu as G
    if not G.IsAvailable():
        pytest.fail("gfx950 device not available (the HIP path must run; there is no fallback)")
    g = G.GPUCompressor(0)
    try:
        for name in sorted(MANIFEST):
            a, frames, sizes, level = load(name)
            got, gsz, _ = g.CompressFrames(a, level)
            assert got == frames and np.array_equal(gsz, sizes), name
        # and all of them in one batched launch
        names = sorted(MANIFEST)
        by_level = {}
        for name in names:
            by_level.setdefault(MANIFEST[name]["level"], []).append(name)
        for level, group in by_level.items():
            res, _ = g.EncodeBatch([load(nm)[0] for nm in group], level)
            for nm, (fr, fs) in zip(group, res):
                assert fr == load(nm)[1] and np.array_equal(fs, load(nm)[2]), nm
    finally:
        g.Close()
