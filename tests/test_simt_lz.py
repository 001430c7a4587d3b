"""The LZ kernel (przm_amd/csrc/przm_lz.hip: bucketing, chain search, the speculative tokenizer
and stream emission) on the CPU under the fiber emulation of HIP in tests/simt, one workgroup of
512 threads at a time, with every buffer it writes followed by a page that faults on access.
Under sweep and greedy thread schedules (tests/test_simt_kernels.py) its six token streams of
every block equal the oracle's (orc_tokenize_streams: przm_tokenizer.cuh:115-466 over the
matches of compress.go:681-707), and a copy of the kernel with one synchronisation taken out
gives a wrong result under some schedule.  The GPU parity tests check the same streams on the
device.

The whole batched encode of a buffer also runs emulated, kernel by kernel in the runtime's launch
order (tests/simt/pipeline_emul.cpp: check, LZ, histograms, coder order, rANS, frame sizes, scan,
frame writer), and its frames equal the oracle's (orc_encode_block per 64 KiB block).
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle_lib import oracle, ptr

from przm_amd import datagen as D

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "simt", "lz_emul.cpp")
ORDERS = (0, 1, 2, 3, 7)
NAMES = ["cmd", "off", "lit", "op", "meta", "var"]


def build(out, src=SRC):
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-I", os.path.join(HERE, "simt", "include"),
                    "-o", out, src], check=True)


def load(out):
    lib = ctypes.CDLL(out)
    lib.simt_lz_const.restype = ctypes.c_uint64
    lib.simt_lz_const.argtypes = [ctypes.c_int]
    lib.simt_lz.restype = ctypes.c_int
    lib.simt_lz.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return lib


PIPE = os.path.join(HERE, "simt", "pipeline_emul.cpp")


@pytest.fixture(scope="module")
def pipe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("simt_pipe") / "libpipe.so")
    build(out, PIPE)
    lib = ctypes.CDLL(out)
    lib.simt_frames_bound.restype = ctypes.c_uint64
    lib.simt_frames_bound.argtypes = [ctypes.c_uint64]
    lib.simt_encode.restype = ctypes.c_int
    lib.simt_encode.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    return lib


@pytest.fixture(scope="module")
def lz(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("simt_lz") / "liblz.so")
    build(out)
    return load(out)


def inputs():
    """one buffer of blocks: mixture, text, long runs, a two-byte period, JSON-like text, random
    bytes, and a partial last block"""
    rng = np.random.default_rng(61)
    text = np.frombuffer(b"the quick brown fox jumps over the lazy dog; " * 1500, np.uint8)[:65536]
    json = np.frombuffer(bytes(rng.choice(list(b'{"name": "value_1", "n": 12.5, "ok": true}, \n'), 65536).tolist()),
                         np.uint8)
    runs = np.repeat(rng.integers(0, 4, 700, dtype=np.uint8), rng.integers(1, 200, 700))[:65536]
    runs = np.concatenate([runs, np.zeros(65536 - runs.size, np.uint8)])
    ab = np.frombuffer(b"ab" * 32768, np.uint8)
    return np.concatenate([D.mixture(2 << 16, 62), text, runs, ab, json, rng.integers(0, 256, 65536, dtype=np.uint8),
                           D.markov(40000, 63)])


def pipeline_input():
    """four blocks of inputs(): mixture, JSON-like text (where the section options pay), long
    runs, and the partial Markov block"""
    a = inputs()
    return np.concatenate([a[: 1 << 16], a[5 << 16: 6 << 16], a[3 << 16: 4 << 16], a[7 << 16:]])


def check_lz(lib, a, order, B=8, grid=1, level=6):
    C = [int(lib.simt_lz_const(i)) for i in range(9)]
    O = oracle()
    depth = int(O.l.orc_search_depth(level))
    nb = (a.size + 65535) // 65536
    ws = np.zeros(nb * C[0], np.uint8)
    meta = np.zeros(nb * C[7], np.uint8)
    assert lib.simt_lz(ptr(a), a.size, depth, B, grid, ptr(ws), ptr(meta), order) == 0, "deadlock"
    for b in range(nb):
        want = O.tokenize(a[b * 65536: (b + 1) * 65536], level)
        lens = meta[b * C[7] + C[8]: b * C[7] + C[8] + 24].view(np.uint32)
        W = ws[b * C[0]: (b + 1) * C[0]]
        for k, name in enumerate(NAMES):
            assert np.array_equal(W[C[1 + k]: C[1 + k] + lens[k]], want[name]), (b, name, order, B)


@pytest.mark.parametrize("order", (1, 2, 7))  # the pipeline tests below run orders 0 and 3
def test_lz_under_schedules(lz, order):
    check_lz(lz, inputs(), order)


@pytest.mark.parametrize("order,prefix,sheaf", [(0, False, False), (3, False, False), (2, True, True),
                                                (1, True, False), (0, False, True)])
def test_encode_pipeline_under_schedules(pipe, order, prefix, sheaf):
    a = pipeline_input()
    cap = int(pipe.simt_frames_bound(a.size))
    out = np.zeros(cap, np.uint8)
    fl = np.zeros((a.size + 65535) // 65536, np.uint32)
    assert pipe.simt_encode(ptr(a), a.size, 16, order, ptr(out), cap, ptr(fl), int(prefix), int(sheaf)) == 0
    of, osz = oracle().encode_buffer(a, 6, prefix, sheaf)
    assert np.array_equal(fl, osz), (order, prefix, sheaf)
    assert out[: int(fl.sum())].tobytes() == of, (order, prefix, sheaf)


def test_lz_b2_two_workgroups(lz):
    """the two-candidate search on a persistent grid of two workgroups (run one after the other:
    the first takes every block, the second finds the counter past the end)"""
    check_lz(lz, inputs()[: 3 << 16], 2, B=2, grid=2)


@pytest.mark.parametrize("level,order", [(1, 3), (11, 1)])
def test_lz_other_depths(lz, level, order):
    """chain depth 2 (level 1: fewer candidates than a search step holds, the two-candidate
    search of large buckets) and 128 (level 11: the deep search of long chains) on three blocks"""
    check_lz(lz, inputs()[: 3 << 16], order, level=level)


MUTANTS = [
    ("peers_no_fence_after_or", "    if (valid) atomicOr((unsigned long long *)&pb[d], (unsigned long long)me);\n    lds_fence();\n",
     "    if (valid) atomicOr((unsigned long long *)&pb[d], (unsigned long long)me);\n"),
    ("bucket_no_sync_after_count",
     "                    if (wb + c + u * 64 + lane < end) atomicAdd(&S.sm.tab[w][dgs[u]], 1u);\n            }\n            __syncthreads();\n",
     "                    if (wb + c + u * 64 + lane < end) atomicAdd(&S.sm.tab[w][dgs[u]], 1u);\n            }\n"),
    ("emit_no_sync_before_copy_out", "    __syncthreads();\n    QTS(7);\n", "    QTS(7);\n"),
]
# The barrier after the speculative walks has no mutant: without it a thread can read the exit
# of the segment before its own ahead of that segment's walk, but no input and schedule tried
# gave a wrong stream (the second entry round and the path's consistency loop read the exits
# again after barriers).

RUNNER = """
import sys
sys.path.insert(0, %r)
sys.path.insert(0, %r)
import test_simt_lz as T
lib = T.load(sys.argv[1])
a = T.inputs()[: 2 << 16]
for order in (2, 3, 0):
    try:
        T.check_lz(lib, a, order)
    except AssertionError as e:
        print("caught under order", order, e)
        sys.exit(1)
sys.exit(0)
"""


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_lz_missing_synchronisation_is_caught(tmp_path, name, old, new):
    root = os.path.dirname(HERE)
    csrc = tmp_path / "przm_amd" / "csrc"
    shutil.copytree(os.path.join(root, "przm_amd", "csrc"), csrc)
    text = (csrc / "przm_lz.hip").read_text()
    assert text.count(old) == 1, name
    (csrc / "przm_lz.hip").write_text(text.replace(old, new))
    (tmp_path / "tests" / "simt").mkdir(parents=True)
    shutil.copy(SRC, tmp_path / "tests" / "simt" / "lz_emul.cpp")
    out = str(tmp_path / "libmut.so")
    build(out, str(tmp_path / "tests" / "simt" / "lz_emul.cpp"))
    try:
        r = subprocess.run(["python3", "-c", RUNNER % (HERE, root), out], capture_output=True, text=True, timeout=180)
    except subprocess.TimeoutExpired:
        return  # a thread that never reaches its next synchronisation: caught as well
    assert r.returncode != 0, (name, r.stdout[-500:])
