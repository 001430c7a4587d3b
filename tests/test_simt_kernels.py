"""Two workgroup kernels run on the CPU under a fiber emulation of HIP (tests/simt): every GPU
thread is a fiber, __syncthreads a barrier over the workgroup, and every wave operation (DPP
moves, readlane, ballot, shuffles, the wave barrier) a rendezvous of the wave's lanes.  The
scheduler resumes the threads in ascending, descending or random order between
synchronisations, by sweeps (the waves move about in step) or greedily (a wave runs ahead
until a barrier stops it), so a missing barrier (a thread reading LDS another thread has not
written yet, or overwriting what another has not read) gives a different result under some
order.

  * przm_sheaf_split_kernel (przm_amd/csrc/przm_sheaf.hip): the fibers of every mode
    (DecomposeFibers, core/literals/sheaf.go:125-152, with the context classes of the oracle),
    their lengths, the zero bytes after each fiber up to its 16-byte boundary, and the
    normalised table and matrix of every fiber long enough to be coded (NormalizeFrequencies
    and the joint-nibble matrix, literals.go:78-184, from the oracle), for literal counts at and
    around the kernel's 4,096-literal rounds.
  * przm_hist_kernel (przm_amd/csrc/przm_entropy.hip): the tables and matrix lengths of the
    four streams of a block, against the oracle's.
  * przm_hist_kernel, przm_rans_order_kernel and przm_rans_kernel in sequence (the coder stages
    of the pipeline): every job's coded stream, length (or overflow past its stream's length)
    and final states, against the oracle's V-ANS and Meta8 containers (literals.go:289-341,
    vans_meta.go:104-150), on streams long enough for the coder's turns that skip the liveness
    and cap tests and on an incompressible one that overflows.
  * the sheaf option after the coder (the split, class, place, count and encode kernels of
    przm_amd/csrc/przm_sheaf.hip, in launch order): every fiber's exact coded length or
    overflow (the counting coder, literals.go:289-341 and 622) and the section's choice and
    length (PackageLiteralsSection, literals.go:667-693), against the oracle's.
  * the V-ANS stage entry point's kernels (przm_hist_grid_kernel, przm_normalize_kernel,
    przm_rans1_kernel, przm_vans_assemble_kernel): the histogram and the container of a whole
    buffer, against numpy's bincount and the oracle's adaptive V-ANS container.
  * przm_offs_prep_kernel and przm_offs_code_kernel (przm_amd/csrc/przm_offsets.hip): the
    distance codes, the extra bits (MSB first) and the codes' table of LEB128 offset streams over
    several 1,024-byte rounds, and the codes' Mode-12 and V-ANS streams, lengths and states,
    against the oracle's DistanceToCode (distance.go:171-196), BitWriter (distance.go:217-247),
    Mode-12 coder (distance.go:323-376) and V-ANS coder.

The GPU parity tests check the same kernels bit for bit on the device; this checks their
synchronisation on the CPU, where a schedule can be chosen.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import oracle, ptr
from sheaf_cases import modal_text, unique_text

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "simt", "sheaf_emul.cpp")
ORDERS = (0, 1, 2, 3, 7)  # sweeps ascending / descending, greedy ascending / descending, random sweeps
MIN_CODED = 22


def build(out, src=SRC, include=os.path.join(HERE, "simt", "include")):
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-I", include, "-o", out, src], check=True)


def load(out):
    lib = ctypes.CDLL(out)
    lib.simt_const.restype = ctypes.c_uint64
    lib.simt_const.argtypes = [ctypes.c_int]
    lib.simt_sheaf_split.restype = ctypes.c_int
    lib.simt_sheaf_split.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int]
    lib.simt_hist.restype = ctypes.c_int
    lib.simt_hist.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_int]
    lib.simt_rans.restype = ctypes.c_int
    lib.simt_rans.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                              ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    lib.simt_sheaf_pipeline.restype = ctypes.c_int
    lib.simt_sheaf_pipeline.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p]
    lib.simt_vans_stage.restype = ctypes.c_int64
    lib.simt_vans_stage.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_uint64, ctypes.c_void_p]
    lib.simt_offs_prep.restype = ctypes.c_int
    lib.simt_offs_prep.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_int]
    return lib


@pytest.fixture(scope="module")
def simt(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("simt") / "libsimt.so")
    build(out)
    return load(out)


def class_table():
    """cls[m][p2 * 256 + p1]: the fiber of a literal after (p2, p1) in mode m (the oracle's
    ContextClassify / ContextClassifyOrder2)"""
    O = oracle()
    return np.array([[O.sheaf_class(p2, p1, m) for p2 in range(256) for p1 in range(256)] for m in range(4)],
                    np.uint8)


@pytest.fixture(scope="module")
def classes():
    return class_table()


def _table(counts):
    """the oracle's normalised frequencies, freq | cdf << 16, and matrix of a histogram"""
    O = oracle()
    f = O.normalize(counts).astype(np.uint32)
    cdf = np.concatenate([[0], np.cumsum(f)[:-1]]).astype(np.uint32)
    return f | (cdf << 16), O.matrix(f.astype(np.uint16))


def _lit_blocks():
    rng = np.random.default_rng(31)
    src = {m: unique_text(m, 9000, 900 + m) for m in range(4)}
    out = [np.zeros(0, np.uint8), src[0][:1], src[1][:17], src[2][:21], src[3][:22]]
    out += [src[0][:4095], src[1][:4096], src[2][:4097], src[3][:9000]]
    out += [modal_text(3, 5000, 41), rng.integers(0, 256, 8200, dtype=np.uint8)]
    return out


def _split(lib, blocks, order):
    off = np.zeros(len(blocks) + 1, np.uint32)
    off[1:] = np.cumsum([b.size for b in blocks])
    lits = np.concatenate(blocks + [np.zeros(16, np.uint8)])
    sws_bytes = int(lib.simt_const(0))
    sws = np.full(len(blocks) * sws_bytes, 0xA5, np.uint8)  # the kernel must write every byte it owns
    assert lib.simt_sheaf_split(ptr(lits), ptr(off), len(blocks), ptr(sws), order) == 0
    return sws.reshape(len(blocks), sws_bytes)


@pytest.mark.parametrize("order", ORDERS)
def test_sheaf_split_under_schedules(simt, classes, order):
    check_split(simt, classes, order)


def check_split(lib, classes, order):
    hdr, ftab, fib, cap = (int(lib.simt_const(i)) for i in (1, 2, 3, 4))
    ftsz, off_n, off_l, off_mat, off_fmat = (int(lib.simt_const(i)) for i in (5, 6, 7, 8, 9))
    blocks = _lit_blocks()
    sws = _split(lib, blocks, order)
    for bi, lits in enumerate(blocks):
        S = sws[bi]
        u32 = lambda o: int(S[o: o + 4].view(np.uint32)[0])
        assert u32(hdr + off_n) == lits.size, bi
        p2 = np.concatenate([[0, 0], lits[:-2]]).astype(np.int64) if lits.size else np.zeros(0, np.int64)
        p1 = np.concatenate([[0], lits[:-1]]).astype(np.int64) if lits.size else np.zeros(0, np.int64)
        for m in range(4):
            cl = classes[m][p2[: lits.size] * 256 + p1[: lits.size]]
            pos = fib + m * cap
            for k in range(4):
                want = lits[cl == k]
                r = 4 * m + k
                assert u32(hdr + off_l + 4 * r) == want.size, (bi, m, k)
                got = S[pos: pos + want.size]
                assert np.array_equal(got, want), (bi, m, k, order)
                pad = (-want.size) % 16
                assert not S[pos + want.size: pos + want.size + pad].any(), (bi, m, k, "padding")
                pos += want.size + pad
                mat_len = u32(hdr + off_mat + 4 * r)
                if want.size < MIN_CODED:
                    assert mat_len == 0, (bi, m, k)
                    continue
                fc, mat = _table(np.bincount(want, minlength=256))
                t = ftab + r * ftsz
                assert np.array_equal(S[t: t + 1024].view(np.uint32), fc), (bi, m, k, "table")
                assert mat_len == len(mat), (bi, m, k, "matrix length")
                assert S[t + off_fmat: t + off_fmat + mat_len].tobytes() == mat, (bi, m, k, "matrix")


@pytest.mark.parametrize("order", ORDERS)
def test_hist_under_schedules(simt, order):
    check_hist(simt, order)


def check_hist(lib, order):
    rng = np.random.default_rng(32)
    streams = [unique_text(0, 3000, 33), modal_text(1, 700, 34), rng.integers(0, 40, 15, dtype=np.uint8),
               rng.integers(0, 256, 2500, dtype=np.uint8),
               np.zeros(0, np.uint8), unique_text(2, 64, 35), modal_text(2, 4100, 36), np.full(300, 7, np.uint8)]
    nb = len(streams) // 4
    off = np.zeros(len(streams) + 1, np.uint32)
    off[1:] = np.cumsum([s.size for s in streams])
    data = np.concatenate(streams + [np.zeros(16, np.uint8)])
    ftsz, off_fmat = int(lib.simt_const(5)), int(lib.simt_const(9))
    ftab = np.zeros(nb * 4 * ftsz, np.uint8)
    mat_len = np.zeros(nb * 4, np.uint32)
    assert lib.simt_hist(ptr(data), ptr(off), nb, ptr(ftab), ptr(mat_len), order) == 0
    for j, s in enumerate(streams):
        if s.size < (1 if j % 4 == 0 else 16):  # the kernel leaves short streams alone
            continue
        fc, mat = _table(np.bincount(s, minlength=256))
        t = ftab[j * ftsz: (j + 1) * ftsz]
        assert np.array_equal(t[:1024].view(np.uint32), fc), (j, order)
        assert mat_len[j] == len(mat) and t[off_fmat: off_fmat + len(mat)].tobytes() == mat, (j, order)


@pytest.mark.parametrize("order", ORDERS)
def test_rans_under_schedules(simt, order):
    check_rans(simt, order)


OVERFLOW = 0xFFFFFFFF
# job kinds: lit V-ANS, op Meta8 / V-ANS, meta Meta8 / V-ANS, var Meta8 / V-ANS
JOBS = [(0, False), (1, True), (1, False), (2, True), (2, False), (3, True), (3, False)]


def check_rans(lib, order):
    O = oracle()
    rng = np.random.default_rng(38)
    streams = [unique_text(0, 40000, 39), rng.integers(0, 6, 5000, dtype=np.uint8) * 7,
               rng.integers(0, 256, 3000, dtype=np.uint8), modal_text(2, 1000, 40),
               rng.integers(0, 256, 9000, dtype=np.uint8), rng.integers(0, 256, 23, dtype=np.uint8),
               rng.integers(0, 30, 24, dtype=np.uint8), rng.integers(0, 4, 16, dtype=np.uint8),
               np.zeros(1, np.uint8), np.zeros(0, np.uint8), rng.integers(0, 9, 15, dtype=np.uint8),
               modal_text(1, 20000, 41)]
    nb = len(streams) // 4
    off = np.zeros(len(streams) + 1, np.uint32)
    off[1:] = np.cumsum([x.size for x in streams])
    data = np.concatenate(streams + [np.zeros(16, np.uint8)])
    nj = int(lib.simt_const(17))
    stride = 2 * max(x.size for x in streams) + 64
    out = np.zeros(nb * nj * stride, np.uint8)
    klen = np.zeros(nb * nj, np.uint32)
    states = np.zeros(nb * nj * 4, np.uint32)
    assert lib.simt_rans(ptr(data), ptr(off), nb, order, ptr(out), stride, ptr(klen), ptr(states)) == 0
    for b in range(nb):
        for kind, (k, m8) in enumerate(JOBS):
            src = streams[4 * b + k]
            j = b * nj + kind
            if src.size < (24 if m8 else 16 if k else 1):
                assert klen[j] == OVERFLOW, (b, kind)
                continue
            if m8:
                c = O.meta8(src)
                head = 24
            else:
                c = O.vans(src)
                head = 8 + len(O.matrix(O.normalize(np.bincount(src, minlength=256)))) + 16
            K = len(c) - head
            if K > src.size:
                assert klen[j] == OVERFLOW, (b, kind, order)
                continue
            assert klen[j] == K, (b, kind, order)
            assert states[4 * j: 4 * j + 4].tobytes() == c[head - 16: head], (b, kind, order)
            assert out[j * stride: j * stride + K].tobytes() == c[head:], (b, kind, order)


@pytest.mark.parametrize("order", ORDERS)
def test_sheaf_pipeline_under_schedules(simt, classes, order):
    check_sheaf_pipeline(simt, classes, order)


def check_sheaf_pipeline(lib, classes, order):
    O = oracle()
    rng = np.random.default_rng(42)
    lits = [unique_text(0, 12000, 43), unique_text(1, 6000, 44), unique_text(2, 3000, 45),
            unique_text(3, 9000, 46), rng.integers(0, 256, 5000, dtype=np.uint8), modal_text(2, 40, 47),
            np.zeros(0, np.uint8), rng.integers(0, 3, 700, dtype=np.uint8)]
    streams = []
    for i, l in enumerate(lits):
        streams += [l, rng.integers(0, 50, 30 + 7 * i, dtype=np.uint8), rng.integers(0, 20, 25, dtype=np.uint8),
                    rng.integers(0, 256, 17, dtype=np.uint8)]
    nb = len(lits)
    off = np.zeros(len(streams) + 1, np.uint32)
    off[1:] = np.cumsum([x.size for x in streams])
    data = np.concatenate(streams + [np.zeros(16, np.uint8)])
    sws_bytes, hdr = int(lib.simt_const(0)), int(lib.simt_const(1))
    off_l, off_mat, off_k, off_choice, off_len = (int(lib.simt_const(i)) for i in (7, 8, 18, 19, 20))
    sws = np.zeros(nb * sws_bytes, np.uint8)
    meta = np.zeros(nb * int(lib.simt_const(10)), np.uint8)
    assert lib.simt_sheaf_pipeline(ptr(data), ptr(off), nb, order, ptr(sws), ptr(meta)) == 0
    for b, lit in enumerate(lits):
        H = sws[b * sws_bytes + hdr: b * sws_bytes + hdr + 512]
        u32 = lambda o: int(H[o: o + 4].view(np.uint32)[0])
        p2 = np.concatenate([[0, 0], lit[:-2]]).astype(np.int64)[: lit.size]
        p1 = np.concatenate([[0], lit[:-1]]).astype(np.int64)[: lit.size]
        for m in range(4):
            cl = classes[m][p2 * 256 + p1]
            for k in range(4):
                f = lit[cl == k]
                r = 4 * m + k
                assert u32(off_l + 4 * r) == f.size, (b, m, k)
                want = OVERFLOW
                if f.size >= MIN_CODED:
                    head = len(O.matrix(O.normalize(np.bincount(f, minlength=256)))) + 16
                    assert u32(off_mat + 4 * r) + 16 == head, (b, m, k)
                    K = len(O.vans(f)) - 8 - head
                    if f.size > head and K < f.size - head:
                        want = K
                assert u32(off_k + 4 * r) == want, (b, m, k, order)
        sec, kind = O.package_literals(lit)
        assert (u32(off_choice), u32(off_len)) == (kind, len(sec)), (b, order)


@pytest.mark.parametrize("order", ORDERS)
def test_vans_stage_under_schedules(simt, order):
    check_vans_stage(simt, order)


def check_vans_stage(lib, order):
    O = oracle()
    rng = np.random.default_rng(48)
    for a, grid in ((unique_text(0, 30000, 49), 3), (modal_text(3, 5000, 50), 1), (np.full(100, 9, np.uint8), 2),
                    (rng.integers(0, 256, 3, dtype=np.uint8), 1)):
        cap = 8 + 600 + 16 + 2 * a.size + 64
        out = np.zeros(cap, np.uint8)
        counts = np.zeros(256, np.uint32)
        n = lib.simt_vans_stage(ptr(a), a.size, grid, order, ptr(out), cap, ptr(counts))
        assert n > 0, (a.size, order)
        assert np.array_equal(counts, np.bincount(a, minlength=256)), (a.size, order)
        assert out[:n].tobytes() == O.vans(a), (a.size, order)


def _leb128(ds):
    out = bytearray()
    for d in ds:
        d = int(d)
        while d >= 0x80:
            out.append((d & 0x7F) | 0x80)
            d >>= 7
        out.append(d)
    return np.frombuffer(bytes(out), np.uint8)


@pytest.mark.parametrize("order", ORDERS)
def test_offs_prep_under_schedules(simt, order):
    check_offs(simt, order)


def check_offs(lib, order):
    O = oracle()
    rng = np.random.default_rng(37)
    dists = [np.zeros(0, np.int64), rng.integers(1, 65536, 10), rng.integers(1, 5, 63), rng.integers(1, 65536, 64),
             rng.integers(1, 300, 1000), np.concatenate([rng.integers(1, 5, 700), rng.integers(16000, 65536, 2300)])]
    streams = [_leb128(d) for d in dists]
    off = np.zeros(len(streams) + 1, np.uint32)
    off[1:] = np.cumsum([s.size for s in streams])
    data = np.concatenate(streams + [np.zeros(16, np.uint8)])
    ows_bytes, o_ftab, o_codes, o_extra = (int(lib.simt_const(i)) for i in (12, 13, 14, 15))
    ftsz, off_fmat = int(lib.simt_const(5)), int(lib.simt_const(9))
    ows = np.full(len(streams) * ows_bytes, 0xA5, np.uint8)
    assert lib.simt_offs_prep(ptr(data), ptr(off), len(streams), ptr(ows), order, 1) == 0
    m12_end, v_end = int(lib.simt_const(21)), int(lib.simt_const(22))
    for j, (d, st) in enumerate(zip(dists, streams)):
        B = ows[j * ows_bytes: (j + 1) * ows_bytes]
        H = B[:64].view(np.uint32)  # n, orig, extra, m12_len, m12_state, v_len, v_state[4], mat_len
        cz = [O.dist_code(int(x)) for x in d]
        codes = np.array([c for c, _, _ in cz], np.uint8)
        bits = "".join(format(e, "0%db" % nb) for _, e, nb in cz if nb)
        bits += "0" * ((-len(bits)) % 8)
        extra = bytes(int(bits[i: i + 8], 2) for i in range(0, len(bits), 8))
        assert (H[0], H[1], H[2]) == (d.size, st.size, len(extra)), (j, order)
        assert B[o_codes: o_codes + d.size].tobytes() == codes.tobytes(), (j, order)
        assert B[o_extra: o_extra + len(extra)].tobytes() == extra, (j, order)
        if d.size:  # the Mode-12 stream: [0x0C][u32 state][bytes]
            m12 = O.dist_m12(codes)
            K = len(m12) - 5
            assert (H[3], H[4]) == (K, int.from_bytes(m12[1:5], "little")), (j, order, "m12")
            assert B[m12_end - K: m12_end].tobytes() == m12[5:], (j, order, "m12 bytes")
        if d.size >= 64:
            fc, mat = _table(np.bincount(codes, minlength=256))
            assert np.array_equal(B[o_ftab: o_ftab + 1024].view(np.uint32), fc), (j, order)
            assert H[10] == len(mat) and B[o_ftab + off_fmat: o_ftab + off_fmat + len(mat)].tobytes() == mat, (j, order)
            v = O.vans(codes)
            head = 8 + len(mat) + 16
            K = len(v) - head
            assert H[5] == K and H[6:10].tobytes() == v[head - 16: head], (j, order, "vans")
            assert B[v_end - K: v_end].tobytes() == v[head:], (j, order, "vans bytes")
        else:
            assert H[10] == 0 and H[5] == OVERFLOW, (j, order)


# Mutants: one synchronisation taken out of a copy of the kernel sources each.  Every one must
# make some schedule give a wrong result (or stop: a deadlock or a crash in the emulation, which
# the subprocess reports as a failure).  The split kernel's barrier before the scatter pass has
# no mutant: the first round writes the staging only after its own first barrier, which every
# wave reaches after its normalisations, so that barrier orders the reuse as well.  Nor has the
# coder step's wave barrier after its store branch: the store reads the completed half of the
# 256-byte staging ring while later steps write the other half, and the step's quad exchanges
# keep the lanes within a step of each other, so that barrier only keeps the compiler's order.
# The sheaf counting kernel's barrier after its table loads has none either: its workgroup is one
# wave, whose shuffles for the longest and shortest fiber order the loads before the counting.
MUTANTS = [
    ("split_no_sync_after_scan", "przm_sheaf.hip",
     "    if (lane == 63u) *(uint4 *)part[w] = make_uint4(in4[0], in4[1], in4[2], in4[3]);\n    __syncthreads();\n",
     "    if (lane == 63u) *(uint4 *)part[w] = make_uint4(in4[0], in4[1], in4[2], in4[3]);\n"),
    ("split_no_sync_after_staging", "przm_sheaf.hip",
     "            P[m] += 1ull << sh;\n        }\n    }\n    __syncthreads();\n",
     "            P[m] += 1ull << sh;\n        }\n    }\n"),
    ("split_no_sync_after_copy", "przm_sheaf.hip",
     "        my_dst = F.dst[tid] + 16u * F.nw[tid];\n    }\n    __syncthreads();\n",
     "        my_dst = F.dst[tid] + 16u * F.nw[tid];\n    }\n"),
    ("split_no_sync_after_counting", "przm_sheaf.hip",
     "        else if (i0 < n) sheaf_hist_round<false>(L, n, i0, kind, sigw, hs);\n    }\n    __syncthreads();\n",
     "        else if (i0 < n) sheaf_hist_round<false>(L, n, i0, kind, sigw, hs);\n    }\n"),
    ("offs_no_sync_after_scan_read", "przm_offsets.hip",
     "    __syncthreads(); /* part is reused by the next call */\n", ""),
    ("offs_no_sync_after_rounds", "przm_offsets.hip",
     "        nbits += tot & 0xFFFFFu;\n    }\n    __syncthreads();\n",
     "        nbits += tot & 0xFFFFFu;\n    }\n"),
    ("rans_no_wave_barrier_before_store", "przm_entropy.hip",
     "    const uint32_t K2 = K + tot;\n    __builtin_amdgcn_wave_barrier();\n",
     "    const uint32_t K2 = K + tot;\n"),
    ("hist_no_wave_barrier", "przm_entropy.hip",
     "    for (int i = 0; i < 4; i++) cnt[w][lane * 4 + i] = 0;\n    __builtin_amdgcn_wave_barrier();\n",
     "    for (int i = 0; i < 4; i++) cnt[w][lane * 4 + i] = 0;\n"),
]

RUNNER = """
import sys
sys.path.insert(0, %r)
import test_simt_kernels as T
lib = T.load(sys.argv[1])
cls = T.class_table()
for order in (2, 3, 0, 1, 7):  # the greedy schedules first: they catch the most
    for name, fn in (("split", lambda: T.check_split(lib, cls, order)), ("hist", lambda: T.check_hist(lib, order)),
                     ("offs", lambda: T.check_offs(lib, order)), ("rans", lambda: T.check_rans(lib, order)),
                     ("sheaf", lambda: T.check_sheaf_pipeline(lib, cls, order)),
                     ("vans", lambda: T.check_vans_stage(lib, order))):
        try:
            fn()
        except AssertionError:
            print("caught by", name, "under order", order)
            sys.exit(1)
sys.exit(0)
"""


@pytest.mark.parametrize("name,fname,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_missing_synchronisation_is_caught(tmp_path, name, fname, old, new):
    import shutil
    root = os.path.dirname(HERE)
    csrc = tmp_path / "przm_amd" / "csrc"
    shutil.copytree(os.path.join(root, "przm_amd", "csrc"), csrc)
    text = (csrc / fname).read_text()
    assert text.count(old) == 1, name
    (csrc / fname).write_text(text.replace(old, new))
    (tmp_path / "tests" / "simt").mkdir(parents=True)
    shutil.copy(SRC, tmp_path / "tests" / "simt" / "sheaf_emul.cpp")
    out = str(tmp_path / "libmut.so")
    build(out, str(tmp_path / "tests" / "simt" / "sheaf_emul.cpp"))
    r = subprocess.run(["python3", "-c", RUNNER % HERE, out], capture_output=True, text=True, timeout=600)
    assert r.returncode != 0, (name, r.stdout[-500:])
