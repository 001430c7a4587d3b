"""ctypes access to the checkers: the C oracle (oracle/liboracle.so) and the
reference C99 decoder compiled from the reference sources (oracle/_ref/).
Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REFDEC = os.path.join(ROOT, "oracle", "_ref", "libprzm_ref_dec.so")
REFTABLES = os.path.join(ROOT, "oracle", "_ref", "libprzm_ref_tables.so")
P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64


class FrameInfo(ctypes.Structure):
    _fields_ = [("kind", U32), ("ncmd", U32), ("noff", U32), ("nlit", U32), ("nop", U32), ("nmeta", U32),
                ("nvar", U32), ("ntok", U32), ("s2", U32), ("s3", U32), ("s4", U32), ("sub", ctypes.c_int32 * 3),
                ("s4v", U32), ("s3_mode", U32), ("s3_code", ctypes.c_int32), ("s4_kind", U32)]


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def u8(x):
    if isinstance(x, np.ndarray):
        return np.ascontiguousarray(x, dtype=np.uint8)
    return np.frombuffer(bytes(x), dtype=np.uint8).copy()


class Oracle:
    def __init__(self, path=ORACLE):
        l = ctypes.CDLL(path)
        l.orc_hash4.restype = U32
        l.orc_hash4.argtypes = [P]
        l.orc_search_depth.restype = ctypes.c_int
        l.orc_search_depth.argtypes = [ctypes.c_int]
        l.orc_match_block.argtypes = [P, U32, ctypes.c_int, P, P]
        l.orc_match_block.restype = None
        l.orc_tokenize_streams.argtypes = [P, U32, ctypes.c_int, P, P, P, P, P, P, P]
        l.orc_tokenize_streams.restype = None
        l.orc_tokenize_desc.argtypes = [P, U32, P, P, P, P, P, P, P, P]
        l.orc_tokenize_desc.restype = None
        l.orc_split_commands.argtypes = [P, U32, P, P, P, P, P, P]
        l.orc_split_commands.restype = ctypes.c_int
        l.orc_normalize.argtypes = [P, P]
        l.orc_normalize.restype = None
        l.orc_joint_nibble.argtypes = [P, P]
        l.orc_joint_nibble.restype = U32
        l.orc_vans4x_encode.argtypes = [P, U32, P, P]
        l.orc_vans4x_encode.restype = U32
        l.orc_vans4x_adaptive.argtypes = [P, U32, P]
        l.orc_vans4x_adaptive.restype = U32
        l.orc_meta8_encode.argtypes = [P, U32, P]
        l.orc_meta8_encode.restype = U32
        l.orc_pack_substream.argtypes = [P, U32, P, ctypes.POINTER(ctypes.c_int)]
        l.orc_pack_substream.restype = U32
        l.orc_check_high_entropy.argtypes = [P, U32]
        l.orc_check_high_entropy.restype = ctypes.c_int
        l.orc_window_log.argtypes = [U32]
        l.orc_window_log.restype = U32
        l.orc_encode_block.argtypes = [P, U32, ctypes.c_int, P, ctypes.POINTER(FrameInfo)]
        l.orc_encode_block.restype = U32
        l.orc_encode_buffer.argtypes = [P, U64, ctypes.c_int, P, U64, P, U64, ctypes.POINTER(U64)]
        l.orc_encode_buffer.restype = U64
        l.orc_package_cmd_stream.argtypes = [P, U32, P, P]
        l.orc_package_cmd_stream.restype = ctypes.c_int64
        l.orc_encode_block_opts.argtypes = [P, U32, ctypes.c_int, ctypes.c_uint, P, ctypes.POINTER(FrameInfo)]
        l.orc_encode_block_opts.restype = U32
        l.orc_encode_buffer_opts.argtypes = [P, U64, ctypes.c_int, ctypes.c_uint, P, U64, P, U64, ctypes.POINTER(U64)]
        l.orc_encode_buffer_opts.restype = U64
        l.orc_sheaf_class.argtypes = [ctypes.c_uint8, ctypes.c_uint8, ctypes.c_int]
        l.orc_sheaf_class.restype = ctypes.c_int
        l.orc_pack_sheaf.argtypes = [P, U32, ctypes.c_int, P]
        l.orc_pack_sheaf.restype = U32
        l.orc_package_literals.argtypes = [P, U32, P, ctypes.POINTER(ctypes.c_int)]
        l.orc_package_literals.restype = U32
        l.orc_dist_code.argtypes = [U32, P, P, P]
        l.orc_dist_code.restype = None
        l.orc_dist_m12_encode.argtypes = [P, U32, P]
        l.orc_dist_m12_encode.restype = U32
        l.orc_pack_distance_prefix.argtypes = [P, U32, P, P]
        l.orc_pack_distance_prefix.restype = U32
        l.orc_package_offsets.argtypes = [P, U32, P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        l.orc_package_offsets.restype = U32
        self.l = l

    def hash4(self, b4):
        a = u8(b4)
        return self.l.orc_hash4(ptr(a))

    def match_block(self, src, level=6, depth=None):
        a = u8(src)
        desc = np.zeros(max(a.size, 1), np.uint32)
        st = np.zeros(1, np.uint64)
        d = depth if depth is not None else self.l.orc_search_depth(level)
        self.l.orc_match_block(ptr(a), a.size, d, ptr(desc), ptr(st))
        return desc[: a.size]

    def tokenize(self, src, level=6, desc=None):
        a = u8(src)
        cap = 2 * a.size + 64
        bufs = [np.zeros(cap, np.uint8) for _ in range(6)]
        lens = np.zeros(7, np.uint32)
        if desc is None:
            self.l.orc_tokenize_streams(ptr(a), a.size, level, *[ptr(b) for b in bufs], ptr(lens))
        else:
            d = np.ascontiguousarray(desc, np.uint32)
            self.l.orc_tokenize_desc(ptr(a), a.size, ptr(d), *[ptr(b) for b in bufs], ptr(lens))
        names = ["cmd", "off", "lit", "op", "meta", "var"]
        out = {n: bufs[i][: lens[i]].copy() for i, n in enumerate(names)}
        out["ntok"] = int(lens[6])
        return out

    def distance_to_code(self, d):
        c, e, nb = ctypes.c_uint8(), ctypes.c_uint32(), ctypes.c_uint8()
        r = self.l.przm_distance_to_code(d, ctypes.byref(c), ctypes.byref(e), ctypes.byref(nb))
        if r != 0:
            raise ValueError(f"distance_to_code error {r}")
        return c.value, e.value, nb.value

    def dist4_decode(self, stream, n):
        s = u8(stream)
        out = np.zeros(max(n, 1), np.uint8)
        r = self.l.przm_vans_decode_dist4(ptr(s), s.size, ptr(out), out.size, n)
        if r != 0:
            raise ValueError(f"dist4 decode error {r}")
        return out[:n].tobytes()

    def prefix_decode(self, container, cap=1 << 20):
        """A distance-prefix container (0x06 / 0x0A / 0x0B) back to its LEB128 varints."""
        c = u8(container)
        out = np.zeros(cap, np.uint8)
        ln = ctypes.c_size_t()
        r = self.l.przm_distance_prefix_decode(ptr(c), c.size, ptr(out), out.size, ctypes.byref(ln))
        if r != 0:
            raise ValueError(f"prefix decode error {r}")
        return out[: ln.value].tobytes()

    def split(self, cmd):
        c = u8(cmd)
        n = max(c.size, 1)
        op, meta, var = (np.zeros(n, np.uint8) for _ in range(3))
        lo, lm, lv = U32(), U32(), U32()
        rc = self.l.orc_split_commands(ptr(c), c.size, ptr(op), ctypes.byref(lo), ptr(meta), ctypes.byref(lm),
                                       ptr(var), ctypes.byref(lv))
        if rc:
            raise ValueError("split failed")
        return op[: lo.value], meta[: lm.value], var[: lv.value]

    def normalize(self, counts):
        c = np.ascontiguousarray(counts, np.uint32)
        f = np.zeros(256, np.uint16)
        self.l.orc_normalize(ptr(c), ptr(f))
        return f

    def matrix(self, freq):
        f = np.ascontiguousarray(freq, np.uint16)
        out = np.zeros(600, np.uint8)
        n = self.l.orc_joint_nibble(ptr(f), ptr(out))
        return out[:n].tobytes()

    def vans(self, src, freq=None):
        a = u8(src)
        out = np.zeros(8 + 600 + 16 + 2 * a.size + 64, np.uint8)
        if freq is None:
            n = self.l.orc_vans4x_adaptive(ptr(a), a.size, ptr(out))
        else:
            f = np.ascontiguousarray(freq, np.uint16)
            n = self.l.orc_vans4x_encode(ptr(a), a.size, ptr(f), ptr(out))
        return out[:n].tobytes()

    def meta8(self, src):
        a = u8(src)
        out = np.zeros(24 + 2 * a.size + 64, np.uint8)
        n = self.l.orc_meta8_encode(ptr(a), a.size, ptr(out))
        return out[:n].tobytes()

    def pack_substream(self, src):
        a = u8(src)
        out = np.zeros(640 + 2 * a.size, np.uint8)
        ch = ctypes.c_int()
        n = self.l.orc_pack_substream(ptr(a), a.size, ptr(out), ctypes.byref(ch))
        return out[:n].tobytes(), ch.value

    def package_cmds(self, cmd):
        """Section 2 of a command stream (split, then Mode 8 or flat) and the substream choices."""
        c = u8(cmd)
        out = np.zeros(16 + 2 * c.size + 3 * 640, np.uint8)
        ch = (ctypes.c_int * 3)(-2, -2, -2)
        n = self.l.orc_package_cmd_stream(ptr(c), c.size, ptr(out), ch)
        if n < 0:
            raise ValueError("command stream does not split")
        return out[:n].tobytes(), list(ch)

    def encode_block(self, src, level=6, s3_prefix=False, sheaf=False):
        """One frame; s3_prefix: Section 3 as PackageOffsetsSection chooses it; sheaf: Section 4
        as PackageLiteralsSection chooses it."""
        a = u8(src)
        out = np.zeros(a.size + 64, np.uint8)
        info = FrameInfo()
        n = self.l.orc_encode_block_opts(ptr(a), a.size, level, int(s3_prefix) | 2 * int(sheaf), ptr(out),
                                         ctypes.byref(info))
        return out[:n].tobytes(), info

    def encode_buffer(self, src, level=6, s3_prefix=False, sheaf=False):
        a = u8(src)
        nb = max((a.size + 65535) // 65536, 1)
        out = np.zeros(a.size + 8 * nb + 64, np.uint8)
        sizes = np.zeros(nb, np.uint32)
        nf = U64()
        tot = self.l.orc_encode_buffer_opts(ptr(a), a.size, level, int(s3_prefix) | 2 * int(sheaf), ptr(out),
                                            out.size, ptr(sizes), nb, ctypes.byref(nf))
        return out[:tot].tobytes(), sizes[: nf.value].copy()

    def sheaf_class(self, p2, p1, mode):
        return self.l.orc_sheaf_class(p2, p1, mode)

    def pack_sheaf(self, lits, mode):
        """PackEtaleSheaf's container of lits in one mode."""
        a = u8(lits)
        out = np.zeros(40 + 2 * a.size + 2400, np.uint8)
        n = self.l.orc_pack_sheaf(ptr(a), a.size, mode, ptr(out))
        return out[:n].tobytes()

    def package_literals(self, lits):
        """PackageLiteralsSection: (section, kind: 0 raw, 1 V-ANS, 2 + mode sheaf)."""
        a = u8(lits)
        out = np.zeros(40 + 2 * a.size + 2400, np.uint8)
        k = ctypes.c_int()
        n = self.l.orc_package_literals(ptr(a), a.size, ptr(out), ctypes.byref(k))
        return out[:n].tobytes(), k.value

    def dist_code(self, d):
        """(code, extra, extra bit count) of distance d (DistanceToCode)."""
        c, e, nb = ctypes.c_uint8(), ctypes.c_uint32(), ctypes.c_uint8()
        self.l.orc_dist_code(d, ctypes.byref(c), ctypes.byref(e), ctypes.byref(nb))
        return c.value, e.value, nb.value

    def dist_m12(self, codes):
        """The Mode-12 code stream [0x0C][u32 state][bytes] of distance codes."""
        c = u8(codes)
        out = np.zeros(2 * c.size + 16, np.uint8)
        n = self.l.orc_dist_m12_encode(ptr(c), c.size, ptr(out))
        return out[:n].tobytes()

    def pack_distance_prefix(self, dists):
        """PackDistancePrefix(auto mode, coded codes): (container, [code kind, code len, extra len, mode])."""
        d = np.ascontiguousarray(dists, dtype=np.uint32)
        out = np.zeros(13 + 5 * d.size + 700, np.uint8)
        info = np.zeros(4, np.uint32)
        n = self.l.orc_pack_distance_prefix(ptr(d), d.size, ptr(out), ptr(info))
        return out[:n].tobytes(), [int(x) for x in info]

    def package_offsets(self, off):
        """Section 3 of a LEB128 offset stream (PackageOffsetsSection): (section, mode byte, code kind)."""
        o = u8(off)
        out = np.zeros(16 + 3 * o.size + 700, np.uint8)
        mode, ck = ctypes.c_int(), ctypes.c_int()
        n = self.l.orc_package_offsets(ptr(o), o.size, ptr(out), ctypes.byref(mode), ctypes.byref(ck))
        return out[:n].tobytes(), mode.value, ck.value


class RefDecoder:
    """The reference C99 decoder (src/przm_dec.c), built from its own sources."""

    def __init__(self):
        l = ctypes.CDLL(REFDEC)
        l.przm_decompress.restype = ctypes.c_int32
        l.przm_decompress.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t]
        l.przm_vans_decode_literals.restype = ctypes.c_int32
        l.przm_vans_decode_literals.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t]
        l.przm_vans_decode_meta8.restype = ctypes.c_int32
        l.przm_vans_decode_meta8.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, ctypes.c_size_t]
        l.przm_distance_to_code.restype = ctypes.c_int32
        l.przm_distance_to_code.argtypes = [U32, P, P, P]
        l.przm_vans_decode_dist4.restype = ctypes.c_int32
        l.przm_vans_decode_dist4.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, ctypes.c_size_t]
        l.przm_distance_prefix_decode.restype = ctypes.c_int32
        l.przm_distance_prefix_decode.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        l.przm_split_commands.restype = ctypes.c_int32
        l.przm_split_commands.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                          P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), P, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)]
        l.przm_get_uncompressed_size.restype = ctypes.c_int32
        l.przm_get_uncompressed_size.argtypes = [P, ctypes.c_size_t]
        self.l = l

    def uncompressed_size(self, data):
        """przm_get_uncompressed_size (src/przm_dec.c:1554-1600): for a PRZS stream it reads the
        u32 at byte 5 (:1564-1567)."""
        f = u8(data)
        return self.l.przm_get_uncompressed_size(ptr(f), f.size)

    def decompress(self, frame, expect_len=None):
        f = u8(frame)
        cap = (expect_len if expect_len is not None else 65536) + 64
        out = np.zeros(cap, np.uint8)
        r = self.l.przm_decompress(ptr(f), f.size, ptr(out), out.size)
        if r < 0:
            raise ValueError(f"decoder error {r}")
        return out[:r].tobytes()

    def decompress_frames(self, frames, sizes, expect_total=None):
        f = u8(frames)
        parts = []
        pos = 0
        for s in sizes:
            parts.append(self.decompress(f[pos: pos + int(s)]))
            pos += int(s)
        assert pos == f.size, (pos, f.size)
        return b"".join(parts)

    def vans_decode(self, container, n):
        c = u8(container)
        out = np.zeros(max(n, 1), np.uint8)
        r = self.l.przm_vans_decode_literals(ptr(c[8:]), c.size - 8, ptr(out), n)
        if r != 0:
            raise ValueError(f"vans decode error {r}")
        return out[:n].tobytes()

    def meta8_decode(self, container, n):
        c = u8(container)
        out = np.zeros(max(n, 1), np.uint8)
        r = self.l.przm_vans_decode_meta8(ptr(c), c.size, ptr(out), out.size, n)
        if r != 0:
            raise ValueError(f"meta8 decode error {r}")
        return out[:n].tobytes()

    def distance_to_code(self, d):
        c, e, nb = ctypes.c_uint8(), ctypes.c_uint32(), ctypes.c_uint8()
        r = self.l.przm_distance_to_code(d, ctypes.byref(c), ctypes.byref(e), ctypes.byref(nb))
        if r != 0:
            raise ValueError(f"distance_to_code error {r}")
        return c.value, e.value, nb.value

    def dist4_decode(self, stream, n):
        s = u8(stream)
        out = np.zeros(max(n, 1), np.uint8)
        r = self.l.przm_vans_decode_dist4(ptr(s), s.size, ptr(out), out.size, n)
        if r != 0:
            raise ValueError(f"dist4 decode error {r}")
        return out[:n].tobytes()

    def prefix_decode(self, container, cap=1 << 20):
        """A distance-prefix container (0x06 / 0x0A / 0x0B) back to its LEB128 varints."""
        c = u8(container)
        out = np.zeros(cap, np.uint8)
        ln = ctypes.c_size_t()
        r = self.l.przm_distance_prefix_decode(ptr(c), c.size, ptr(out), out.size, ctypes.byref(ln))
        if r != 0:
            raise ValueError(f"prefix decode error {r}")
        return out[: ln.value].tobytes()

    def split(self, cmd):
        c = u8(cmd)
        n = max(c.size, 1)
        bufs = [np.zeros(n, np.uint8) for _ in range(3)]
        lens = [ctypes.c_size_t() for _ in range(3)]
        r = self.l.przm_split_commands(ptr(c), c.size, ptr(bufs[0]), n, ctypes.byref(lens[0]), ptr(bufs[1]), n,
                                       ctypes.byref(lens[1]), ptr(bufs[2]), n, ctypes.byref(lens[2]))
        if r != 0:
            raise ValueError(f"split error {r}")
        return tuple(b[: l.value] for b, l in zip(bufs, lens))


def stream_of(frames, sizes, total):
    """The PRZS stream of a buffer's frames: the 17-byte header of ParseStreamHeader
    (decompress.go:849-866: 'P','R','Z','S', flags, u32 chunk count, u64 total uncompressed size),
    then every frame behind its u32 length (the chunk walk of cmd/przm-viz/analyzer.go:242-256).
    A buffer of one frame stays that frame, as the reference writes a single frame for input that
    fits one chunk (tests/stress/stress_test.go:1934-1940, tests/e2e/gpu_tier3_interactions_test.go:62)."""
    sizes = [int(x) for x in sizes]
    if len(sizes) <= 1:
        return bytes(frames)
    out = [b"PRZS", b"\x00", len(sizes).to_bytes(4, "little"), int(total).to_bytes(8, "little")]
    pos = 0
    for n in sizes:
        out.append(n.to_bytes(4, "little"))
        out.append(bytes(frames[pos: pos + n]))
        pos += n
    assert pos == len(frames)
    return b"".join(out)


class RefInfo(ctypes.Structure):  # orc_ref_info (oracle/przm_refcpu.c)
    _fields_ = [(k, U32) for k in ("stride", "bmin", "wlog", "uncompressed", "kind", "ncmd", "noff", "nlit",
                                   "s2", "s3", "s4")] + [("edges", U32 * 7), ("dict", U32)]


class RefCPU:
    """The C restatement of the reference Go CPU compressor (compressInternal at a level,
    oracle/przm_refcpu.c), with the static dictionary when its tables were built into
    oracle/_ref/ (from the reference decoder's own table header)."""

    def __init__(self, lib):
        l = lib
        I, Pi = ctypes.c_int, ctypes.POINTER(ctypes.c_int)
        l.orc_ref_compress_ex.argtypes = [P, U64, I, P, ctypes.POINTER(RefInfo), P, U64]
        l.orc_ref_compress_ex.restype = U64
        l.orc_ref_load_dict.argtypes = [ctypes.c_char_p]
        l.orc_ref_load_dict.restype = I
        l.orc_ref_dict_entries.restype = I
        l.orc_ref_sublengths.argtypes = [I, I, Pi]
        l.orc_ref_sublengths.restype = I
        l.orc_ref_lit_cost.argtypes = [I, I]
        l.orc_ref_lit_cost.restype = U32
        l.orc_ref_dist_cost.argtypes = [I]
        l.orc_ref_dist_cost.restype = U32
        l.orc_ref_cost_explicit.argtypes = [I, U32, I, I]
        l.orc_ref_cost_explicit.restype = U32
        l.orc_ref_cost_repeat.argtypes = [I, I, I, I]
        l.orc_ref_cost_repeat.restype = U32
        l.orc_ref_detect_stride.argtypes = [P, I]
        l.orc_ref_detect_stride.restype = U32
        l.orc_ref_hash.argtypes = [I, P]
        l.orc_ref_hash.restype = U32
        l.orc_ref_find_dict.argtypes = [P, I, P, P, P, I]
        l.orc_ref_find_dict.restype = I
        self.l = l
        self.dict = os.path.exists(REFTABLES) and l.orc_ref_load_dict(REFTABLES.encode()) == 0

    def compress(self, src, level=6, with_cmds=False):
        """one frame of src (one compressInternal call); with_cmds: also the command stream
        before Section 2 packs it"""
        a = u8(src)
        out = np.zeros(a.size + 64, np.uint8)
        info = RefInfo()
        cmds = np.zeros(4 * a.size + 64, np.uint8) if with_cmds else None
        n = self.l.orc_ref_compress_ex(ptr(a), a.size, level, ptr(out), ctypes.byref(info),
                                       ptr(cmds) if with_cmds else None, cmds.size if with_cmds else 0)
        if with_cmds:
            return out[:n].tobytes(), info, cmds[: info.ncmd].tobytes()
        return out[:n].tobytes(), info

    def compress_frames(self, src, level=6, threads=1, block=65536):
        """src as independent frames of `block` bytes (frames, sizes), on `threads` threads"""
        a = u8(src)
        parts = [a[o: o + block] for o in range(0, max(a.size, 1), block)] if a.size else [a]
        if threads > 1:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(threads) as ex:
                res = list(ex.map(lambda b: self.compress(b, level)[0], parts))
        else:
            res = [self.compress(b, level)[0] for b in parts]
        return b"".join(res), np.array([len(f) for f in res], np.uint32)

    def sublengths(self, L, bmin):
        out = (ctypes.c_int * 32)()
        n = self.l.orc_ref_sublengths(L, bmin, out)
        return list(out[:n])

    def find_dict(self, s):
        a = u8(s)
        cap = 256
        wid, tau, ln = np.zeros(cap, np.uint32), np.zeros(cap, np.uint8), np.zeros(cap, np.int32)
        n = self.l.orc_ref_find_dict(ptr(a), a.size, ptr(wid), ptr(tau), ptr(ln), cap)
        if n < 0:
            return None
        return [(int(wid[i]), int(tau[i]), int(ln[i])) for i in range(min(n, cap))]


class RefTables:
    """The static dictionary tables of the reference C99 decoder (src/brotli_tables.h, built
    into oracle/_ref/libprzm_ref_tables.so), read as data for independent checks."""

    def __init__(self):
        l = ctypes.CDLL(REFTABLES)
        l.przm_ref_table.restype = P
        l.przm_ref_table.argtypes = [ctypes.c_int]
        at = lambda k, t, n: (t * n).from_address(l.przm_ref_table(k))
        self.word_base = list(at(0, ctypes.c_uint16, 22))
        self.dict_offsets = list(at(1, ctypes.c_uint32, 21))
        tr = at(2, ctypes.c_uint8, 121 * 3)
        self.transforms = [tuple(tr[3 * i: 3 * i + 3]) for i in range(121)]
        cp = at(3, ctypes.c_uint16, 2048 * 2)
        self.compact = [(cp[2 * i], cp[2 * i + 1]) for i in range(2048)]

        class Str(ctypes.Structure):
            _fields_ = [("s", ctypes.c_void_p), ("n", ctypes.c_uint8)]
        self.ps = [ctypes.string_at(e.s, e.n) if e.n else b"" for e in (Str * 50).from_address(l.przm_ref_table(4))]
        self.top = [ctypes.string_at(e.s, e.n) if e.n else b"" for e in (Str * 512).from_address(l.przm_ref_table(5))]
        self.data = bytes(at(6, ctypes.c_uint8, 122784))
        self.l = l


_oracle = None
_ref = None
_refcpu = None


def refcpu():
    global _refcpu
    if _refcpu is None:
        _refcpu = RefCPU(oracle().l)
    return _refcpu


def have_reftables():
    return os.path.exists(REFTABLES)


def oracle():
    global _oracle
    if _oracle is None:
        _oracle = Oracle()
    return _oracle


def refdec():
    global _ref
    if _ref is None:
        _ref = RefDecoder()
    return _ref


def have_refdec():
    return os.path.exists(REFDEC)


def encode_many(bufs, level=6, threads=None):
    """The oracle's frames for many buffers, on a thread pool (the C calls release the GIL):
    list of (frames, sizes)."""
    from concurrent.futures import ThreadPoolExecutor
    O = oracle()
    threads = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda b: O.encode_buffer(b, level), bufs))
