"""Python mirror of the reference's Go `gpu` package API (gpu/bridge.go,
gpu/types.go, gpu/memory.go) over the C ABI of include/przm_gpu.h.

Names, argument meaning and error behaviour follow the Go package: every
method returns its results plus a profile, and raises GPUError where the Go
method returns an error.  There is no CPU fallback in here: the fallback of
the reference lives in its caller (compressInternal, compress.go:1036), and
`compress_with_fallback` below shows that wiring with a caller-supplied CPU
compressor.
"""
import ctypes
import threading
from dataclasses import dataclass, field
from typing import List, NamedTuple, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L

# gpu/types.go:8-16
Window64K = 64 * 1024
Window1M = 1024 * 1024
Window8M = 8 * 1024 * 1024
MinMatchLen = 4
MaxMatchLen32 = 4095
MaxOffset32 = 1048575
MaxMatchLen64 = 65535
# frame block of this implementation (one self-contained frame per block)
BlockSize = 64 * 1024
# Section 3 of the frames (GPUCompressor.SetOffsetsMode)
OffsetsFlat = 0    # 0x09 + the LEB128 varints
OffsetsPrefix = 1  # the distance-prefix container when shorter
# Section 4 of the frames (GPUCompressor.SetLiteralsMode)
LiteralsVANS = 0   # adaptive V-ANS when shorter, else raw
LiteralsSheaf = 1  # also the four sheaf containers
# Layout of an encoded buffer (GPUCompressor.SetContainer)
ContainerFrames = 0  # frames back to back, sizes beside them
ContainerStream = 1  # a buffer of more than one block as one PRZS stream
MagicPRZS = 0x535A5250  # decompress.go:25
StreamHeaderSize = 17   # decompress.go:849-866
# bridge.go:148-153, the reference tokenizer's tile and per-tile stream slots.  Here the tile is
# the 64 KiB block (window and extension bounded by it), so the reference's GPU tests that
# expect a 1 MiB window and a 4095 length cap (gpu_tier2_boundaries_test.go) see 64 KiB and 1024.
TokenizerTileSize = 32768
MaxTileCmdBytes = 36864
MaxTileOffBytes = 36864
MaxTileLitBytes = 36864


class GPUError(Exception):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"przm gpu error {code}: {what or L.last_error()}")


class LZ77Match(NamedTuple):  # gpu/types.go:19-24
    Position: int
    Offset: int
    Length: int
    Literal: int


def PackMatch32(length: int, offset: int) -> int:  # gpu/types.go:27-34
    if length < MinMatchLen or offset == 0:
        return 0
    return (min(length, MaxMatchLen32) << 20) | (offset & MaxOffset32)


def UnpackMatch32(desc: int) -> Tuple[int, int]:  # gpu/types.go:36-39
    return desc >> 20, desc & MaxOffset32


def IsMatch32(desc: int) -> bool:  # gpu/types.go:41-43
    return (desc >> 20) >= MinMatchLen


@dataclass
class MatchDesc64:  # gpu/types.go:46-52
    Length: int = 0
    Offset: int = 0
    Literal: int = 0
    Flags: int = 0  # 1: is_match
    Reserved: int = 0


@dataclass
class DeviceTokenStreams:  # gpu/types.go:90-100 (device pointers borrowed from the compressor)
    DFinalCmds: int = 0
    DFinalOffs: int = 0
    DFinalLits: int = 0
    DTotals: int = 0
    CmdLen: int = 0
    OffLen: int = 0
    LitLen: int = 0
    NumChunks: int = 0
    Stream: int = 0
    DBlockLens: int = 0
    raw: object = None  # the C struct, handed back to EncodeD2D


@dataclass
class PinnedBuffer:  # gpu/memory.go:93-97
    Bytes: np.ndarray
    Ptr: int


def AllocPinnedHost(size: int) -> Tuple[np.ndarray, int]:  # memory.go:51
    if size <= 0:
        raise GPUError(L.ERR_ARG, f"invalid size {size}")
    p = ctypes.c_void_p()
    rc = L.lib().przm_pinned_alloc(size, ctypes.byref(p))
    if rc:
        raise GPUError(rc)
    return _pinned_view(p.value, size), p.value


def FreePinnedHost(ptr: int) -> None:  # memory.go:78
    if ptr:
        rc = L.lib().przm_pinned_free(ctypes.c_void_p(ptr))
        if rc:
            raise GPUError(rc)


def _pinned_view(ptr: int, size: int) -> np.ndarray:
    return np.ctypeslib.as_array((ctypes.c_uint8 * size).from_address(ptr))


class PinnedBufferPool:  # memory.go:99-165
    def __init__(self, bufSize: int, initialCap: int = 0, _handle=None):
        self._lib = L.lib()
        self.bufSize = bufSize
        if _handle is not None:
            self._h = _handle
            self._owned = False
            return
        h = ctypes.c_void_p()
        rc = self._lib.przm_pinned_pool_create(bufSize, initialCap, ctypes.byref(h))
        if rc:
            raise GPUError(rc)
        self._h = h
        self._owned = True

    def Get(self) -> PinnedBuffer:
        p = ctypes.c_void_p()
        rc = self._lib.przm_pinned_pool_get(self._h, ctypes.byref(p))
        if rc:
            raise GPUError(rc)
        return PinnedBuffer(_pinned_view(p.value, self.bufSize), p.value)

    def Put(self, buf: PinnedBuffer) -> None:
        if buf.Ptr:
            self._lib.przm_pinned_pool_put(self._h, ctypes.c_void_p(buf.Ptr))

    def Close(self) -> None:
        if self._owned and self._h:
            self._lib.przm_pinned_pool_destroy(self._h)
        self._h = None


def NewPinnedBufferPool(bufSize: int, initialCap: int = 0) -> PinnedBufferPool:  # memory.go:110
    return PinnedBufferPool(bufSize, initialCap)


@dataclass
class GPUProfile:  # gpu/types.go:55-72
    IngressDMABytes: int = 0
    IngressDMATimeMs: float = 0.0
    IngressDMABandwidthGBps: float = 0.0
    KernelTimeMs: float = 0.0
    InternalVRAMBandwidthGBps: float = 0.0
    EgressDMABytes: int = 0
    EgressDMATimeMs: float = 0.0
    EgressDMABandwidthGBps: float = 0.0
    TotalRoundTripMs: float = 0.0
    EndToEndBandwidthGBps: float = 0.0
    PCIeOverheadPercent: float = 0.0
    DeviceName: str = ""
    SMCount: int = 0
    TotalVRAMBytes: int = 0
    PipelinedChunks: int = 0
    GraphLaunch: bool = False
    StageMs: List[float] = field(default_factory=list)


@dataclass
class TokenizerProfile:  # gpu/types.go:75-85
    KernelTimeMs: float = 0.0
    Pass1Ms: float = 0.0
    Pass2Ms: float = 0.0
    Pass3Ms: float = 0.0
    TotalRoundTripMs: float = 0.0
    ThroughputGBps: float = 0.0
    CmdBytes: int = 0
    OffsetBytes: int = 0
    LiteralBytes: int = 0
    NumChunks: int = 0


_mock_unavailable = False
_mock_err: Optional[Exception] = None
_global_mu = threading.Lock()
_global: Optional["GPUCompressor"] = None


def SetMockUnavailableForTest(unavailable: bool) -> None:  # bridge.go:45
    global _mock_unavailable
    _mock_unavailable = bool(unavailable)
    L.lib().przm_gpu_set_mock_unavailable(1 if unavailable else 0)


def SetMockMatchErrorForTest(err: Optional[Exception]) -> None:  # bridge.go:50
    global _mock_err
    _mock_err = err
    L.lib().przm_gpu_set_mock_error(L.ERR_MOCK if err is not None else 0)


def IsAvailable() -> bool:  # bridge.go:61
    if _mock_unavailable:
        return False
    try:
        lib = L.lib()
    except (OSError, RuntimeError):
        return False
    info = L.GpuInfo()
    return lib.przm_gpu_probe(ctypes.byref(info)) == 0


def NewGPUCompressor(deviceID: int = 0) -> "GPUCompressor":  # bridge.go:291
    """A compressor bound to one device (GPUError when it cannot be created)."""
    return GPUCompressor(deviceID)


def GetGlobalCompressor() -> "GPUCompressor":  # bridge.go:96
    global _global
    with _global_mu:
        if _global is not None:
            return _global
        if not IsAvailable():
            raise GPUError(L.ERR_UNAVAILABLE, "gpu acceleration not available")
        info = L.GpuInfo()
        L.lib().przm_gpu_probe(ctypes.byref(info))
        _global = GPUCompressor(info.device)
        return _global


def ResetGlobalCompressorForTest() -> None:  # bridge.go:117
    global _global
    with _global_mu:
        if _global is not None:
            _global.Close()
        _global = None


def _u8(a) -> np.ndarray:
    if isinstance(a, np.ndarray) and a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]:
        return a
    return np.frombuffer(bytes(a), dtype=np.uint8)


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def _profile(p: L.GpuProfile, g: "GPUCompressor") -> GPUProfile:
    return GPUProfile(p.ingress_dma_bytes, p.ingress_dma_ms, p.ingress_dma_gbps, p.kernel_ms, p.internal_vram_gbps,
                      p.egress_dma_bytes, p.egress_dma_ms, p.egress_dma_gbps, p.total_round_trip_ms,
                      p.end_to_end_gbps, p.pcie_overhead_percent, g.DeviceName, g.SMCount, g.TotalVRAMBytes,
                      p.pipelined_chunks, bool(p.graph_launch), list(p.stage_ms))


class GPUCompressor:
    """NewGPUCompressor / (*GPUCompressor) of bridge.go:291-3392."""

    def __init__(self, deviceID: int = 0):
        self._lib = L.lib()
        h = ctypes.c_void_p()
        rc = self._lib.przm_gpu_create(deviceID, ctypes.byref(h))
        if rc:
            raise GPUError(rc)
        self._h = h
        self._mu = threading.Lock()
        info = L.GpuInfo()
        self._lib.przm_gpu_get_info(h, ctypes.byref(info))
        self.ComputeMajor = info.compute_major
        self.ComputeMinor = info.compute_minor
        self.DeviceName = info.device_name.decode()
        self.Arch = info.arch.decode()
        self.SMCount = info.sm_count
        self.TotalVRAMBytes = info.total_vram_bytes

    # -- lifecycle --
    def Close(self):
        if self._h:
            self._lib.przm_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def SetOffsetsMode(self, mode: int) -> None:
        """Section 3 of the frames written from now on: OffsetsFlat (LEB128, the default) or
        OffsetsPrefix (the distance-prefix container when shorter, PackageOffsetsSection of
        core/distance/distance.go:825-849, as the reference CPU path writes it)."""
        self._check(self._lib.przm_gpu_set_offsets_mode(self._h, int(mode)))

    def SetLiteralsMode(self, mode: int) -> None:
        """Section 4 of the frames written from now on: LiteralsVANS (the default) or LiteralsSheaf
        (PackageLiteralsSection of core/literals/literals.go:667-693, with the four sheaf
        containers, as the reference CPU path writes it)."""
        self._check(self._lib.przm_gpu_set_literals_mode(self._h, int(mode)))

    def SetContainer(self, mode: int) -> None:
        """Layout of the buffers encoded from now on: ContainerFrames (frames back to back, the
        default) or ContainerStream (a buffer of more than one 64 KiB block as one PRZS stream:
        the header of decompress.go:849-866, then each frame behind its u32 length, as the
        reference writes multi-chunk input)."""
        self._check(self._lib.przm_gpu_set_container(self._h, int(mode)))
        self._container = int(mode)

    def _bound(self, n: int) -> int:
        if getattr(self, "_container", ContainerFrames) == ContainerStream:
            return int(self._lib.przm_gpu_stream_bound(n))
        return int(self._lib.przm_gpu_frames_bound(n))

    def _check(self, rc):
        if rc:
            raise GPUError(rc)

    def _mock(self):
        if _mock_err is not None:
            raise _mock_err

    # -- stages --
    def MatchDescriptors(self, src, level: int = 6) -> Tuple[np.ndarray, GPUProfile]:
        """Per-position (len << 16 | dist) of the mode's best match, 0 if none."""
        self._mock()
        a = _u8(src)
        desc = np.zeros(max(a.size, 1), np.uint32)
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_match(self._h, _ptr(a), a.size, level, _ptr(desc), ctypes.byref(p)))
        return desc[:a.size], _profile(p, self)

    def MatchLZ77(self, src, level: int = 6) -> Tuple[List[LZ77Match], GPUProfile]:  # bridge.go:1480
        a = _u8(src)
        if a.size == 0:
            return [], GPUProfile()
        if a.size > 16 << 20:  # as the reference: large payloads through the streaming path, 8 MiB chunks
            return self.MatchLZ77Streaming(a, 8 << 20, level)
        desc, prof = self.MatchDescriptors(a, level)
        pos = np.nonzero(desc)[0]
        out = [LZ77Match(int(p), int(desc[p] & 0xFFFF), int(desc[p] >> 16), int(a[p])) for p in pos]
        return out, prof

    def MatchDescriptorsStreaming(self, src, chunkSize: int = 0, level: int = 6):
        """MatchDescriptors through the double-buffered chunked path."""
        self._mock()
        a = _u8(src)
        desc = np.zeros(max(a.size, 1), np.uint32)
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_match_streaming(self._h, _ptr(a), a.size, level, chunkSize, _ptr(desc),
                                                       ctypes.byref(p)))
        return desc[:a.size], _profile(p, self)

    def MatchLZ77Streaming(self, src, chunkSize: int = 0, level: int = 6):  # bridge.go:1461
        a = _u8(src)
        if a.size == 0:
            return [], GPUProfile()
        desc, prof = self.MatchDescriptorsStreaming(a, chunkSize, level)
        pos = np.nonzero(desc)[0]
        return [LZ77Match(int(p), int(desc[p] & 0xFFFF), int(desc[p] >> 16), int(a[p])) for p in pos], prof

    # -- device-resident streams and device-buffer stages --
    def InputBuffer(self, n: int) -> int:
        """The compressor's device input buffer of at least n bytes (the reference's slot.dIn)."""
        d = ctypes.c_void_p()
        self._check(self._lib.przm_gpu_input_buffer(self._h, n, ctypes.byref(d)))
        return d.value

    def _streams(self, raw: L.DeviceTokenStreams) -> DeviceTokenStreams:
        return DeviceTokenStreams(raw.d_cmds, raw.d_offs, raw.d_lits, raw.d_totals, raw.cmd_len, raw.off_len,
                                  raw.lit_len, raw.num_chunks, raw.stream, raw.d_block_lens, raw)

    def TokenizeDevice(self, src, level: int = 6) -> Tuple[DeviceTokenStreams, TokenizerProfile]:  # bridge.go:2139
        self._mock()
        a = _u8(src)
        if a.size == 0:
            return DeviceTokenStreams(), TokenizerProfile()
        raw = L.DeviceTokenStreams()
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_tokenize_device(self._h, _ptr(a), a.size, level, ctypes.byref(raw),
                                                       ctypes.byref(p)))
        return self._streams(raw), self._tok_profile(raw, p, a.size)

    def TokenizeDeviceZeroCopy(self, src, level: int = 6):  # bridge.go:2247
        return self.TokenizeDevice(src, level)

    def TokenizePointersDeviceOnly(self, n: int, dIn: Optional[int] = None, level: int = 6):  # bridge.go:2163
        """Token streams of n bytes already in device memory (the compressor's input buffer by default)."""
        if n <= 0:
            return DeviceTokenStreams(), TokenizerProfile()
        d = dIn if dIn is not None else self.InputBuffer(n)
        raw = L.DeviceTokenStreams()
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_tokenize_device_resident(self._h, ctypes.c_void_p(d), n, level,
                                                                ctypes.byref(raw), ctypes.byref(p)))
        return self._streams(raw), self._tok_profile(raw, p, n)

    def TokenizeDeviceOnly(self, n: int, dIn: Optional[int] = None, level: int = 6):  # bridge.go:1915
        """Host copies of the cmd / offset / literal streams of n bytes already in device memory."""
        if n <= 0:
            return b"", b"", b"", TokenizerProfile()
        st, prof = self.TokenizePointersDeviceOnly(n, dIn, level)
        return (self.DownloadDeviceBytes(st.DFinalCmds, st.CmdLen), self.DownloadDeviceBytes(st.DFinalOffs, st.OffLen),
                self.DownloadDeviceBytes(st.DFinalLits, st.LitLen), prof)

    @staticmethod
    def _tok_profile(raw, p, n) -> TokenizerProfile:
        return TokenizerProfile(KernelTimeMs=p.kernel_ms, Pass1Ms=p.stage_ms[1], Pass2Ms=p.kernel_ms - p.stage_ms[1],
                                TotalRoundTripMs=p.kernel_ms,
                                ThroughputGBps=n / (p.kernel_ms * 1e6) if p.kernel_ms > 0 else 0.0,
                                CmdBytes=raw.cmd_len, OffsetBytes=raw.off_len, LiteralBytes=raw.lit_len,
                                NumChunks=raw.num_chunks)

    def DownloadDeviceBytes(self, dPtr: int, size: int) -> bytes:  # bridge.go:2253
        if size == 0:
            return b""
        out = np.zeros(size, np.uint8)
        self._check(self._lib.przm_gpu_download(self._h, _ptr(out), dPtr, size))
        return out.tobytes()

    def EncodeD2D(self, streams: DeviceTokenStreams, uSize: int, windowLog: int = 0):  # bridge.go:2466
        """Frames of the latest token streams: (frames, frame sizes, profile)."""
        if uSize <= 0:
            return b"", np.zeros(0, np.uint32), GPUProfile()
        bound = int(self._lib.przm_gpu_frames_bound(uSize))
        out = np.zeros(bound + 64, np.uint8)
        nb = max((uSize + BlockSize - 1) // BlockSize, 1)
        sizes = np.zeros(nb, np.uint32)
        olen, nf = ctypes.c_uint64(), ctypes.c_uint64()
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_encode_streams(self._h, ctypes.byref(streams.raw), uSize, windowLog, _ptr(out),
                                                      out.size, ctypes.byref(olen), _ptr(sizes), nb, ctypes.byref(nf),
                                                      ctypes.byref(p)))
        return out[: olen.value].tobytes(), sizes[: nf.value].copy(), _profile(p, self)

    def ComputeHistogramGPU(self, dIn: int, n: int, stream: int = 0):  # bridge.go:2890
        """(counts, kernel ms, GB/s) of device bytes dIn[0, n); the compressor's own stream is used."""
        counts = np.zeros(256, np.uint32)
        if n <= 0:
            return counts, 0.0, 0.0
        ms = ctypes.c_double()
        self._check(self._lib.przm_gpu_histogram_device(self._h, ctypes.c_void_p(dIn), n, _ptr(counts),
                                                        ctypes.byref(ms)))
        return counts, ms.value, n / (ms.value * 1e6) if ms.value > 0 else 0.0

    def PackStreamFlatGPU(self, dIn: int, dOut: int, n: int, stream: int = 0):  # bridge.go:3033
        if n <= 0:
            return 0, 0.0, 0.0
        olen, ms = ctypes.c_uint64(), ctypes.c_double()
        self._check(self._lib.przm_gpu_pack_flat_device(self._h, ctypes.c_void_p(dIn), ctypes.c_void_p(dOut), n,
                                                        ctypes.byref(olen), ctypes.byref(ms)))
        return olen.value, ms.value, 2 * n / (ms.value * 1e6) if ms.value > 0 else 0.0

    def VANSEncode4xGPU(self, dIn: int, dOut: int, n: int, freq, cdf, matBytes, stream: int = 0,
                        outCap: Optional[int] = None):  # bridge.go:3161
        if n <= 0:
            return 0, 0.0, 0.0
        f = np.ascontiguousarray(freq, np.uint16)
        c = np.ascontiguousarray(cdf, np.uint16)
        m = np.frombuffer(bytes(matBytes), np.uint8) if len(matBytes) else np.zeros(1, np.uint8)
        olen, ms = ctypes.c_uint64(), ctypes.c_double()
        cap = outCap if outCap is not None else 8 + len(matBytes) + 16 + 2 * n + 16
        self._check(self._lib.przm_gpu_vans_encode_device(self._h, ctypes.c_void_p(dIn), ctypes.c_void_p(dOut), n,
                                                          _ptr(f), _ptr(c), _ptr(m), len(matBytes), cap,
                                                          ctypes.byref(olen), ctypes.byref(ms)))
        return olen.value, ms.value, n / (ms.value * 1e6) if ms.value > 0 else 0.0

    # -- probes --
    def BenchmarkVRAMBandwidth(self, sizeMB: int) -> float:  # bridge.go:2674
        v = ctypes.c_double()
        self._check(self._lib.przm_gpu_bench_vram(self._h, sizeMB, ctypes.byref(v)))
        return v.value

    def BenchmarkD2DEntropy(self, sizeMB: int = 16, iterations: int = 10) -> float:  # bridge.go:2547
        v = ctypes.c_double()
        self._check(self._lib.przm_gpu_bench_d2d_entropy(self._h, sizeMB, iterations, ctypes.byref(v)))
        return v.value

    def BenchmarkEntropyVRAM(self, sizeMB: int = 16, iterations: int = 10) -> float:  # bridge.go:3325
        v = ctypes.c_double()
        self._check(self._lib.przm_gpu_bench_entropy_vram(self._h, sizeMB, iterations, ctypes.byref(v)))
        return v.value

    # -- pinned memory --
    def GetPinned(self, size: int) -> Tuple[PinnedBuffer, PinnedBufferPool]:  # bridge.go:452
        p, pool = ctypes.c_void_p(), ctypes.c_void_p()
        self._check(self._lib.przm_gpu_get_pinned(self._h, size, ctypes.byref(p), ctypes.byref(pool)))
        cap = 1
        while cap < size:
            cap <<= 1
        return PinnedBuffer(_pinned_view(p.value, size), p.value), PinnedBufferPool(cap, _handle=pool)

    def TokenizeStreams(self, src, level: int = 6):
        """Every stream of every block: dict of cmd/off/lit/op/meta/var byte arrays + per-block lengths."""
        self._mock()
        a = _u8(src)
        nb = max((a.size + BlockSize - 1) // BlockSize, 1)
        caps = [nb * 36864, nb * 49152, nb * 65536, nb * 36864, nb * 18432, nb * 36864]
        bufs = [np.zeros(c, np.uint8) for c in caps]
        lens = np.zeros((nb, 7), np.uint32)
        p = L.GpuProfile()
        args = []
        for b, c in zip(bufs, caps):
            args += [_ptr(b), c]
        self._check(self._lib.przm_gpu_tokenize(self._h, _ptr(a), a.size, level, *args, _ptr(lens), nb,
                                                ctypes.byref(p)))
        if a.size == 0:
            lens = lens[:0]
        names = ["cmd", "off", "lit", "op", "meta", "var"]
        tot = lens.sum(axis=0) if lens.size else np.zeros(7, np.uint64)
        out = {n: bufs[i][: int(tot[i])].copy() for i, n in enumerate(names)}
        return out, lens, _profile(p, self)

    def Tokenize(self, src, level: int = 6):  # bridge.go:1895
        s, lens, prof = self.TokenizeStreams(src, level)
        tp = TokenizerProfile(KernelTimeMs=prof.KernelTimeMs, CmdBytes=int(s["cmd"].size),
                              OffsetBytes=int(s["off"].size), LiteralBytes=int(s["lit"].size),
                              NumChunks=int(lens.shape[0]))
        return s["cmd"].tobytes(), s["off"].tobytes(), s["lit"].tobytes(), tp

    def ComputeHistogramHost(self, src) -> Tuple[np.ndarray, GPUProfile]:  # bridge.go:2961
        self._mock()
        a = _u8(src)
        counts = np.zeros(256, np.uint32)
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_histogram(self._h, _ptr(a), a.size, _ptr(counts), ctypes.byref(p)))
        return counts, _profile(p, self)

    def VANSEncode4xHost(self, src) -> Tuple[bytes, GPUProfile]:  # bridge.go:3233
        self._mock()
        a = _u8(src)
        out = np.zeros(8 + 600 + 16 + 2 * a.size + 64, np.uint8)
        n = ctypes.c_uint64()
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_vans_encode(self._h, _ptr(a), a.size, _ptr(out), out.size, ctypes.byref(n),
                                                   ctypes.byref(p)))
        return out[: n.value].tobytes(), _profile(p, self)

    def PackStreamFlatHost(self, src) -> Tuple[bytes, GPUProfile]:  # bridge.go:3079
        self._mock()
        a = _u8(src)
        out = np.zeros(a.size + 5, np.uint8)
        n = ctypes.c_uint64()
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_pack_flat(self._h, _ptr(a), a.size, _ptr(out), out.size, ctypes.byref(n),
                                                 ctypes.byref(p)))
        return out[: n.value].tobytes(), _profile(p, self)

    # -- encode --
    def CompressFrames(self, src, level: int = 6) -> Tuple[bytes, np.ndarray, GPUProfile]:
        """Frames of src (one per 64 KiB block) and their sizes; with ContainerStream and more than
        one block, the PRZS stream of them and the bytes each block occupies in it."""
        self._mock()
        a = _u8(src)
        bound = self._bound(a.size)
        out = np.zeros(bound + 64, np.uint8)
        nb = max((a.size + BlockSize - 1) // BlockSize, 1)
        sizes = np.zeros(nb, np.uint32)
        olen, nf = ctypes.c_uint64(), ctypes.c_uint64()
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_encode(self._h, _ptr(a), a.size, level, _ptr(out), out.size, ctypes.byref(olen),
                                              _ptr(sizes), nb, ctypes.byref(nf), ctypes.byref(p)))
        return out[: olen.value].tobytes(), sizes[: nf.value].copy(), _profile(p, self)

    def CompressD2D(self, src, level: int = 6) -> Tuple[bytes, GPUProfile]:  # bridge.go:2486
        frames, _, prof = self.CompressFrames(src, level)
        return frames, prof

    def EncodeBatch(self, bufs: Sequence, level: int = 6):
        """Many host buffers through the batched (graph) path: list of (frames, sizes)."""
        self._mock()
        arrs = [_u8(b) for b in bufs]
        offs = np.zeros(len(arrs) + 1, np.uint64)
        for i, a in enumerate(arrs):
            offs[i + 1] = offs[i] + a.size
        arena = np.concatenate(arrs) if arrs and offs[-1] else np.zeros(1, np.uint8)
        spans = (L.Span * max(len(arrs), 1))()
        nb = 0
        bound = 0
        for i, a in enumerate(arrs):
            spans[i].offset, spans[i].len = int(offs[i]), a.size
            nb += max((a.size + BlockSize - 1) // BlockSize, 1)
            bound += self._bound(a.size)
        out = np.zeros(bound + 64, np.uint8)
        out_off = np.zeros(len(arrs) + 1, np.uint64)
        sizes = np.zeros(max(nb, 1), np.uint32)
        nf = ctypes.c_uint64()
        p = L.GpuProfile()
        self._check(self._lib.przm_gpu_encode_batch(self._h, _ptr(arena), spans, len(arrs), level, _ptr(out), out.size,
                                                    _ptr(out_off), _ptr(sizes), nb, ctypes.byref(nf),
                                                    ctypes.byref(p)))
        res = []
        fi = 0
        for i, a in enumerate(arrs):
            k = max((a.size + BlockSize - 1) // BlockSize, 1)
            res.append((out[int(out_off[i]): int(out_off[i + 1])].tobytes(), sizes[fi: fi + k].copy()))
            fi += k
        return res, _profile(p, self)


def NormalizeFreq(counts, targetM: int = 4096) -> np.ndarray:  # bridge.go:2777
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    f = np.zeros(256, np.uint16)
    rc = L.lib().przm_normalize_freq(_ptr(c), targetM, _ptr(f))
    if rc:
        raise GPUError(rc, "NormalizeFreq")
    return f


def BuildCDF(freq) -> np.ndarray:  # bridge.go:2832
    f = np.ascontiguousarray(freq, dtype=np.uint16)
    c = np.zeros(257, np.uint16)
    L.lib().przm_build_cdf(_ptr(f), _ptr(c))
    return c


def SerializeMatrix(freq) -> bytes:  # bridge.go:2844
    f = np.ascontiguousarray(freq, dtype=np.uint16)
    out = np.zeros(600, np.uint8)
    n = ctypes.c_uint64()
    rc = L.lib().przm_serialize_matrix(_ptr(f), _ptr(out), out.size, ctypes.byref(n))
    if rc:
        raise GPUError(rc, "SerializeMatrix")
    return out[: n.value].tobytes()


def compress_with_fallback(src, cpu_compress, level: int = 6, gpu_acceleration: bool = True):
    """compressInternal (compress.go:1036) with GPUAcceleration wired: the GPU
    frames when the device is usable and the call succeeds, otherwise the
    caller's CPU compressor, with no error surfaced (the reference's silent
    fallback, tests/e2e/challenger_gpu_fault_tolerance_test.go:183-214)."""
    if gpu_acceleration and IsAvailable():
        try:
            g = GetGlobalCompressor()
            frames, _ = g.CompressD2D(src, level)
            return frames, "gpu"
        except Exception:
            pass
    return cpu_compress(src, level), "cpu"
