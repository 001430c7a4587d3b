"""ctypes binding of include/przm_decode.h (libprzm_decode.so): a CPU decoder of
the frames the encode path writes, with the reference's DecompressTo / C99
przm_decompress semantics for that subset of the format."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprzm_decode.so")

ERR_CORRUPT = -1
ERR_MAGIC = -2
ERR_CAPACITY = -3
ERR_UNSUPPORTED = -4
ERR_RANGE = -5
ERR_SIZE = -6

_lib = None


class DecodeError(Exception):
    def __init__(self, code):
        self.code = code
        super().__init__(f"przm decode error {code}")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with przm_amd/_build.py")
        l = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        l.przm_decode_size.restype = ctypes.c_int64
        l.przm_decode_size.argtypes = [P, ctypes.c_size_t]
        l.przm_decode_frame.restype = ctypes.c_int64
        l.przm_decode_frame.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t]
        l.przm_decode_frames.restype = ctypes.c_int64
        l.przm_decode_frames.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, P, ctypes.c_size_t]
        l.przm_decode_stream_size.restype = ctypes.c_int64
        l.przm_decode_stream_size.argtypes = [P, ctypes.c_size_t]
        l.przm_decode_stream.restype = ctypes.c_int64
        l.przm_decode_stream.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t]
        _lib = l
    return _lib


def _u8(x):
    if isinstance(x, np.ndarray):
        return np.ascontiguousarray(x, dtype=np.uint8)
    return np.frombuffer(bytes(x), dtype=np.uint8)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def GetUncompressedSize(frame) -> int:  # api.go GetUncompressedSize
    f = _u8(frame)
    r = lib().przm_decode_size(_p(f), f.size)
    if r < 0:
        raise DecodeError(r)
    return int(r)


def DecompressFrame(frame) -> bytes:
    f = _u8(frame)
    n = GetUncompressedSize(f)
    out = np.zeros(max(n, 1), np.uint8)
    r = lib().przm_decode_frame(_p(f), f.size, _p(out), out.size)
    if r < 0:
        raise DecodeError(r)
    return out[:r].tobytes()


def DecompressFrames(frames, sizes, total=None) -> bytes:
    """Frames back to back (sizes[i] bytes each) -> the concatenated input."""
    f = _u8(frames)
    s = np.ascontiguousarray(sizes, np.uint32)
    if total is None:
        total, pos = 0, 0
        for n in s:
            total += GetUncompressedSize(f[pos: pos + int(n)])
            pos += int(n)
    out = np.zeros(max(total, 1), np.uint8)
    r = lib().przm_decode_frames(_p(f), f.size, _p(s), s.size, _p(out), out.size)
    if r < 0:
        raise DecodeError(r)
    return out[:r].tobytes()


def ParseStreamHeader(stream):
    """(chunk count, total uncompressed size, flags) of a PRZS stream's 17-byte header
    (decompress.go:849-866 ParseStreamHeader)."""
    f = _u8(stream)
    if f.size < 17:
        raise DecodeError(ERR_CORRUPT)
    if int.from_bytes(f[:4].tobytes(), "little") != 0x535A5250:
        raise DecodeError(ERR_MAGIC)
    return (int.from_bytes(f[5:9].tobytes(), "little"), int.from_bytes(f[9:17].tobytes(), "little"), int(f[4]))


def Decompress(data) -> bytes:
    """A PRZS stream or a single frame -> the input (Decompress of decompress.go for these two forms)."""
    f = _u8(data)
    n = lib().przm_decode_stream_size(_p(f), f.size)
    if n < 0:
        raise DecodeError(n)
    out = np.zeros(max(int(n), 1), np.uint8)
    r = lib().przm_decode_stream(_p(f), f.size, _p(out), out.size)
    if r < 0:
        raise DecodeError(r)
    return out[:r].tobytes()
