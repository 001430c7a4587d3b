"""ctypes binding of include/przm_gpu.h (libprzm_gpu.so, built in-tree).

Loading fails loudly: there is no CPU fallback inside this package.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprzm_gpu.so")

P = ctypes.c_void_p
U8P = ctypes.POINTER(ctypes.c_uint8)
U16P = ctypes.POINTER(ctypes.c_uint16)
U32P = ctypes.POINTER(ctypes.c_uint32)
U64P = ctypes.POINTER(ctypes.c_uint64)
I32 = ctypes.c_int32
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64

OK = 0
ERR_UNAVAILABLE = -1
ERR_ARG = -2
ERR_CAPACITY = -3
ERR_HIP = -4
ERR_MOCK = -5
ERR_OOM = -6
ERR_BUSY = -7


class GpuInfo(ctypes.Structure):
    _fields_ = [("device", I32), ("compute_major", I32), ("compute_minor", I32), ("sm_count", I32),
                ("total_vram_bytes", U64), ("device_name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64)]


class GpuProfile(ctypes.Structure):
    _fields_ = [("ingress_dma_bytes", U64), ("ingress_dma_ms", ctypes.c_double),
                ("ingress_dma_gbps", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("internal_vram_gbps", ctypes.c_double), ("egress_dma_bytes", U64),
                ("egress_dma_ms", ctypes.c_double), ("egress_dma_gbps", ctypes.c_double),
                ("total_round_trip_ms", ctypes.c_double), ("end_to_end_gbps", ctypes.c_double),
                ("pcie_overhead_percent", ctypes.c_double), ("pipelined_chunks", I32), ("graph_launch", I32),
                ("stage_ms", ctypes.c_double * 8), ("lz_in_bytes", U64), ("lz_stream_bytes", U64),
                ("frame_bytes", U64), ("raw_frames", U64)]


class Span(ctypes.Structure):
    _fields_ = [("offset", U64), ("len", U64)]


class DeviceTokenStreams(ctypes.Structure):  # przm_device_token_streams (gpu/types.go:90-100)
    _fields_ = [("d_cmds", U64), ("d_offs", U64), ("d_lits", U64), ("d_totals", U64), ("cmd_len", U32),
                ("off_len", U32), ("lit_len", U32), ("num_chunks", U32), ("stream", U64), ("d_block_lens", U64),
                ("uncompressed_size", U64), ("generation", U64)]


class Result(ctypes.Structure):
    _fields_ = [("id", U64), ("status", I32), ("nframes", U32), ("out_len", U64), ("submit_ms", ctypes.c_double),
                ("start_ms", ctypes.c_double), ("done_ms", ctypes.c_double)]


# name -> (restype, argtypes); every symbol declared in include/przm_gpu.h
SIGNATURES = {
    "przm_gpu_probe": (I32, [ctypes.POINTER(GpuInfo)]),
    "przm_gpu_create": (I32, [I32, ctypes.POINTER(P)]),
    "przm_gpu_destroy": (I32, [P]),
    "przm_gpu_get_info": (I32, [P, ctypes.POINTER(GpuInfo)]),
    "przm_gpu_set_mock_unavailable": (None, [I32]),
    "przm_gpu_set_mock_error": (None, [I32]),
    "przm_gpu_match": (I32, [P, P, U64, I32, P, ctypes.POINTER(GpuProfile)]),
    "przm_gpu_match_streaming": (I32, [P, P, U64, I32, U64, P, ctypes.POINTER(GpuProfile)]),
    "przm_gpu_tokenize": (I32, [P, P, U64, I32, P, U64, P, U64, P, U64, P, U64, P, U64, P, U64, P, U64,
                                ctypes.POINTER(GpuProfile)]),
    "przm_gpu_histogram": (I32, [P, P, U64, P, ctypes.POINTER(GpuProfile)]),
    "przm_gpu_vans_encode": (I32, [P, P, U64, P, U64, U64P, ctypes.POINTER(GpuProfile)]),
    "przm_gpu_pack_flat": (I32, [P, P, U64, P, U64, U64P, ctypes.POINTER(GpuProfile)]),
    "przm_gpu_histogram_device": (I32, [P, P, U64, P, ctypes.POINTER(ctypes.c_double)]),
    "przm_gpu_pack_flat_device": (I32, [P, P, P, U64, U64P, ctypes.POINTER(ctypes.c_double)]),
    "przm_gpu_vans_encode_device": (I32, [P, P, P, U64, P, P, P, U32, U64, U64P, ctypes.POINTER(ctypes.c_double)]),
    "przm_gpu_input_buffer": (I32, [P, U64, ctypes.POINTER(P)]),
    "przm_gpu_tokenize_device": (I32, [P, P, U64, I32, ctypes.POINTER(DeviceTokenStreams),
                                       ctypes.POINTER(GpuProfile)]),
    "przm_gpu_tokenize_device_resident": (I32, [P, P, U64, I32, ctypes.POINTER(DeviceTokenStreams),
                                                ctypes.POINTER(GpuProfile)]),
    "przm_gpu_download": (I32, [P, P, U64, U64]),
    "przm_gpu_encode_streams": (I32, [P, ctypes.POINTER(DeviceTokenStreams), U64, I32, P, U64, U64P, P, U64, U64P,
                                      ctypes.POINTER(GpuProfile)]),
    "przm_gpu_bench_vram": (I32, [P, U32, ctypes.POINTER(ctypes.c_double)]),
    "przm_gpu_bench_d2d_entropy": (I32, [P, U32, U32, ctypes.POINTER(ctypes.c_double)]),
    "przm_gpu_bench_entropy_vram": (I32, [P, U32, U32, ctypes.POINTER(ctypes.c_double)]),
    "przm_pinned_pool_create": (I32, [U64, U32, ctypes.POINTER(P)]),
    "przm_pinned_pool_get": (I32, [P, ctypes.POINTER(P)]),
    "przm_pinned_pool_put": (I32, [P, P]),
    "przm_pinned_pool_destroy": (I32, [P]),
    "przm_gpu_get_pinned": (I32, [P, U64, ctypes.POINTER(P), ctypes.POINTER(P)]),
    "przm_gpu_encode": (I32, [P, P, U64, I32, P, U64, U64P, P, U64, U64P, ctypes.POINTER(GpuProfile)]),
    "przm_gpu_encode_device": (I32, [P, P, U64, I32, P, U64, U64P, P, U64, U64P, ctypes.POINTER(GpuProfile)]),
    "przm_gpu_encode_device_noalloc": (I32, [P, P, U64, I32, P, U64, U64P, P, U64, U64P]),
    "przm_gpu_encode_batch_device": (I32, [P, P, ctypes.POINTER(Span), U32, I32, P, U64, P, P, U64, U64P,
                                           ctypes.POINTER(GpuProfile)]),
    "przm_gpu_encode_batch": (I32, [P, P, ctypes.POINTER(Span), U32, I32, P, U64, P, P, U64, U64P,
                                    ctypes.POINTER(GpuProfile)]),
    "przm_gpu_frames_bound": (U64, [U64]),
    "przm_gpu_autotune": (I32, [P, P, ctypes.POINTER(Span), U32, I32, ctypes.POINTER(I32)]),
    "przm_gpu_set_tuning": (I32, [P, I32, I32]),
    "przm_gpu_set_kernel_shape": (I32, [P, I32, I32]),
    "przm_gpu_set_offsets_mode": (I32, [P, I32]),
    "przm_gpu_set_literals_mode": (I32, [P, I32]),
    "przm_gpu_set_container": (I32, [P, I32]),
    "przm_gpu_stream_bound": (U64, [U64]),
    "przm_gpu_get_tuning": (I32, [P, ctypes.POINTER(I32)]),
    "przm_gpu_set_options": (I32, [P, I32]),
    "przm_gpu_malloc": (I32, [P, U64, ctypes.POINTER(P)]),
    "przm_gpu_free": (I32, [P, P]),
    "przm_gpu_h2d": (I32, [P, P, P, U64]),
    "przm_gpu_d2h": (I32, [P, P, P, U64]),
    "przm_gpu_copy_d2d": (I32, [P, P, P, U64]),
    "przm_gpu_sync": (I32, [P]),
    "przm_pinned_alloc": (I32, [U64, ctypes.POINTER(P)]),
    "przm_pinned_free": (I32, [P]),
    "przm_normalize_freq": (I32, [P, U32, P]),
    "przm_build_cdf": (I32, [P, P]),
    "przm_serialize_matrix": (I32, [P, P, U64, U64P]),
    "przm_serve_create": (I32, [P, I32, U64, U64, U32, ctypes.POINTER(P)]),
    "przm_serve_submit": (I32, [P, P, U64, U64P]),
    "przm_serve_wait": (I32, [P, U64, ctypes.c_int64, P, U64, P, U64, ctypes.POINTER(Result)]),
    "przm_serve_stats": (I32, [P, U64P]),
    "przm_serve_destroy": (I32, [P]),
    "przm_gpu_last_error": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with przm_amd/_build.py (no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def last_error():
    e = lib().przm_gpu_last_error()
    return e.decode(errors="replace") if e else ""
