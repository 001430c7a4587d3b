"""Seeded synthetic inputs of the encode workloads (libprzm_datagen.so).

markov(n, seed)   order-1 Markov text with 30% copies (16 MiB CPU case, 256 MiB plain GPU case)
mixture(n, seed)  JSON / int32 table / motif-random / uniform blocks, 4:3:2:1 (1 GiB case)
batch(count, lo, hi, seed)  buffers with log-uniform sizes, mixture content (batched / serving cases)
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libprzm_datagen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with przm_amd/_build.py")
        l = ctypes.CDLL(path)
        V = ctypes.c_void_p
        U64 = ctypes.c_uint64
        l.przm_gen_markov.argtypes = [V, U64, U64, ctypes.c_int]
        l.przm_gen_mixture.argtypes = [V, U64, U64, ctypes.c_int]
        l.przm_gen_mixture_kinds.argtypes = [U64, U64, V]
        l.przm_gen_loguniform_sizes.argtypes = [V, U64, U64, U64, U64]
        l.przm_gen_batch.argtypes = [V, V, V, U64, U64, ctypes.c_int]
        for f in (l.przm_gen_markov, l.przm_gen_mixture, l.przm_gen_mixture_kinds, l.przm_gen_loguniform_sizes,
                  l.przm_gen_batch):
            f.restype = None
        _lib = l
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def markov(n: int, seed: int = 1, threads: int = 0) -> np.ndarray:
    out = np.empty(max(n, 1), np.uint8)
    lib().przm_gen_markov(_p(out), n, seed, threads)
    return out[:n]


def mixture(n: int, seed: int = 1, threads: int = 0) -> np.ndarray:
    out = np.empty(max(n, 1), np.uint8)
    lib().przm_gen_mixture(_p(out), n, seed, threads)
    return out[:n]


def mixture_kinds(nblocks: int, seed: int = 1) -> np.ndarray:
    out = np.empty(max(nblocks, 1), np.uint8)
    lib().przm_gen_mixture_kinds(seed, nblocks, _p(out))
    return out[:nblocks]


def loguniform_sizes(count: int, lo: int, hi: int, seed: int = 1) -> np.ndarray:
    out = np.empty(max(count, 1), np.uint64)
    lib().przm_gen_loguniform_sizes(_p(out), count, lo, hi, seed)
    return out[:count]


def batch(count: int, lo: int, hi: int, seed: int = 1, threads: int = 0, align: int = 256):
    """(arena, offsets, sizes): buffer i = arena[offsets[i]: offsets[i] + sizes[i]], offsets aligned."""
    sizes = loguniform_sizes(count, lo, hi, seed)
    padded = (sizes + (align - 1)) // align * align
    offs = np.zeros(count, np.uint64)
    if count:
        offs[1:] = np.cumsum(padded)[:-1]
    total = int(offs[-1] + padded[-1]) if count else 0
    arena = np.zeros(max(total, 1), np.uint8)
    lib().przm_gen_batch(_p(arena), _p(offs), _p(sizes), count, seed, threads)
    return arena, offs, sizes
