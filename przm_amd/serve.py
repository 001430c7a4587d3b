"""Continuous-batching serving loop (przm_serve_* of include/przm_gpu.h).

Requests are admitted into launch slots as slots and HBM budget free up;
up to `streams` launches are in flight at once on separate HIP streams, so a
request that arrives while others run is encoded in the next launch, which
overlaps the running one.  `poisson_load` drives an open-loop Poisson
arrival process against a server and reports latency percentiles.
"""
import ctypes
import threading
import time

import numpy as np

from . import _lib as L
from .gpu import GPUError


class Server:
    def __init__(self, compressor, level=6, hbm_budget=64 << 30, max_batch_bytes=256 << 20, streams=3):
        self._lib = L.lib()
        h = ctypes.c_void_p()
        rc = self._lib.przm_serve_create(compressor.handle, level, hbm_budget, max_batch_bytes, streams,
                                         ctypes.byref(h))
        if rc:
            raise GPUError(rc)
        self._h = h

    def submit(self, data) -> int:
        a = data if isinstance(data, np.ndarray) else np.frombuffer(bytes(data), np.uint8)
        rid = ctypes.c_uint64()
        ptr = ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)
        rc = self._lib.przm_serve_submit(self._h, ptr, a.size, ctypes.byref(rid))
        if rc:
            raise GPUError(rc)
        return rid.value

    def wait(self, rid, want_output=True, timeout_ms=-1, cap=None):
        """(frames bytes or None, frame sizes or None, result record)."""
        res = L.Result()
        if want_output:
            cap = cap or (1 << 20)
            while True:
                out = np.zeros(cap, np.uint8)
                fs = np.zeros(cap // 8 + 16, np.uint32)
                rc = self._lib.przm_serve_wait(self._h, rid, timeout_ms, ctypes.c_void_p(out.ctypes.data), out.size,
                                               ctypes.c_void_p(fs.ctypes.data), fs.size, ctypes.byref(res))
                if rc == L.ERR_CAPACITY:
                    cap = max(cap * 2, int(res.out_len) + 64)
                    continue
                if rc:
                    raise GPUError(rc)
                return out[: res.out_len].tobytes(), fs[: res.nframes].copy(), res
        rc = self._lib.przm_serve_wait(self._h, rid, timeout_ms, None, 0, None, 0, ctypes.byref(res))
        if rc:
            raise GPUError(rc)
        return None, None, res

    def stats(self):
        s = (ctypes.c_uint64 * 8)()
        self._lib.przm_serve_stats(self._h, s)
        keys = ["launches", "done", "bytes_in", "bytes_out", "peak_admitted", "max_batch_reqs", "overlapped",
                "reserved"]
        return dict(zip(keys, list(s)))

    def close(self):
        if self._h:
            self._lib.przm_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_load(server, pool, rate=2000.0, duration_s=10.0, lo=4 << 10, hi=16 << 20, seed=7):
    """Open-loop Poisson arrivals at `rate` req/s for `duration_s`; request
    sizes log-uniform in [lo, hi], contents are slices of `pool`.  Returns a
    dict with latencies (ms, scheduled arrival -> frames ready, so time a
    request spends waiting to be submitted counts), input GB/s, counts."""
    rng = np.random.default_rng(seed)
    n_req = int(rate * duration_s)
    gaps = rng.exponential(1.0 / rate, n_req)
    arrivals = np.cumsum(gaps)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_req)).astype(np.int64)
    offs = (rng.uniform(0, 1, n_req) * (pool.size - sizes)).astype(np.int64)
    ids = [None] * n_req
    late = np.zeros(n_req)  # ms between a request's scheduled arrival and its submit call
    lat = np.zeros(n_req)
    done_at = np.zeros(n_req)
    submitted = threading.Event()
    err = []

    def producer():
        t0 = time.perf_counter()
        for i in range(n_req):
            wait = arrivals[i] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            late[i] = max(0.0, (time.perf_counter() - t0) - arrivals[i]) * 1e3
            ids[i] = server.submit(pool[offs[i]: offs[i] + sizes[i]])
        submitted.set()

    def consumer():
        try:
            for i in range(n_req):
                while ids[i] is None:
                    time.sleep(0.0002)
                _, _, res = server.wait(ids[i], want_output=False)
                lat[i] = res.done_ms - res.submit_ms + late[i]  # arrival -> frames ready
                done_at[i] = res.done_ms
        except Exception as e:  # surfaced to the caller
            err.append(e)

    t_start = time.perf_counter()
    tp = threading.Thread(target=producer)
    tc = threading.Thread(target=consumer)
    tp.start()
    tc.start()
    tp.join()
    tc.join()
    wall = time.perf_counter() - t_start
    if err:
        raise err[0]
    total = int(sizes.sum())
    return {
        "requests": n_req, "offered_req_s": rate, "duration_s": duration_s, "wall_s": wall,
        "input_bytes": total, "input_gbps": total / wall / 1e9,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat.max()), "mean_ms": float(lat.mean()),
        "submit_late_p99_ms": float(np.percentile(late, 99)),
    }
