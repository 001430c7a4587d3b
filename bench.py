#!/usr/bin/env python3
"""Benchmark of the MI355X PRZM encode path: one JSON line.

Workload (the batched form): 10,000 independent buffers with sizes
log-uniform in [1 KiB, 8 MiB] (about 8.7 GiB), contents from the mixture
generator (JSON-like records, int32 tables, motif-random, uniform random),
64 KiB frames at the default level 6, inputs resident in HBM.

  step      one pass over every buffer: block list upload + one hipGraph
            launch of the persistent stage pipeline + frame sizes back.
  value     whole-job input GB/s of the batched form (K timed steps).
  plain     the same buffers one at a time (one launch per stage each),
            frames compared byte for byte with the batched ones.
  plain_256MiB  one 256 MiB Markov-text buffer, 4096 blocks, one launch
            per stage.
  section_options  the same batch with the frame options of the reference CPU
            path: Section 3 as a distance-prefix container where shorter
            (przm_gpu_set_offsets_mode), Section 4 with the sheaf trials
            (przm_gpu_set_literals_mode), and both: GB/s, ratio, bytes saved,
            the options' kernel time, a decoded sample.
  serving   open-loop Poisson arrivals at 2,000 req/s, sizes log-uniform in
            [4 KiB, 16 MiB], 64 GiB HBM admission budget: p50/p99 latency.
  roofline  the dominant kernel (przm_lz_kernel) against the 8 TB/s HBM
            peak: algorithmic bytes per launch = (input bytes + frame bytes)
            of the blocks one launch processes; its average launch time
            comes from an un-graphed pass of the same batch right after the
            timed steps,
            with HIP events between the launches on their stream (events
            recorded inside a captured graph cannot be timed on this ROCm);
            HBM traffic per block from the committed rocprofv3 --pmc summary.
  --profile-only  one torch copy kernel of known size (the PMC scale) and one
            batched pass over the first --profile-buffers buffers, nothing
            else: the program the rocprofv3 --pmc passes run.
  cpu_baseline  the CPU oracle of the same mode on the host's cores over a
            time-bounded sample of the same buffers; beside it, config 0 (16 MiB
            of Markov text as 256 frames) through the C restatement of the
            reference CPU compressor (the trellis path at level 6) and through
            the plain GPU path.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "GPU encode input GB/s, batched vs plain (bit-exact); p50/p99 latency @2k req/s"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cpu():
    """Core counts and model of this host: nproc, the cores this process may run on, the model
    name from /proc/cpuinfo."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count() or 1, "usable_cores": usable, "cpu_model": model}


def cpu_threads():
    """Every core this process may use, capped by the host's share for one GPU when the
    environment states it (OMP_NUM_THREADS)."""
    usable = host_cpu()["usable_cores"]
    share = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(usable, int(share))) if share and share.isdigit() else usable


def cpu_baseline(bufs, seconds, threads=None, level=6):
    """The C oracle of the same mode (oracle/liboracle.so: the CPU fallback restated) on the
    buffers `bufs` in order, until `seconds` have passed or all are done; independent buffers
    spread over `threads` threads.  Returns (input GB/s, buffers done, bytes done, wall s)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from concurrent.futures import ThreadPoolExecutor
    from oracle_lib import oracle
    O = oracle()
    threads = threads or cpu_threads()
    done_bytes = 0
    t0 = time.perf_counter()
    i = 0
    with ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < seconds and i < len(bufs):
            batch = list(range(i, min(i + threads * 2, len(bufs))))
            i += len(batch)
            futs = [ex.submit(O.encode_buffer, bufs[j], level) for j in batch]
            for f, j in zip(futs, batch):
                f.result()
                done_bytes += int(bufs[j].size)
    wall = time.perf_counter() - t0
    return done_bytes / wall / 1e9, i, done_bytes, wall


def refcpu_leg(src, threads, level=6):
    """The C restatement of the reference Go CPU compressor (oracle/przm_refcpu.c: the trellis
    path of compressInternal at `level`, with the static dictionary when its tables are built)
    on src as independent 64 KiB frames, on `threads` threads and on one thread over the first
    16 frames; every frame decoded by the reference C99 decoder when it is built."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import refcpu, refdec, have_refdec
    R = refcpu()
    t0 = time.perf_counter()
    frames, sizes = R.compress_frames(src, level, threads)
    w = time.perf_counter() - t0
    one = src[: 16 * 65536]
    t0 = time.perf_counter()
    R.compress_frames(one, level, 1)
    w1 = time.perf_counter() - t0
    ok = None
    if have_refdec():
        ok = refdec().decompress_frames(frames, sizes) == src.tobytes()
    return {"value": round(src.size / w / 1e9, 5), "unit": "GB/s", "threads": threads,
            "one_thread": round(one.size / w1 / 1e9, 5), "compression_ratio": round(src.size / len(frames), 4),
            "frames": int(sizes.size), "dictionary": bool(R.dict), "decoded_by_reference_decoder": ok,
            "wall_s": round(w, 3)}


def pmc_traffic():
    """HBM bytes per block of przm_lz_kernel from the committed rocprofv3 --pmc summary, if present."""
    p = os.path.join(ROOT, "profiles", "lz_kernel_pmc.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d
    except Exception:
        return None


def profile_only(args, g, lib, L, D, G):
    """The program of the rocprofv3 --pmc passes: a copy kernel of known size, then one batched pass."""
    if args.lz_grid > 0 and args.chunk > 0:
        lib.przm_gpu_set_tuning(g.handle, args.lz_grid, args.chunk)
    lib.przm_gpu_set_kernel_shape(g.handle, args.search_batch, args.rans_jobs)
    n = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for x in (a, b):
        if lib.przm_gpu_malloc(g.handle, n, ctypes.byref(x)):
            raise G.GPUError(-6)
    lib.przm_gpu_copy_d2d(g.handle, b, a, n)  # warm
    lib.przm_gpu_copy_d2d(g.handle, b, a, n)  # reads n bytes, writes n bytes
    lib.przm_gpu_free(g.handle, a)
    lib.przm_gpu_free(g.handle, b)
    log(f"copy kernel: {n} bytes")
    arena, offs, sizes = D.batch(args.profile_buffers, 1024, 8 << 20, args.seed)
    bufs = [arena[int(o): int(o) + int(s)] for o, s in zip(offs, sizes)]
    res, prof = g.EncodeBatch(bufs, args.level)
    nblocks = sum(max((int(s) + 65535) // 65536, 1) for s in sizes)
    log(f"profile pass: {len(bufs)} buffers, {int(sizes.sum())} bytes, {nblocks} blocks, "
        f"{prof.PipelinedChunks if hasattr(prof, 'PipelinedChunks') else '?'} chunks")
    print(json.dumps({"profile_buffers": len(bufs), "input_bytes": int(sizes.sum()), "blocks": nblocks,
                      "copy_bytes": n}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--buffers", type=int, default=10000)
    ap.add_argument("--level", type=int, default=6)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-batch-mib", type=int, default=64, help="serving: largest launch (MiB)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-options", action="store_true", help="skip the section-option legs")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--lz-grid", type=int, default=0, help="fixed LZ workgroups (skips the autotuner with --chunk)")
    ap.add_argument("--chunk", type=int, default=0, help="fixed blocks per launch (skips the autotuner with --lz-grid)")
    ap.add_argument("--profile-only", action="store_true")
    ap.add_argument("--search-batch", type=int, default=0, help="fixed chain candidates per search step (2, 4, 8)")
    ap.add_argument("--rans-jobs", type=int, default=0, help="fixed coder jobs per workgroup (16, 32)")
    ap.add_argument("--profile-buffers", type=int, default=1500)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); frames are independent "
              f"per 64 KiB block, so the path would shard by buffer with no exchange step", file=sys.stderr)
        sys.exit(2)

    from przm_amd import _lib as L
    from przm_amd import datagen as D
    from przm_amd import gpu as G
    from przm_amd import serve as S

    lib = L.lib()
    g = G.GPUCompressor(0)
    log("device", g.DeviceName, g.Arch, g.SMCount)

    if args.profile_only:
        profile_only(args, g, lib, L, D, G)
        return

    t = time.perf_counter()
    arena, offs, sizes = D.batch(args.buffers, 1024, 8 << 20, args.seed)
    total_in = int(sizes.sum())
    log(f"generated {args.buffers} buffers, {total_in / 2**30:.2f} GiB in {time.perf_counter() - t:.1f}s")

    def dmalloc(n):
        p = ctypes.c_void_p()
        rc = lib.przm_gpu_malloc(g.handle, n, ctypes.byref(p))
        if rc:
            raise G.GPUError(rc)
        return p

    d_src = dmalloc(arena.size + 64)
    lib.przm_gpu_h2d(g.handle, d_src, ctypes.c_void_p(arena.ctypes.data), arena.size)
    nbuf = len(sizes)
    spans = (L.Span * nbuf)()
    nblocks = 0
    bound = 0
    bound_off = np.zeros(nbuf + 1, np.int64)
    for i in range(nbuf):
        spans[i].offset, spans[i].len = int(offs[i]), int(sizes[i])
        nblocks += max((int(sizes[i]) + 65535) // 65536, 1)
        bound += int(lib.przm_gpu_frames_bound(int(sizes[i])))
        bound_off[i + 1] = bound
    d_out = dmalloc(bound + 64)
    out_off = np.zeros(nbuf + 1, np.uint64)
    fsz = np.zeros(nblocks, np.uint32)
    nf = ctypes.c_uint64()

    tuned = (ctypes.c_int32 * 4)()
    if args.lz_grid > 0 and args.chunk > 0:
        lib.przm_gpu_set_tuning(g.handle, args.lz_grid, args.chunk)
        lib.przm_gpu_set_kernel_shape(g.handle, args.search_batch, args.rans_jobs)
        tuned[0], tuned[1] = args.lz_grid, args.chunk
        log(f"fixed tuning lz_grid={args.lz_grid} chunk={args.chunk}")
    elif not args.no_tune:
        t = time.perf_counter()
        rc = lib.przm_gpu_autotune(g.handle, d_src, spans, nbuf, args.level, tuned)
        if rc:
            raise G.GPUError(rc)
        log(f"autotune lz_grid={tuned[0]} chunk={tuned[1]} rans_jobs={tuned[2]} search_batch={tuned[3]} in "
            f"{time.perf_counter() - t:.1f}s")
    lib.przm_gpu_get_tuning(g.handle, tuned)

    def batched(prof=None):
        rc = lib.przm_gpu_encode_batch_device(g.handle, d_src, spans, nbuf, args.level, d_out, bound + 64,
                                              ctypes.c_void_p(out_off.ctypes.data), ctypes.c_void_p(fsz.ctypes.data),
                                              nblocks, ctypes.byref(nf), ctypes.byref(prof) if prof else None)
        if rc:
            raise G.GPUError(rc)

    for _ in range(args.warmup):
        batched()
    lib.przm_gpu_sync(g.handle)
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p = L.GpuProfile()
        batched(p)
        kms.append(p.kernel_ms)  # HIP events around the graph launch
    lib.przm_gpu_sync(g.handle)
    wall = time.perf_counter() - t0
    ms_step = wall * 1000 / args.steps
    value = total_in * args.steps / wall / 1e9
    total_frames = int(out_off[-1])
    log(f"batched: {value:.2f} GB/s input, {ms_step:.1f} ms/step, device {np.mean(kms):.1f} ms, ratio "
        f"{total_in / total_frames:.3f}")

    # profiled pass of the same batch without the graph: HIP events between the stages of every
    # chunk give each kernel's device time; the runtime also reports the LZ kernel's byte counts
    lib.przm_gpu_set_options(g.handle, 0)
    prof = L.GpuProfile()
    batched(prof)
    lib.przm_gpu_set_options(g.handle, 1)
    stage_names = ["check", "lz", "hist", "rans", "frame_size", "scan", "frame_write"]
    stages = {k: round(float(v), 3) for k, v in zip(stage_names, prof.stage_ms[:7])}
    n_lz = int(prof.pipelined_chunks)
    log(f"stages of the un-graphed pass (ms, summed over {n_lz} chunks): {stages}, device {prof.kernel_ms:.1f} ms")
    bat_frames = None

    # plain: one buffer at a time, one launch per stage
    plain = None
    d_out2 = None
    if not args.no_plain:
        d_out2 = dmalloc(bound + 64)
        lib.przm_gpu_sync(g.handle)
        t0 = time.perf_counter()
        olen = ctypes.c_uint64()
        nf2 = ctypes.c_uint64()
        for i in range(nbuf):
            rc = lib.przm_gpu_encode_device(g.handle, ctypes.c_void_p(d_src.value + int(offs[i])), int(sizes[i]),
                                            args.level, ctypes.c_void_p(d_out2.value + int(bound_off[i])),
                                            int(bound_off[i + 1] - bound_off[i]), ctypes.byref(olen), None, 0,
                                            ctypes.byref(nf2), None)
            if rc:
                raise G.GPUError(rc, G.L.last_error() if hasattr(G, "L") else "")
        lib.przm_gpu_sync(g.handle)
        pwall = time.perf_counter() - t0
        a1 = np.zeros(total_frames, np.uint8)
        a2 = np.zeros(bound, np.uint8)
        lib.przm_gpu_d2h(g.handle, ctypes.c_void_p(a1.ctypes.data), d_out, total_frames)
        lib.przm_gpu_d2h(g.handle, ctypes.c_void_p(a2.ctypes.data), d_out2, bound)
        exact = True
        for i in range(nbuf):
            o0, o1 = int(out_off[i]), int(out_off[i + 1])
            b0 = int(bound_off[i])
            if not np.array_equal(a1[o0:o1], a2[b0: b0 + (o1 - o0)]):
                exact = False
                log(f"plain/batched frames differ for buffer {i}")
                break
        del a2
        plain = {"value": round(total_in / pwall / 1e9, 4), "unit": "GB/s", "wall_s": round(pwall, 3),
                 "launches_per_buffer": 8, "bit_exact_vs_batched": exact}
        log(f"plain: {plain['value']:.3f} GB/s, bit-exact vs batched: {exact}")

    # the same batch with the section options (przm_gpu_set_offsets_mode / _literals_mode: Section 3
    # as a distance-prefix container, Section 4 with the sheaf trials, as the reference CPU path
    # writes them): timed like the headline form, one profiled pass for the options' kernels, and
    # the first buffers decoded by the library's decoder
    def option_leg(offsets, literals):
        from przm_amd import decode as X
        lib.przm_gpu_set_offsets_mode(g.handle, offsets)
        lib.przm_gpu_set_literals_mode(g.handle, literals)
        try:
            batched()
            lib.przm_gpu_sync(g.handle)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                batched()
            lib.przm_gpu_sync(g.handle)
            pw = time.perf_counter() - t0
            fb = int(out_off[-1])
            lib.przm_gpu_set_options(g.handle, 0)
            pprof = L.GpuProfile()
            batched(pprof)
            lib.przm_gpu_set_options(g.handle, 1)
            nchk = min(nbuf, 200)
            blob = np.zeros(int(out_off[nchk]), np.uint8)
            lib.przm_gpu_d2h(g.handle, ctypes.c_void_p(blob.ctypes.data), d_out, blob.size)
            ok, fi = True, 0
            for i in range(nchk):
                k = max((int(sizes[i]) + 65535) // 65536, 1)
                got = X.DecompressFrames(blob[int(out_off[i]): int(out_off[i + 1])], fsz[fi: fi + k])
                ok = ok and got == arena[int(offs[i]): int(offs[i]) + int(sizes[i])].tobytes()
                fi += k
        finally:
            lib.przm_gpu_set_offsets_mode(g.handle, 0)
            lib.przm_gpu_set_literals_mode(g.handle, 0)
        pst = {k: round(float(v), 3) for k, v in zip(stage_names + ["options"], list(pprof.stage_ms[:7]) +
                                                     [pprof.stage_ms[7]])}
        return {"value": round(total_in * args.steps / pw / 1e9, 4), "unit": "GB/s",
                "ms_per_step": round(pw * 1000 / args.steps, 3), "compression_ratio": round(total_in / fb, 4),
                "frame_bytes": fb, "bytes_saved_vs_default": total_frames - fb, "stages_ms": pst,
                "decoded_sample_buffers": nchk, "decoded_sample_ok": ok}

    options = None
    if not args.no_options:
        options = {}
        for name, o, l in (("prefix_offsets", 1, 0), ("sheaf_literals", 0, 1), ("both", 1, 1)):
            try:
                options[name] = option_leg(o, l)
                r = options[name]
                log(f"{name}: {r['value']:.2f} GB/s, ratio {r['compression_ratio']}, saved "
                    f"{r['bytes_saved_vs_default']} bytes, option kernels {r['stages_ms']['options']} ms, "
                    f"decode ok {r['decoded_sample_ok']}")
            except G.GPUError as e:
                options[name] = {"error": str(e)}
                log(f"{name}: {e}")

    # roofline of the dominant kernel: algorithmic bytes (block bytes read once + token-stream bytes
    # written once) per launch / average launch duration, from the profiled batched pass
    roofline = None
    if prof.stage_ms[1] > 0:
        launches = n_lz
        avg_ms = float(prof.stage_ms[1]) / launches  # events around each LZ launch, un-graphed pass
        # algorithmic bytes per unit: input bytes read once + frame bytes written once (intermediates
        # stay on chip in the ideal pipeline); one launch processes nblocks / n_lz blocks
        alg = (total_in + total_frames) / max(n_lz, 1)
        ach = alg / (avg_ms * 1e6)
        tr = pmc_traffic()
        traffic = int(tr["hbm_bytes_per_block"] * nblocks / max(n_lz, 1)) if tr else None
        kname = {2: "przm_lz_kernel_b2", 8: "przm_lz_kernel_b8"}.get(int(tuned[3]), "przm_lz_kernel")
        roofline = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBPS, 5), "traffic": traffic,
                    "kernel": kname, "launches": launches, "avg_launch_ms": round(avg_ms, 3),
                    "algorithmic_bytes_per_launch": int(alg), "blocks_per_launch": round(nblocks / max(n_lz, 1), 1),
                    "traffic_source": "profiles/lz_kernel_pmc.json" if tr else None}

    # 256 MiB Markov text, one buffer, one launch per stage
    plain256 = None
    if not args.no_plain:
        m = D.markov(256 << 20, args.seed)
        d_m = dmalloc(m.size + 64)
        lib.przm_gpu_h2d(g.handle, d_m, ctypes.c_void_p(m.ctypes.data), m.size)
        mb = int(lib.przm_gpu_frames_bound(m.size))
        d_mo = dmalloc(mb + 64)
        olen = ctypes.c_uint64()
        nf3 = ctypes.c_uint64()
        for rep in range(2):
            pp = L.GpuProfile()
            lib.przm_gpu_sync(g.handle)
            t0 = time.perf_counter()
            rc = lib.przm_gpu_encode_device(g.handle, d_m, m.size, args.level, d_mo, mb + 64, ctypes.byref(olen), None,
                                            0, ctypes.byref(nf3), ctypes.byref(pp))
            lib.przm_gpu_sync(g.handle)
            w = time.perf_counter() - t0
            if rc:
                raise G.GPUError(rc)
        plain256 = {"value": round(m.size / w / 1e9, 4), "unit": "GB/s", "device_ms": round(pp.kernel_ms, 3),
                    "ratio": round(m.size / olen.value, 4),
                    "stage_ms": [round(x, 3) for x in pp.stage_ms[:7]],
                    "lz_roofline_frac": round((pp.lz_in_bytes + pp.lz_stream_bytes) / (pp.stage_ms[1] * 1e6) /
                                              HBM_PEAK_GBPS, 5) if pp.stage_ms[1] > 0 else None}
        log(f"plain 256MiB markov: {plain256['value']:.3f} GB/s, stages {plain256['stage_ms']}")
        lib.przm_gpu_free(g.handle, d_m)
        lib.przm_gpu_free(g.handle, d_mo)

    serving = None
    if not args.no_serve and args.serve_seconds > 0:
        pool = D.mixture(512 << 20, args.seed + 7)
        srv = S.Server(g, level=args.level, hbm_budget=64 << 30, max_batch_bytes=args.serve_batch_mib << 20, streams=3)
        res = S.poisson_load(srv, pool, rate=2000.0, duration_s=args.serve_seconds)
        res["server"] = srv.stats()
        srv.close()
        serving = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res.items()}
        log(f"serving: p50 {serving['p50_ms']:.2f} ms p99 {serving['p99_ms']:.2f} ms, {serving['input_gbps']:.2f} GB/s")

    cpu = None
    if not args.no_cpu:
        hc = host_cpu()
        nt = cpu_threads()
        bufs = [arena[int(o): int(o) + int(s)] for o, s in zip(offs, sizes)]
        v, nb_done, by_done, w = cpu_baseline(bufs, args.cpu_seconds, nt, args.level)
        v1, nb1, by1, w1 = cpu_baseline(bufs, max(args.cpu_seconds / 3, 2.0), 1, args.level)
        # config 0: one 16 MiB buffer of Markov text as 256 independent 64 KiB frames
        m0 = D.markov(16 << 20, args.seed)
        frames0 = [m0[k: k + 65536] for k in range(0, m0.size, 65536)]
        c0, _, _, c0w = cpu_baseline(frames0, 1e9, nt, args.level)
        c01, _, _, c01w = cpu_baseline(frames0, 1e9, 1, args.level)
        cpu = {"value": round(v, 4), "unit": "GB/s", "cores": nt, "kind": "port",
               "label": "CPU fallback, same mode",
               "sample": f"the first {nb_done} buffers of the batched workload ({by_done} bytes) through "
                         f"oracle/liboracle.so (the C restatement of the reference CPU path in this mode), "
                         f"{nt} threads, {w:.1f} s",
               "nproc": hc["nproc"], "usable_cores": hc["usable_cores"], "cpu_model": hc["cpu_model"],
               "one_thread": {"value": round(v1, 4), "sample": f"first {nb1} buffers ({by1} bytes), {w1:.1f} s"},
               "markov_16MiB": {"value": round(c0, 4), "one_thread": round(c01, 4), "unit": "GB/s",
                                "sample": f"16 MiB order-1 Markov text (seed {args.seed}) as 256 frames of "
                                          f"64 KiB, level {args.level}: {nt} threads {c0w:.2f} s, 1 thread "
                                          f"{c01w:.2f} s"}}
        # config 0 through the reference CPU path itself (restated in C), and through the GPU
        ref0 = refcpu_leg(m0, nt, args.level)
        ref0["label"] = "reference CPU path (C restatement of compressInternal, level 6)"
        ref0["sample"] = (f"16 MiB order-1 Markov text (seed {args.seed}) as 256 frames of 64 KiB, "
                          f"{nt} threads; one thread on the first 16 frames")
        d0 = dmalloc(m0.size)
        lib.przm_gpu_h2d(g.handle, d0, ctypes.c_void_p(m0.ctypes.data), m0.size)
        o0 = dmalloc(lib.przm_gpu_frames_bound(m0.size) + 64)
        olen0, nf0 = ctypes.c_uint64(), ctypes.c_uint64()
        s0 = np.zeros(256, np.uint32)
        lib.przm_gpu_encode_device(g.handle, d0, m0.size, args.level, o0, lib.przm_gpu_frames_bound(m0.size) + 64,
                                   ctypes.byref(olen0), ctypes.c_void_p(s0.ctypes.data), 256, ctypes.byref(nf0), None)
        p0 = L.GpuProfile()
        lib.przm_gpu_encode_device(g.handle, d0, m0.size, args.level, o0, lib.przm_gpu_frames_bound(m0.size) + 64,
                                   ctypes.byref(olen0), ctypes.c_void_p(s0.ctypes.data), 256, ctypes.byref(nf0),
                                   ctypes.byref(p0))
        lib.przm_gpu_free(g.handle, d0)
        lib.przm_gpu_free(g.handle, o0)
        ref0["gpu_same_input"] = {"value": round(m0.size / (p0.kernel_ms * 1e6), 4), "unit": "GB/s",
                                  "compression_ratio": round(m0.size / olen0.value, 4),
                                  "path": "przm_gpu_encode_device, one launch per stage (mode G)"}
        cpu["reference_cpu_path_markov_16MiB"] = ref0
        log(f"cpu baseline: {cpu}")

    # measured copy bandwidth (the STREAM-copy figure beside the 8 TB/s peak)
    ncopy = 1 << 30
    ca, cb = dmalloc(ncopy), dmalloc(ncopy)
    lib.przm_gpu_copy_d2d(g.handle, cb, ca, ncopy)
    t0 = time.perf_counter()
    for _ in range(5):
        lib.przm_gpu_copy_d2d(g.handle, cb, ca, ncopy)
    cw = (time.perf_counter() - t0) / 5
    copy_gbps = round(2 * ncopy / cw / 1e9, 1)
    lib.przm_gpu_free(g.handle, ca)
    lib.przm_gpu_free(g.handle, cb)
    log(f"copy kernel: {copy_gbps} GB/s (read + write)")

    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "batched: 10,000 buffers, sizes log-uniform [1 KiB, 8 MiB], mixture content "
                               "(JSON / int32 tables / motif-random / uniform), 64 KiB frames, level 6 (chain depth 16), "
                               "one hipGraph launch per step",
                   "buffers": nbuf, "input_bytes": total_in, "frame_bytes": total_frames, "blocks": nblocks,
                   "level": args.level, "lz_grid": int(tuned[0]), "chunk_blocks": int(tuned[1]),
                   "rans_jobs_per_workgroup": int(tuned[2]), "search_batch": int(tuned[3]),
                   "tuning": "fixed" if (args.lz_grid > 0 and args.chunk > 0) else
                             ("defaults" if args.no_tune else "autotuned")},
        "device_ms_per_step": round(float(np.mean(kms)), 3),
        "stages_ms": stages,
        "ungraphed_device_ms": round(prof.kernel_ms, 3),
        "compression_ratio": round(total_in / total_frames, 4),
        "plain": plain,
        "batched_vs_plain": round(value / plain["value"], 3) if plain else None,
        "plain_256MiB_markov": plain256,
        "section_options": options,
        "serving": serving,
        "roofline": roofline,
        "copy_kernel_gbps": copy_gbps,
        "cpu_baseline": cpu,
        "device": {"name": g.DeviceName, "arch": g.Arch, "cus": g.SMCount},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
