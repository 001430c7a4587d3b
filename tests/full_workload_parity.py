"""Every buffer of the batched workload of bench.py (10,000 buffers, sizes log-uniform in
[1 KiB, 8 MiB], mixture content, seed 1, 9,473,959,928 bytes): the frames of one GPU batched
encode (przm_gpu_encode_batch through GPUCompressor.EncodeBatch) against the oracle's frames of
every buffer (orc_encode_buffer), compared by BLAKE2 digests computed in a pool of 16 oracle
workers.  With --options both section options are on (the Section-3 prefix container and the
Section-4 sheaf trials) on the GPU and in the oracle.  With --config2 the buffer is instead the
1 GiB mixture buffer (seed 2) through the plain entry point (CompressFrames), every one of its
16,384 frames against the oracle's.  A one-off check run on a GPU machine
(python3 tests/full_workload_parity.py [--options | --config2]; a few minutes), not collected by
pytest; exits non-zero on any mismatch."""
import hashlib, os, sys, time
import numpy as np
from multiprocessing import Pool
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))  # the repository
sys.path.insert(0, ROOT); sys.path.insert(0, os.path.join(ROOT, "tests"))
from przm_amd import datagen as D

ARENA = None
def init(arena, offs, sizes):
    global ARENA, OFFS, SIZES, O
    from oracle_lib import oracle
    ARENA, OFFS, SIZES = arena, offs, sizes
    O = oracle()
def work(job):
    lo, hi, opt = job
    out = []
    for i in range(lo, hi):
        a = ARENA[int(OFFS[i]): int(OFFS[i]) + int(SIZES[i])]
        f, s = O.encode_buffer(a, 6, opt, opt)
        out.append((i, hashlib.blake2b(f, digest_size=16).hexdigest(), int(len(f))))
    return out

def config2():
    """the 1 GiB mixture buffer through the plain entry point: every frame against the oracle's"""
    from przm_amd import gpu as G
    t = time.time()
    a = D.mixture(1 << 30, 2)
    g = G.GPUCompressor(0)
    frames, sizes, _ = g.CompressFrames(a)
    offs = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))])
    gh = [hashlib.blake2b(frames[offs[b]: offs[b + 1]], digest_size=16).hexdigest() for b in range(sizes.size)]
    print("[parity] 1 GiB plain encode: %d frames, %d bytes, %.1fs" % (sizes.size, len(frames), time.time() - t),
          flush=True)
    blk_offs = np.arange(sizes.size, dtype=np.int64) * 65536
    blk_sizes = np.minimum(65536, a.size - blk_offs)
    t = time.time()
    bad, seen = [], 0
    with Pool(16, initializer=init, initargs=(a, blk_offs, blk_sizes)) as pool:
        for out in pool.imap_unordered(work, [(k, min(k + 256, sizes.size), False) for k in range(0, sizes.size, 256)]):
            bad += [i for i, h, l in out if h != gh[i] or l != int(sizes[i])]
            seen += len(out)
    print("[parity] RESULT config2 frames %d compared %d bytes %d frame bytes %d mismatches %d %s (%.0fs)" % (
        sizes.size, seen, a.size, len(frames), len(bad), bad[:10], time.time() - t), flush=True)
    sys.exit(1 if bad or seen != sizes.size else 0)


if __name__ == "__main__":
    if "--config2" in sys.argv:
        config2()
    t0 = time.time()
    arena, offs, sizes = D.batch(10000, 1024, 8 << 20, 1)
    print("[parity] generated %.2f GiB in %.1fs" % (int(sizes.sum()) / 2**30, time.time() - t0), flush=True)
    from przm_amd import gpu as G
    g = G.GPUCompressor(0)
    opt = "--options" in sys.argv
    if opt:
        g.SetOffsetsMode(G.OffsetsPrefix)
        g.SetLiteralsMode(G.LiteralsSheaf)
    bufs = [arena[int(offs[i]): int(offs[i]) + int(sizes[i])] for i in range(10000)]
    t = time.time()
    res, _ = g.EncodeBatch(bufs, 6)
    gh = [hashlib.blake2b(f, digest_size=16).hexdigest() for f, s in res]
    glen = [len(f) for f, s in res]
    print("[parity] GPU batch in %.1fs, %d frames bytes" % (time.time() - t, sum(glen)), flush=True)
    t = time.time()
    chunks = [(k, min(k + 50, 10000), opt) for k in range(0, 10000, 50)]
    bad, done = [], 0
    with Pool(16, initializer=init, initargs=(arena, offs, sizes)) as pool:
        for out in pool.imap_unordered(work, chunks):
            for i, h, l in out:
                if h != gh[i] or l != glen[i]:
                    bad.append(i)
            done += len(out)
            if done % 1000 == 0:
                print("[parity] oracle %d buffers, %d mismatches, %.0fs" % (done, len(bad), time.time() - t), flush=True)
    print("[parity] RESULT options %s buffers 10000 compared %d bytes %d frame bytes %d mismatches %d %s" % (
        opt, done, int(sizes.sum()), sum(glen), len(bad), bad[:10]), flush=True)
    sys.exit(1 if bad or done != 10000 else 0)
