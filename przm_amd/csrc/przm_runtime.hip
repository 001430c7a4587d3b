/*
 * przm_runtime.hip -- host runtime behind include/przm_gpu.h.
 *
 * Replaces the reference's Go driver-API bridge (gpu/bridge.go,
 * gpu/driver*.go, gpu/memory.go): device/context set-up, grow-only device
 * arenas, pinned staging, the stage launch sequence, hipGraph capture of
 * the batched form, an autotuner for the launch shapes, and a
 * continuous-batching serving loop.  Kernels live in przm_lz.hip and
 * przm_entropy.hip (one translation unit, so no relocatable device code).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/przm_gpu.h"
#include "przm_entropy.hip"
#include "przm_lz.hip"
#include "przm_offsets.hip"
#include "przm_sheaf.hip"

namespace {

thread_local char g_err[512];
#ifdef PRZM_DIAG
uint64_t *g_phase_ts = nullptr; /* per-block phase timestamps of the LZ kernel (device memory) */
#endif
std::atomic<int32_t> g_mock_unavailable{0};
std::atomic<int32_t> g_mock_error{0};

#define HIPCHK(x)                                                                             \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) {                                                               \
            snprintf(g_err, sizeof g_err, "%s:%d %s: %s", __FILE__, __LINE__, #x,             \
                     hipGetErrorString(e_));                                                  \
            return PRZM_GPU_ERR_HIP;                                                          \
        }                                                                                     \
    } while (0)

#define CHK(x)                  \
    do {                        \
        int32_t r_ = (x);       \
        if (r_ != 0) return r_; \
    } while (0)


struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int32_t ensure(size_t n, uint64_t *gen) {
        if (n <= cap && p) return 0;
        if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
        size_t c = n < 4096 ? 4096 : n;
        if (hipMalloc(&p, c) != hipSuccess) {
            p = nullptr;
            snprintf(g_err, sizeof g_err, "hipMalloc(%zu) failed", c);
            return PRZM_GPU_ERR_OOM;
        }
        cap = c;
        if (gen) (*gen)++;
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

/* One set of device buffers + stream: the handle owns one, the server one per slot. */
struct Ctx {
    hipStream_t st = nullptr;
    DevBuf in, out, blocks, ws, meta, ftab, flen, foff, run, scratch, dbg, perm, tok, tmp, ows, sws;
    uint64_t gen = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    struct { const void *src; void *out; uint32_t nblocks, depth, grid, chunk, sb, rnt, s3, s4; uint64_t gen; } gkey{};
    std::vector<hipEvent_t> ev;
    void release() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        for (DevBuf *b : {&in, &out, &blocks, &ws, &meta, &ftab, &flen, &foff, &run, &scratch, &dbg, &perm, &tok, &tmp, &ows,
                          &sws})
            b->release();
        for (auto e : ev) (void)hipEventDestroy(e);
        ev.clear();
        if (st) (void)hipStreamDestroy(st);
        st = nullptr;
    }
};

__global__ void przm_zero_kernel(uint64_t *p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

/* device-to-device copy: 16-byte vectors when both ends are 16-byte aligned, bytes otherwise */
__global__ void przm_copy_kernel(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, uint64_t n) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
    if ((((uintptr_t)src | (uintptr_t)dst) & 15u) == 0) {
        const uint64_t n16 = n >> 4;
        for (uint64_t i = t; i < n16; i += st) ((uint4 *)dst)[i] = ((const uint4 *)src)[i];
        for (uint64_t i = (n16 << 4) + t; i < n; i += st) dst[i] = src[i];
    } else {
        for (uint64_t i = t; i < n; i += st) dst[i] = src[i];
    }
}

/* desc[p] = (len << 16) | dist from the LZ kernel's dist output */
__global__ void przm_desc_kernel(const uint8_t *__restrict__ src, const przm_block_t *__restrict__ blocks,
                                 uint32_t nblocks, const uint16_t *__restrict__ dist, uint32_t *__restrict__ desc,
                                 uint64_t desc_base) {
    const uint32_t blk = blockIdx.y;
    if (blk >= nblocks) return;
    const przm_block_t bd = blocks[blk];
    const uint8_t *g = src + bd.src_off;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < bd.len; p += gridDim.x * blockDim.x) {
        const uint32_t d = dist[(uint64_t)blk * PRZM_BLOCK + p];
        uint32_t v = 0;
        if (d) {
            const uint32_t cap = bd.len - p < PRZM_MAX_MATCH ? bd.len - p : PRZM_MAX_MATCH;
            uint32_t l = 0;
            while (l < cap && g[p + l] == g[p - d + l]) l++;
            v = (l << 16) | d;
        }
        desc[(uint64_t)blk * PRZM_BLOCK + p - desc_base] = v;
    }
}

/*
 * Scratch of the entry point that allocates nothing (przm_gpu_encode_device_noalloc): every
 * device buffer the stage kernels need for at most NA_BLOCKS blocks on at most NA_GRID match
 * workgroups, carved out of one static device array.  Host code never passes the array's name to
 * a kernel or a copy: it uses the device address hipGetSymbolAddress returns for the current
 * device (na_scratch_ptr), looked up once per device when a handle is created.
 */
constexpr uint32_t NA_BLOCKS = 16; /* 1 MiB of input */
constexpr uint32_t NA_GRID = 8;
constexpr uint64_t na_al(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t NA_OFF_BLOCKS = 0;
constexpr uint64_t NA_OFF_META = NA_OFF_BLOCKS + na_al(NA_BLOCKS * sizeof(przm_block_t));
constexpr uint64_t NA_OFF_FTAB = NA_OFF_META + na_al(NA_BLOCKS * sizeof(przm_meta_t));
constexpr uint64_t NA_OFF_PERM = NA_OFF_FTAB + na_al(NA_BLOCKS * 4ull * sizeof(przm_ftab_t));
constexpr uint64_t NA_OFF_FLEN = NA_OFF_PERM + na_al(NA_BLOCKS * (PRZM_J_N + 16) * 4ull + PRZM_SHEAF_BINS * 4ull + 4ull);
constexpr uint64_t NA_OFF_FOFF = NA_OFF_FLEN + na_al(NA_BLOCKS * 4ull);
constexpr uint64_t NA_OFF_RUN = NA_OFF_FOFF + na_al(NA_BLOCKS * 8ull);
constexpr uint64_t NA_OFF_WS = NA_OFF_RUN + 256;
constexpr uint64_t NA_OFF_LZ = NA_OFF_WS + na_al((uint64_t)NA_BLOCKS * PRZM_WS_BYTES);
constexpr uint64_t NA_OFF_OWS = NA_OFF_LZ + (uint64_t)NA_GRID * LZ_SCRATCH_STRIDE;
constexpr uint64_t NA_OFF_SWS = NA_OFF_OWS + (uint64_t)NA_BLOCKS * PRZM_OWS_BYTES;
constexpr uint64_t NA_BYTES = NA_OFF_SWS + (uint64_t)NA_BLOCKS * PRZM_SWS_BYTES;

__device__ __align__(256) uint8_t scratch[NA_BYTES];

std::mutex g_na_mu;
void *g_na_ptr[64]; /* device address of `scratch` per device (null until looked up) */

/* looks the array up on device `dev` (which is current) once; 0 or an error code */
int32_t na_scratch_lookup(int dev) {
    if (dev < 0 || dev >= 64) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[dev]) return 0;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        snprintf(g_err, sizeof g_err, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        return PRZM_GPU_ERR_HIP;
    }
    g_na_ptr[dev] = ptr;
    return 0;
}

int depth_of_level(int level) { /* compress.go:34-51, default level 6 (compress.go:1038-1040) */
    if (level <= 0) level = 6;
    if (level <= 1) return 2;
    if (level <= 3) return 4;
    if (level <= 5) return 8;
    if (level <= 7) return 16;
    if (level <= 9) return 32;
    if (level <= 10) return 64;
    return 128;
}

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct przm_pinned_pool;

struct przm_gpu {
    int dev = 0;
    przm_gpu_info info{};
    int cus = 0;
    int lz_grid = 0;
    uint32_t chunk = 8192;
    uint32_t search_batch = 4; /* chain candidates per search step: 2, 4 or 8 (which LZ kernel) */
    uint32_t rans_nt = 64;     /* coder threads per workgroup: 64 or 128 (which coder kernel) */
    bool s3_prefix = false;    /* Section 3 as PackageOffsetsSection chooses it (else flat LEB128) */
    bool s4_sheaf = false;     /* Section 4 as PackageLiteralsSection chooses it (else V-ANS or raw) */
    bool przs = false;         /* a buffer of more than one block as one PRZS stream (else bare frames) */
    bool use_graph = true;
    Ctx c;
    std::mutex mu;
    /* calls that launch work on c's buffers, counted: token streams left in c by
     * przm_gpu_tokenize_device are valid only while this has not moved on */
    uint64_t epoch = 0;
    struct { uint64_t epoch = 0, n = 0; const uint8_t *src = nullptr; uint32_t nblocks = 0; int level = 6; } tok;
    std::map<uint64_t, przm_pinned_pool *> pools; /* GetPinned's pools by power-of-two size */
    /* the two slots of the double-buffered streaming match (bridge.go:157-236 pipelineSlot):
     * device buffers and stream, pinned staging of one chunk's input and descriptors */
    struct Pipe {
        Ctx c;
        DevBuf desc;
        uint8_t *h_in = nullptr;
        uint32_t *h_desc = nullptr;
        uint64_t h_cap = 0;
        hipEvent_t done = nullptr;
        void release() {
            c.release();
            desc.release();
            if (h_in) (void)hipHostFree(h_in);
            if (h_desc) (void)hipHostFree(h_desc);
            if (done) (void)hipEventDestroy(done);
            h_in = nullptr;
            h_desc = nullptr;
            done = nullptr;
            h_cap = 0;
        }
    } pipe[2];
};

namespace {

int32_t mock_check() {
    if (g_mock_unavailable.load()) return PRZM_GPU_ERR_UNAVAILABLE;
    const int32_t e = g_mock_error.load();
    if (e) return e < 0 ? e : PRZM_GPU_ERR_MOCK;
    return 0;
}

/* The block table of the buffers.  stream: every buffer of more than one block becomes one PRZS
 * stream (its blocks are length-prefixed chunks, its first block carries the header and the chunk
 * count); a buffer of one block (or none) stays a bare frame, as the reference writes a single
 * frame for input that fits one chunk. */
void build_blocks(const przm_span *bufs, uint32_t nbufs, std::vector<przm_block_t> &blocks,
                  std::vector<uint64_t> &first, bool stream = false) {
    blocks.clear();
    first.assign(nbufs + 1, 0);
    for (uint32_t i = 0; i < nbufs; i++) {
        first[i] = blocks.size();
        const uint64_t n = bufs[i].len;
        if (n == 0) {
            blocks.push_back({bufs[i].offset, 0, PRZM_BF_RAW});
            continue;
        }
        const uint64_t cnt = (n + PRZM_BLOCK - 1) / PRZM_BLOCK;
        const uint32_t sf = stream && cnt > 1 ? PRZM_BF_CHUNK : 0u;
        for (uint64_t o = 0; o < n; o += PRZM_BLOCK) {
            const uint32_t l = (uint32_t)std::min<uint64_t>(PRZM_BLOCK, n - o);
            const uint32_t hf = sf && o == 0 ? (PRZM_BF_HEAD | (uint32_t)(cnt << PRZM_BF_COUNT_SHIFT)) : 0u;
            blocks.push_back({bufs[i].offset + o, l, sf | hf});
        }
    }
    first[nbufs] = blocks.size();
}

/* output bytes of a block: its frame's bound, plus the stream prefix and header it carries */
uint64_t block_bound(const przm_block_t &b) {
    return 8ull + b.len + ((b.flags & PRZM_BF_CHUNK) ? PRZM_PRZS_PREFIX : 0u) +
           ((b.flags & PRZM_BF_HEAD) ? PRZM_PRZS_HDR : 0u);
}

/* the stream header's chunk count is 24 bits of the flag word: 2^24 - 1 blocks (just under 1 TiB) */
int32_t check_stream_sizes(const przm_span *bufs, uint32_t nbufs) {
    for (uint32_t i = 0; i < nbufs; i++)
        if (bufs[i].len > (uint64_t)((1u << (32 - PRZM_BF_COUNT_SHIFT)) - 1u) * PRZM_BLOCK) {
            snprintf(g_err, sizeof g_err, "buffer %u: %llu bytes exceed a PRZS stream's chunk count", i,
                     (unsigned long long)bufs[i].len);
            return PRZM_GPU_ERR_CAPACITY;
        }
    return 0;
}

/* device buffers for nblocks blocks processed in chunks of `chunk`; offs / sheaf: the per-block
 * workspaces of the Section-3 prefix and Section-4 sheaf options too */
int32_t ensure_ctx(Ctx &c, uint32_t nblocks, uint32_t chunk, int grid, bool dbg, bool offs = false,
                   bool sheaf = false) {
    const uint32_t ch = std::min<uint32_t>(std::max<uint32_t>(nblocks, 1), chunk);
    CHK(c.blocks.ensure((size_t)std::max<uint32_t>(nblocks, 1) * sizeof(przm_block_t), &c.gen));
    CHK(c.ws.ensure((size_t)ch * PRZM_WS_BYTES, &c.gen));
    CHK(c.meta.ensure((size_t)std::max<uint32_t>(nblocks, 1) * sizeof(przm_meta_t), &c.gen));
    CHK(c.ftab.ensure((size_t)ch * 4 * sizeof(przm_ftab_t), &c.gen));
    CHK(c.perm.ensure((size_t)ch * (PRZM_J_N + (sheaf ? 16u : 0u)) * 4 + (sheaf ? PRZM_SHEAF_BINS * 4u + 4u : 0u), &c.gen));
    CHK(c.flen.ensure((size_t)std::max<uint32_t>(nblocks, 1) * 4, &c.gen));
    CHK(c.foff.ensure((size_t)std::max<uint32_t>(nblocks, 1) * 8, &c.gen));
    /* run: the running frame offset (u64) at 0, then one block counter (u32) per chunk at 64 */
    CHK(c.run.ensure(64 + 4ull * ((std::max<uint32_t>(nblocks, 1) + ch - 1) / ch + 2), &c.gen));
    CHK(c.scratch.ensure((size_t)grid * LZ_SCRATCH_STRIDE, &c.gen));
    if (dbg) CHK(c.dbg.ensure((size_t)std::max<uint32_t>(nblocks, 1) * PRZM_BLOCK * 2, &c.gen));
    if (offs) CHK(c.ows.ensure((size_t)ch * PRZM_OWS_BYTES, &c.gen));
    if (sheaf) CHK(c.sws.ensure((size_t)ch * PRZM_SWS_BYTES, &c.gen));
    if (!c.st) HIPCHK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    return 0;
}

/*
 * The stage launch sequence for nblocks blocks already in c.blocks.
 * stage_ev (optional): 8 events per chunk recorded between stages.
 */
int32_t enqueue(przm_gpu *g, Ctx &c, const uint8_t *d_src, uint32_t nblocks, uint32_t depth, uint8_t *d_out,
                bool check, bool stop_after_lz, bool dbg, int grid, uint32_t chunk, std::vector<hipEvent_t> *stage_ev,
                bool skip_lz = false) {
    hipStream_t st = c.st;
    przm_block_t *blocks = (przm_block_t *)c.blocks.p;
    przm_meta_t *meta = (przm_meta_t *)c.meta.p;
    uint32_t *flen = (uint32_t *)c.flen.p;
    uint64_t *foff = (uint64_t *)c.foff.p;
    uint64_t *run = (uint64_t *)c.run.p;
    size_t evi = 0;
    auto rec = [&]() -> int32_t {
        if (stage_ev) {
            if (evi >= stage_ev->size()) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                stage_ev->push_back(e);
            }
            HIPCHK(hipEventRecord((*stage_ev)[evi++], st));
        }
        return 0;
    };
    /* the Section-3 prefix and Section-4 sheaf options need their workspaces for a chunk */
    const bool offs = g->s3_prefix && !stop_after_lz, sheaf = g->s4_sheaf && !stop_after_lz;
    uint8_t *ows = offs ? (uint8_t *)c.ows.p : nullptr;
    uint8_t *sws = sheaf ? (uint8_t *)c.sws.p : nullptr;
    const size_t chb = (size_t)std::min(chunk, std::max<uint32_t>(nblocks, 1));
    if ((offs && (!ows || c.ows.cap < chb * PRZM_OWS_BYTES)) ||
        (sheaf && (!sws || c.sws.cap < chb * PRZM_SWS_BYTES || c.perm.cap < chb * (PRZM_J_N + 16) * 4 + PRZM_SHEAF_BINS * 4u + 4u))) {
        snprintf(g_err, sizeof g_err, "a section option needs its workspace (set the option before the call)");
        return PRZM_GPU_ERR_ARG;
    }
    CHK(rec());
    if (check) hipLaunchKernelGGL(przm_check_kernel, dim3(nblocks), dim3(64), 0, st, d_src, blocks, nblocks);
    const uint32_t nchunks = (nblocks + chunk - 1) / chunk;
    hipLaunchKernelGGL(przm_zero_kernel, dim3(1), dim3(64), 0, st, run, 8u + (nchunks + 2u) / 2u);
    HIPCHK(hipGetLastError());
    for (uint32_t c0 = 0; c0 < nblocks; c0 += chunk) {
        const uint32_t nb = std::min(chunk, nblocks - c0);
        const uint32_t gz = (uint32_t)std::min<int>(grid, (int)nb);
        CHK(rec());
        uint32_t *counter = (uint32_t *)((uint8_t *)c.run.p + 64) + c0 / chunk;
        if (skip_lz) {
        } else if (dbg) {
            hipLaunchKernelGGL(przm_lz_match_kernel, dim3(gz), dim3(LZ_NT), 0, st, d_src, blocks + c0, nb, depth,
                               (uint8_t *)c.ws.p, meta + c0, (uint8_t *)c.scratch.p, (uint64_t)LZ_SCRATCH_STRIDE,
                               (uint16_t *)c.dbg.p + (uint64_t)c0 * PRZM_BLOCK, counter);
#ifdef PRZM_DIAG
        } else if (g_phase_ts) {
            hipLaunchKernelGGL(przm_lz_diag_kernel, dim3(gz), dim3(LZ_NT), 0, st, d_src, blocks + c0, nb, depth,
                               (uint8_t *)c.ws.p, meta + c0, (uint8_t *)c.scratch.p, (uint64_t)LZ_SCRATCH_STRIDE,
                               g_phase_ts + (uint64_t)c0 * 32, counter);
#endif
        } else {
            auto *k = g->search_batch == 2 ? przm_lz_kernel_b2 : g->search_batch == 8 ? przm_lz_kernel_b8 : przm_lz_kernel;
            hipLaunchKernelGGL(k, dim3(gz), dim3(LZ_NT), 0, st, d_src, blocks + c0, nb, depth, (uint8_t *)c.ws.p,
                               meta + c0, (uint8_t *)c.scratch.p, (uint64_t)LZ_SCRATCH_STRIDE, counter);
        }
        HIPCHK(hipGetLastError());
        if (stop_after_lz) continue;
        CHK(rec());
        hipLaunchKernelGGL(przm_hist_kernel, dim3(nb), dim3(256), 0, st, (uint8_t *)c.ws.p, meta + c0,
                           (przm_ftab_t *)c.ftab.p, nb);
        CHK(rec());
        hipLaunchKernelGGL(przm_rans_order_kernel, dim3(PRZM_J_N), dim3(1024), 0, st, (const przm_meta_t *)(meta + c0), nb,
                           (uint32_t *)c.perm.p);
        const uint32_t jobs = g->rans_nt / 4;
        hipLaunchKernelGGL(g->rans_nt == 128 ? przm_rans_kernel_128 : przm_rans_kernel,
                           dim3(PRZM_J_N * ((nb + jobs - 1) / jobs)), dim3(g->rans_nt), 0, st,
                           (uint8_t *)c.ws.p, meta + c0, (const przm_ftab_t *)c.ftab.p, nb, (const uint32_t *)c.perm.p);
        if (offs || sheaf) CHK(rec());
        if (offs) { /* Section 3 as a distance-prefix container: codes, extra bits, code streams */
            hipLaunchKernelGGL(przm_offs_prep_kernel, dim3(nb), dim3(OFFS_NT), 0, st, (const uint8_t *)c.ws.p,
                               (const przm_meta_t *)(meta + c0), ows, nb);
            const uint32_t m12_groups = (nb + 63) / 64;
            hipLaunchKernelGGL(przm_offs_code_kernel, dim3(m12_groups + (nb + 15) / 16), dim3(64), 0, st, ows, nb,
                               m12_groups);
        }
        if (sheaf) { /* Section 4 with the sheaf trials: fibers, their coded lengths, the choice */
            uint32_t *sperm = (uint32_t *)c.perm.p + (size_t)PRZM_J_N * nb;
            hipLaunchKernelGGL(przm_sheaf_split_kernel, dim3(nb), dim3(SHEAF_NT), 0, st, (const uint8_t *)c.ws.p,
                               (const przm_meta_t *)(meta + c0), sws, nb);
            uint32_t *sbins = (uint32_t *)c.perm.p + (((size_t)(PRZM_J_N + 16) * nb + 1u) & ~(size_t)1);
            hipLaunchKernelGGL(przm_zero_kernel, dim3(1), dim3(256), 0, st, (uint64_t *)sbins, PRZM_SHEAF_BINS / 2u);
            hipLaunchKernelGGL(przm_sheaf_class_kernel, dim3((nb + 15) / 16), dim3(256), 0, st, (const uint8_t *)sws,
                               nb, sbins);
            hipLaunchKernelGGL(przm_sheaf_place_kernel, dim3((nb + 15) / 16), dim3(256), 0, st, (const uint8_t *)sws,
                               nb, sbins, sperm);
            hipLaunchKernelGGL(przm_sheaf_count_kernel, dim3(nb), dim3(64), 0, st, sws, nb, (const uint32_t *)sperm);
            hipLaunchKernelGGL(przm_sheaf_encode_kernel, dim3((nb + 3) / 4), dim3(64), 0, st,
                               (const przm_meta_t *)(meta + c0), sws, nb);
        }
        CHK(rec());
        hipLaunchKernelGGL(przm_frame_size_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, blocks + c0, meta + c0,
                           (const uint8_t *)c.ws.p, nb, flen + c0, (const uint8_t *)ows, (const uint8_t *)sws);
        CHK(rec());
        hipLaunchKernelGGL(przm_scan_kernel, dim3(1), dim3(1024), 0, st, flen + c0, nb, foff + c0, run);
        CHK(rec());
        hipLaunchKernelGGL(przm_frame_write_kernel, dim3(nb), dim3(256), 0, st, d_src, blocks + c0, meta + c0,
                           (const uint8_t *)c.ws.p, (const przm_ftab_t *)c.ftab.p, foff + c0, d_out, nb,
                           (const uint8_t *)ows, (const uint8_t *)sws);
        HIPCHK(hipGetLastError());
    }
    CHK(rec());
    if (stage_ev && stage_ev->size() > evi) { /* a longer earlier sequence left spare events */
        for (size_t k = evi; k < stage_ev->size(); k++) (void)hipEventDestroy((*stage_ev)[k]);
        stage_ev->resize(evi);
    }
    return 0;
}

struct BatchOut {
    std::vector<uint32_t> flen;
    uint64_t total = 0;
};

/*
 * Encode the buffers `bufs` of the device arena d_src into d_out.  With
 * use_graph the whole launch sequence is captured once per (arena, output,
 * block count, depth, shape) and replayed.
 */
int32_t run_batch(przm_gpu *g, Ctx &c, const uint8_t *d_src, const przm_span *bufs, uint32_t nbufs, int level,
                  uint8_t *d_out, uint64_t out_cap, bool use_graph, std::vector<przm_block_t> &blocks,
                  std::vector<uint64_t> &first, BatchOut &bo, przm_gpu_profile *prof) {
    g->epoch++;
    if (g->przs) CHK(check_stream_sizes(bufs, nbufs));
    build_blocks(bufs, nbufs, blocks, first, g->przs);
    const uint32_t nblocks = (uint32_t)blocks.size();
    uint64_t bound = 0;
    for (auto &b : blocks) bound += block_bound(b);
    if (bound > out_cap) {
        snprintf(g_err, sizeof g_err, "output capacity %llu < bound %llu", (unsigned long long)out_cap,
                 (unsigned long long)bound);
        return PRZM_GPU_ERR_CAPACITY;
    }
    const uint32_t depth = (uint32_t)depth_of_level(level);
    const int grid = g->lz_grid;
    const uint32_t chunk = g->chunk;
    CHK(ensure_ctx(c, nblocks, chunk, grid, false, g->s3_prefix, g->s4_sheaf));
    HIPCHK(hipMemcpyAsync(c.blocks.p, blocks.data(), blocks.size() * sizeof(przm_block_t), hipMemcpyHostToDevice, c.st));
    if (c.ev.size() < 2) {
        for (int i = 0; i < 2; i++) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            c.ev.push_back(e);
        }
    }
    std::vector<hipEvent_t> stage_ev;
    const bool want_stages = prof && !use_graph;
    const bool want_bytes = prof && !use_graph; /* stream byte counts need a meta read-back */
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    if (use_graph) {
        const bool hit = c.gexec && c.gkey.src == d_src && c.gkey.out == d_out && c.gkey.nblocks == nblocks &&
                         c.gkey.depth == depth && c.gkey.grid == (uint32_t)grid && c.gkey.chunk == chunk &&
                         c.gkey.sb == g->search_batch && c.gkey.rnt == g->rans_nt &&
                         c.gkey.s3 == (uint32_t)g->s3_prefix && c.gkey.s4 == (uint32_t)g->s4_sheaf &&
                         c.gkey.gen == c.gen;
        if (!hit) {
            if (c.gexec) { (void)hipGraphExecDestroy(c.gexec); c.gexec = nullptr; }
            if (c.graph) { (void)hipGraphDestroy(c.graph); c.graph = nullptr; }
            HIPCHK(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
            const int32_t r = enqueue(g, c, d_src, nblocks, depth, d_out, true, false, false, grid, chunk, nullptr);
            hipGraph_t gr = nullptr;
            const hipError_t e2 = hipStreamEndCapture(c.st, &gr);
            if (r) { if (gr) (void)hipGraphDestroy(gr); return r; }
            HIPCHK(e2);
            c.graph = gr;
            HIPCHK(hipGraphInstantiate(&c.gexec, c.graph, nullptr, nullptr, 0));
            c.gkey = {d_src, d_out, nblocks, depth, (uint32_t)grid, chunk, g->search_batch, g->rans_nt,
                      (uint32_t)g->s3_prefix, (uint32_t)g->s4_sheaf, c.gen};
        }
        HIPCHK(hipGraphLaunch(c.gexec, c.st));
    } else {
        CHK(enqueue(g, c, d_src, nblocks, depth, d_out, true, false, false, grid, chunk, want_stages ? &stage_ev : nullptr));
    }
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    bo.flen.resize(nblocks);
    HIPCHK(hipMemcpyAsync(bo.flen.data(), c.flen.p, (size_t)nblocks * 4, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    bo.total = 0;
    for (uint32_t v : bo.flen) bo.total += v;
    if (prof) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
        prof->kernel_ms = ms;
        prof->graph_launch = use_graph ? 1 : 0;
        prof->pipelined_chunks = (int32_t)((nblocks + chunk - 1) / chunk);
        uint64_t in_bytes = 0;
        for (auto &b : blocks) in_bytes += b.len;
        prof->internal_vram_gbps = ms > 0 ? (double)(in_bytes + bo.total) / (ms * 1e6) : 0;
        for (int s = 0; s < 8; s++) prof->stage_ms[s] = 0;
        /* per-stage times need events between the launches: un-graphed runs only (the events an
         * event-record node of a captured graph records are not accepted by hipEventElapsedTime) */
        const std::vector<hipEvent_t> &sev = stage_ev;
        if (want_stages && sev.size() >= 2) {
            /* layout: [start][check+zero] then per chunk [lz][hist][rans]([options])[size][scan][write]
             * ... [end]; the kernels of the section options (prefix offsets, sheaf) are stage 7 */
            static const int flat_map[6] = {1, 2, 3, 4, 5, 6}, pre_map[7] = {1, 2, 3, 7, 4, 5, 6};
            const bool opt = g->s3_prefix || g->s4_sheaf;
            const int per = opt ? 7 : 6;
            const int *map = opt ? pre_map : flat_map;
            float t;
            HIPCHK(hipEventElapsedTime(&t, sev[0], sev[1]));
            prof->stage_ms[0] += t;
            size_t i = 1;
            while (i + per < sev.size()) {
                for (int s = 0; s < per; s++) {
                    HIPCHK(hipEventElapsedTime(&t, sev[i + s], sev[i + s + 1]));
                    prof->stage_ms[map[s]] += t;
                }
                i += per;
            }
        }
        for (auto e : stage_ev) (void)hipEventDestroy(e);
        prof->frame_bytes = bo.total;
        if (want_bytes) {
            std::vector<przm_meta_t> mt(nblocks);
            HIPCHK(hipMemcpy(mt.data(), c.meta.p, (size_t)nblocks * sizeof(przm_meta_t), hipMemcpyDeviceToHost));
            uint64_t lin = 0, lst = 0, raw = 0;
            for (uint32_t b = 0; b < nblocks; b++) {
                if (mt[b].frame_kind != 0) raw++;
                const uint32_t s6 = mt[b].len[0] + mt[b].len[1] + mt[b].len[2] + mt[b].len[3] + mt[b].len[4] + mt[b].len[5];
                if (s6) { lin += blocks[b].len; lst += s6; }
            }
            prof->lz_in_bytes = lin;
            prof->lz_stream_bytes = lst;
            prof->raw_frames = raw;
        }
    }
    return 0;
}

void fill_outputs(const przm_span *bufs, uint32_t nbufs, const std::vector<uint64_t> &first, const BatchOut &bo,
                  uint64_t *out_off, uint32_t *frame_sizes, uint64_t *nframes) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < nbufs; i++) {
        if (out_off) out_off[i] = acc;
        for (uint64_t b = first[i]; b < first[i + 1]; b++) acc += bo.flen[b];
    }
    if (out_off) out_off[nbufs] = acc;
    if (frame_sizes) std::memcpy(frame_sizes, bo.flen.data(), bo.flen.size() * 4);
    if (nframes) *nframes = bo.flen.size();
    (void)bufs;
}

int32_t init_handle(przm_gpu *g, int device) {
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return PRZM_GPU_ERR_UNAVAILABLE;
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    g->dev = device;
    g->cus = p.multiProcessorCount;
    g->lz_grid = p.multiProcessorCount;
    g->info.device = device;
    g->info.compute_major = p.major;
    g->info.compute_minor = p.minor;
    g->info.sm_count = p.multiProcessorCount;
    g->info.total_vram_bytes = p.totalGlobalMem;
    snprintf(g->info.device_name, sizeof g->info.device_name, "%s", p.name);
    snprintf(g->info.arch, sizeof g->info.arch, "%s", p.gcnArchName);
    HIPCHK(hipStreamCreateWithFlags(&g->c.st, hipStreamNonBlocking));
    CHK(na_scratch_lookup(device));
    return 0;
}

}  // namespace

/* ======================= C ABI ======================= */

extern "C" {

const char *przm_gpu_last_error(void) { return g_err; }

#ifdef PRZM_DIAG
/* Diagnostics builds only (not part of the library's interface): when set, the LZ kernel writes
 * up to 32 s_memrealtime stamps (100 MHz) per block into this device array (32 x blocks u64). */
extern "C" void przm_gpu_debug_phase_ts(void *d) { g_phase_ts = (uint64_t *)d; }
#endif

void przm_gpu_set_mock_unavailable(int32_t on) { g_mock_unavailable.store(on ? 1 : 0); }
void przm_gpu_set_mock_error(int32_t code) { g_mock_error.store(code); }

int32_t przm_gpu_probe(przm_gpu_info *info) {
    if (g_mock_unavailable.load()) return PRZM_GPU_ERR_UNAVAILABLE;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return PRZM_GPU_ERR_UNAVAILABLE;
    for (int d = 0; d < count; d++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) != hipSuccess) continue;
        if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) continue; /* kernels are built for gfx950 only */
        if (info) {
            info->device = d;
            info->compute_major = p.major;
            info->compute_minor = p.minor;
            info->sm_count = p.multiProcessorCount;
            info->total_vram_bytes = p.totalGlobalMem;
            snprintf(info->device_name, sizeof info->device_name, "%s", p.name);
            snprintf(info->arch, sizeof info->arch, "%s", p.gcnArchName);
        }
        return 0;
    }
    return PRZM_GPU_ERR_UNAVAILABLE;
}

int32_t przm_gpu_create(int32_t device, przm_gpu **out) {
    if (!out) return PRZM_GPU_ERR_ARG;
    *out = nullptr;
    if (g_mock_unavailable.load()) return PRZM_GPU_ERR_UNAVAILABLE;
    przm_gpu *g = new (std::nothrow) przm_gpu();
    if (!g) return PRZM_GPU_ERR_OOM;
    const int32_t r = init_handle(g, device);
    if (r) {
        delete g;
        return r;
    }
    *out = g;
    return 0;
}

int32_t przm_gpu_destroy(przm_gpu *g) {
    if (!g) return 0;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        (void)hipSetDevice(g->dev);
        (void)hipStreamSynchronize(g->c.st);
        g->c.release();
        for (auto &kv : g->pools) (void)przm_pinned_pool_destroy(kv.second);
        g->pools.clear();
        for (auto &pp : g->pipe) {
            if (pp.c.st) (void)hipStreamSynchronize(pp.c.st);
            pp.release();
        }
    }
    delete g;
    return 0;
}

int32_t przm_gpu_get_info(przm_gpu *g, przm_gpu_info *info) {
    if (!g || !info) return PRZM_GPU_ERR_ARG;
    *info = g->info;
    return 0;
}

uint64_t przm_gpu_frames_bound(uint64_t n) {
    if (n == 0) return 8;
    return n + 8 * ((n + PRZM_BLOCK - 1) / PRZM_BLOCK);
}

uint64_t przm_gpu_stream_bound(uint64_t n) {
    if (n <= PRZM_BLOCK) return przm_gpu_frames_bound(n);
    return przm_gpu_frames_bound(n) + PRZM_PRZS_HDR + PRZM_PRZS_PREFIX * ((n + PRZM_BLOCK - 1) / PRZM_BLOCK);
}

int32_t przm_gpu_set_container(przm_gpu *g, int32_t mode) {
    if (!g || (mode != PRZM_CONTAINER_FRAMES && mode != PRZM_CONTAINER_STREAM)) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    g->przs = mode == PRZM_CONTAINER_STREAM;
    return 0;
}

int32_t przm_gpu_set_tuning(przm_gpu *g, int32_t lz_grid, int32_t chunk_blocks) {
    if (!g) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    if (lz_grid > 0) g->lz_grid = lz_grid;
    if (chunk_blocks > 0) g->chunk = (uint32_t)chunk_blocks;
    return 0;
}

int32_t przm_gpu_set_kernel_shape(przm_gpu *g, int32_t search_batch, int32_t rans_jobs) {
    if (!g) return PRZM_GPU_ERR_ARG;
    if (search_batch && search_batch != 2 && search_batch != 4 && search_batch != 8) return PRZM_GPU_ERR_ARG;
    if (rans_jobs && rans_jobs != 16 && rans_jobs != 32) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    if (search_batch) g->search_batch = (uint32_t)search_batch;
    if (rans_jobs) g->rans_nt = (uint32_t)rans_jobs * 4;
    return 0;
}

int32_t przm_gpu_set_offsets_mode(przm_gpu *g, int32_t mode) {
    if (!g || (mode != PRZM_OFFSETS_FLAT && mode != PRZM_OFFSETS_PREFIX)) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    g->s3_prefix = mode == PRZM_OFFSETS_PREFIX;
    return 0;
}

int32_t przm_gpu_set_literals_mode(przm_gpu *g, int32_t mode) {
    if (!g || (mode != PRZM_LITERALS_VANS && mode != PRZM_LITERALS_SHEAF)) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    g->s4_sheaf = mode == PRZM_LITERALS_SHEAF;
    return 0;
}

int32_t przm_gpu_get_tuning(przm_gpu *g, int32_t tuned[4]) {
    if (!g || !tuned) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    tuned[0] = g->lz_grid;
    tuned[1] = (int32_t)g->chunk;
    tuned[2] = (int32_t)(g->rans_nt / 4);
    tuned[3] = (int32_t)g->search_batch;
    return 0;
}

int32_t przm_gpu_set_options(przm_gpu *g, int32_t use_graph) {
    if (!g) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    g->use_graph = use_graph != 0;
    return 0;
}

int32_t przm_gpu_malloc(przm_gpu *g, uint64_t bytes, void **d) {
    if (!g || !d) return PRZM_GPU_ERR_ARG;
    HIPCHK(hipSetDevice(g->dev));
    if (hipMalloc(d, bytes ? bytes : 1) != hipSuccess) return PRZM_GPU_ERR_OOM;
    return 0;
}
int32_t przm_gpu_free(przm_gpu *g, void *d) {
    if (!g) return PRZM_GPU_ERR_ARG;
    if (d) HIPCHK(hipFree(d));
    return 0;
}
int32_t przm_gpu_h2d(przm_gpu *g, void *d, const void *h, uint64_t bytes) {
    if (!g) return PRZM_GPU_ERR_ARG;
    HIPCHK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    return 0;
}
int32_t przm_gpu_d2h(przm_gpu *g, void *h, const void *d, uint64_t bytes) {
    if (!g) return PRZM_GPU_ERR_ARG;
    HIPCHK(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
    return 0;
}
int32_t przm_gpu_copy_d2d(przm_gpu *g, void *dst, const void *src, uint64_t bytes) {
    if (!g || ((!dst || !src) && bytes)) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    if (!bytes) return 0;
    if (!g->c.st) HIPCHK(hipStreamCreateWithFlags(&g->c.st, hipStreamNonBlocking));
    const uint64_t n16 = bytes >> 4;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n16 + 255) / 256 + 1, (uint64_t)g->cus * 8);
    hipLaunchKernelGGL(przm_copy_kernel, dim3(grid), dim3(256), 0, g->c.st, (const uint8_t *)src, (uint8_t *)dst, bytes);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g->c.st));
    return 0;
}
int32_t przm_gpu_sync(przm_gpu *g) {
    if (!g) return PRZM_GPU_ERR_ARG;
    HIPCHK(hipDeviceSynchronize());
    return 0;
}
int32_t przm_pinned_alloc(uint64_t bytes, void **p) {
    if (!p) return PRZM_GPU_ERR_ARG;
    HIPCHK(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault));
    return 0;
}
int32_t przm_pinned_free(void *p) {
    if (p) HIPCHK(hipHostFree(p));
    return 0;
}

/* ---- encode ---- */

int32_t przm_gpu_encode_batch_device(przm_gpu *g, const void *d_src, const przm_span *bufs, uint32_t nbufs,
                                     int32_t level, void *d_out, uint64_t out_cap, uint64_t *out_off,
                                     uint32_t *frame_sizes, uint64_t frames_cap, uint64_t *nframes,
                                     przm_gpu_profile *prof) {
    if (!g || (!bufs && nbufs)) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    std::vector<przm_block_t> blocks;
    std::vector<uint64_t> first;
    uint64_t nb = 0;
    for (uint32_t i = 0; i < nbufs; i++) nb += bufs[i].len ? (bufs[i].len + PRZM_BLOCK - 1) / PRZM_BLOCK : 1;
    if (frame_sizes && nb > frames_cap) return PRZM_GPU_ERR_CAPACITY;
    BatchOut bo;
    const double t0 = now_ms();
    CHK(run_batch(g, g->c, (const uint8_t *)d_src, bufs, nbufs, level, (uint8_t *)d_out, out_cap,
                  g->use_graph && nbufs > 1, blocks, first, bo, prof));
    fill_outputs(bufs, nbufs, first, bo, out_off, frame_sizes, nframes);
    if (prof) {
        prof->total_round_trip_ms = now_ms() - t0;
        uint64_t in = 0;
        for (uint32_t i = 0; i < nbufs; i++) in += bufs[i].len;
        prof->end_to_end_gbps = prof->total_round_trip_ms > 0 ? in / (prof->total_round_trip_ms * 1e6) : 0;
    }
    return 0;
}

int32_t przm_gpu_encode_device(przm_gpu *g, const void *d_src, uint64_t n, int32_t level, void *d_out,
                               uint64_t out_cap, uint64_t *out_len, uint32_t *frame_sizes, uint64_t frames_cap,
                               uint64_t *nframes, przm_gpu_profile *prof) {
    if (!g) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    przm_span sp{0, n};
    const uint64_t nb = n ? (n + PRZM_BLOCK - 1) / PRZM_BLOCK : 1;
    if (frame_sizes && nb > frames_cap) return PRZM_GPU_ERR_CAPACITY;
    std::vector<przm_block_t> blocks;
    std::vector<uint64_t> first;
    BatchOut bo;
    const double t0 = now_ms();
    CHK(run_batch(g, g->c, (const uint8_t *)d_src, &sp, 1, level, (uint8_t *)d_out, out_cap, false, blocks, first, bo,
                  prof));
    fill_outputs(&sp, 1, first, bo, nullptr, frame_sizes, nframes);
    if (out_len) *out_len = bo.total;
    if (prof) prof->total_round_trip_ms = now_ms() - t0;
    return 0;
}

int32_t przm_gpu_encode_device_noalloc(przm_gpu *g, const void *d_src, uint64_t n, int32_t level, void *d_out,
                                       uint64_t out_cap, uint64_t *out_len, uint32_t *frame_sizes, uint64_t frames_cap,
                                       uint64_t *nframes) {
    if (!g || (!d_src && n) || !d_out || !frame_sizes) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    if (n > (uint64_t)NA_BLOCKS * PRZM_BLOCK) {
        snprintf(g_err, sizeof g_err, "przm_gpu_encode_device_noalloc: %llu bytes above its maximum of %llu",
                 (unsigned long long)n, (unsigned long long)NA_BLOCKS * PRZM_BLOCK);
        return PRZM_GPU_ERR_CAPACITY;
    }
    const uint32_t nb = n ? (uint32_t)((n + PRZM_BLOCK - 1) / PRZM_BLOCK) : 1u;
    std::lock_guard<std::mutex> lk(g->mu);
    const bool stream = g->przs && nb > 1;
    if (nb > frames_cap || (stream ? przm_gpu_stream_bound(n) : przm_gpu_frames_bound(n)) > out_cap)
        return PRZM_GPU_ERR_CAPACITY;
    HIPCHK(hipSetDevice(g->dev));
    uint8_t *base = (uint8_t *)g_na_ptr[g->dev];
    if (!base) return PRZM_GPU_ERR_UNAVAILABLE;
    /* the block table on the stack, uploaded into the array (no device allocation) */
    przm_block_t blocks[NA_BLOCKS];
    if (n == 0) blocks[0] = {0, 0, PRZM_BF_RAW};
    for (uint32_t b = 0; b < nb && n; b++) {
        const uint32_t sf = stream ? (PRZM_BF_CHUNK | (b == 0 ? PRZM_BF_HEAD | (nb << PRZM_BF_COUNT_SHIFT) : 0u)) : 0u;
        blocks[b] = {(uint64_t)b * PRZM_BLOCK, (uint32_t)std::min<uint64_t>(PRZM_BLOCK, n - (uint64_t)b * PRZM_BLOCK),
                     sf};
    }
    /* a context whose buffers are views into the array: never ensure()d or release()d */
    Ctx v;
    v.st = g->c.st;
    v.blocks.p = base + NA_OFF_BLOCKS;
    v.meta.p = base + NA_OFF_META;
    v.ftab.p = base + NA_OFF_FTAB;
    v.perm.p = base + NA_OFF_PERM;
    v.perm.cap = (size_t)NA_BLOCKS * (PRZM_J_N + 16) * 4 + PRZM_SHEAF_BINS * 4u + 4u;
    v.flen.p = base + NA_OFF_FLEN;
    v.foff.p = base + NA_OFF_FOFF;
    v.run.p = base + NA_OFF_RUN;
    v.ws.p = base + NA_OFF_WS;
    v.scratch.p = base + NA_OFF_LZ;
    v.ows.p = base + NA_OFF_OWS;
    v.ows.cap = (size_t)NA_BLOCKS * PRZM_OWS_BYTES;
    v.sws.p = base + NA_OFF_SWS;
    v.sws.cap = (size_t)NA_BLOCKS * PRZM_SWS_BYTES;
    HIPCHK(hipMemcpyAsync(v.blocks.p, blocks, nb * sizeof(przm_block_t), hipMemcpyHostToDevice, v.st));
    const int grid = std::min<int>(std::min<int>(g->lz_grid, (int)NA_GRID), (int)nb);
    /* every counter the kernels start from zero is cleared by a fill kernel inside enqueue */
    CHK(enqueue(g, v, (const uint8_t *)d_src, nb, (uint32_t)depth_of_level(level), (uint8_t *)d_out, true, false, false,
                grid, nb, nullptr));
    HIPCHK(hipMemcpyAsync(frame_sizes, v.flen.p, nb * 4ull, hipMemcpyDeviceToHost, v.st));
    HIPCHK(hipStreamSynchronize(v.st));
    uint64_t total = 0;
    for (uint32_t b = 0; b < nb; b++) total += frame_sizes[b];
    if (out_len) *out_len = total;
    if (nframes) *nframes = nb;
    return 0;
}

int32_t przm_gpu_encode_batch(przm_gpu *g, const uint8_t *src, const przm_span *bufs, uint32_t nbufs, int32_t level,
                              uint8_t *out, uint64_t out_cap, uint64_t *out_off, uint32_t *frame_sizes,
                              uint64_t frames_cap, uint64_t *nframes, przm_gpu_profile *prof) {
    if (!g || (!bufs && nbufs)) return PRZM_GPU_ERR_ARG;
    for (uint32_t i = 0; i < nbufs; i++)
        if (bufs[i].len && !src) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    const double t0 = now_ms();
    /* pack buffers 16-byte aligned in the device arena */
    std::vector<przm_span> dev(nbufs);
    uint64_t arena = 0, in_bytes = 0, nb = 0;
    for (uint32_t i = 0; i < nbufs; i++) {
        dev[i] = {arena, bufs[i].len};
        arena += (bufs[i].len + 255) & ~255ull;
        in_bytes += bufs[i].len;
        nb += bufs[i].len ? (bufs[i].len + PRZM_BLOCK - 1) / PRZM_BLOCK : 1;
    }
    if (frame_sizes && nb > frames_cap) return PRZM_GPU_ERR_CAPACITY;
    uint64_t bound = in_bytes + 8 * nb;
    if (g->przs) /* chunk prefixes and stream headers */
        for (uint32_t i = 0; i < nbufs; i++)
            if (bufs[i].len > PRZM_BLOCK)
                bound += PRZM_PRZS_HDR + PRZM_PRZS_PREFIX * ((bufs[i].len + PRZM_BLOCK - 1) / PRZM_BLOCK);
    if (bound > out_cap && out) {
        /* the exact size is known only afterwards; require the bound like the device form */
    }
    Ctx &c = g->c;
    CHK(c.in.ensure(arena + 64, &c.gen));
    CHK(c.out.ensure(bound + 64, &c.gen));
    const double h0 = now_ms();
    for (uint32_t i = 0; i < nbufs; i++)
        if (bufs[i].len)
            HIPCHK(hipMemcpyAsync((uint8_t *)c.in.p + dev[i].offset, src + bufs[i].offset, bufs[i].len,
                                  hipMemcpyHostToDevice, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    const double h1 = now_ms();
    std::vector<przm_block_t> blocks;
    std::vector<uint64_t> first;
    BatchOut bo;
    CHK(run_batch(g, c, (const uint8_t *)c.in.p, dev.data(), nbufs, level, (uint8_t *)c.out.p, bound + 64, nbufs > 1,
                  blocks, first, bo, prof));
    if (bo.total > out_cap) return PRZM_GPU_ERR_CAPACITY;
    const double d0 = now_ms();
    HIPCHK(hipMemcpy(out, c.out.p, bo.total, hipMemcpyDeviceToHost));
    const double d1 = now_ms();
    fill_outputs(bufs, nbufs, first, bo, out_off, frame_sizes, nframes);
    if (prof) {
        prof->ingress_dma_bytes = in_bytes;
        prof->ingress_dma_ms = h1 - h0;
        prof->ingress_dma_gbps = (h1 - h0) > 0 ? in_bytes / ((h1 - h0) * 1e6) : 0;
        prof->egress_dma_bytes = bo.total;
        prof->egress_dma_ms = d1 - d0;
        prof->egress_dma_gbps = (d1 - d0) > 0 ? bo.total / ((d1 - d0) * 1e6) : 0;
        prof->total_round_trip_ms = now_ms() - t0;
        prof->end_to_end_gbps = prof->total_round_trip_ms > 0 ? in_bytes / (prof->total_round_trip_ms * 1e6) : 0;
        const double dma = (h1 - h0) + (d1 - d0);
        prof->pcie_overhead_percent = prof->total_round_trip_ms > 0 ? 100.0 * dma / prof->total_round_trip_ms : 0;
    }
    return 0;
}

int32_t przm_gpu_encode(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level, uint8_t *out, uint64_t out_cap,
                        uint64_t *out_len, uint32_t *frame_sizes, uint64_t frames_cap, uint64_t *nframes,
                        przm_gpu_profile *prof) {
    przm_span sp{0, n};
    uint64_t off[2] = {0, 0};
    const int32_t r = przm_gpu_encode_batch(g, src, &sp, 1, level, out, out_cap, off, frame_sizes, frames_cap,
                                            nframes, prof);
    if (r == 0 && out_len) *out_len = off[1];
    return r;
}

/* ---- stages ---- */

namespace {
int32_t stage_upload(przm_gpu *g, const uint8_t *src, uint64_t n, std::vector<przm_block_t> &blocks) {
    Ctx &c = g->c;
    g->epoch++;
    przm_span sp{0, n};
    std::vector<uint64_t> first;
    build_blocks(&sp, 1, blocks, first);
    for (auto &b : blocks) b.flags = 0; /* stage APIs tokenize every block, high-entropy or not */
    if (n == 0) blocks.clear();
    const uint32_t nb = (uint32_t)blocks.size();
    CHK(c.in.ensure(n + 64, &c.gen));
    CHK(ensure_ctx(c, nb ? nb : 1, std::max<uint32_t>(nb, 1), g->lz_grid, true));
    if (n) HIPCHK(hipMemcpyAsync(c.in.p, src, n, hipMemcpyHostToDevice, c.st));
    if (nb) HIPCHK(hipMemcpyAsync(c.blocks.p, blocks.data(), nb * sizeof(przm_block_t), hipMemcpyHostToDevice, c.st));
    return 0;
}
}  // namespace

int32_t przm_gpu_match(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level, uint32_t *desc,
                       przm_gpu_profile *prof) {
    if (!g || (!src && n) || (!desc && n)) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    if (n == 0) return 0;
    std::vector<przm_block_t> blocks;
    CHK(stage_upload(g, src, n, blocks));
    Ctx &c = g->c;
    const uint32_t nb = (uint32_t)blocks.size();
    if (c.ev.size() < 2) {
        for (int i = 0; i < 2; i++) { hipEvent_t e; HIPCHK(hipEventCreate(&e)); c.ev.push_back(e); }
    }
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    CHK(enqueue(g, c, (const uint8_t *)c.in.p, nb, (uint32_t)depth_of_level(level), (uint8_t *)c.out.p, false, true,
                true, g->lz_grid, nb, nullptr));
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    DevBuf d;
    CHK(d.ensure(n * 4, nullptr));
    hipLaunchKernelGGL(przm_desc_kernel, dim3(64, nb), dim3(256), 0, c.st, (const uint8_t *)c.in.p,
                       (const przm_block_t *)c.blocks.p, nb, (const uint16_t *)c.dbg.p, (uint32_t *)d.p, 0ull);
    const hipError_t le = hipGetLastError();
    if (le == hipSuccess) (void)hipMemcpyAsync(desc, d.p, n * 4, hipMemcpyDeviceToHost, c.st);
    const hipError_t se = hipStreamSynchronize(c.st);
    d.release();
    HIPCHK(le);
    HIPCHK(se);
    if (prof) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
        prof->kernel_ms = ms;
    }
    return 0;
}

int32_t przm_gpu_tokenize(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level, uint8_t *cmd, uint64_t cmd_cap,
                          uint8_t *off, uint64_t off_cap, uint8_t *lit, uint64_t lit_cap, uint8_t *op,
                          uint64_t op_cap, uint8_t *meta, uint64_t meta_cap, uint8_t *var, uint64_t var_cap,
                          uint32_t *block_lens, uint64_t nblocks_cap, przm_gpu_profile *prof) {
    if (!g || (!src && n)) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    if (n == 0) return 0;
    std::vector<przm_block_t> blocks;
    CHK(stage_upload(g, src, n, blocks));
    Ctx &c = g->c;
    const uint32_t nb = (uint32_t)blocks.size();
    if (nb > nblocks_cap) return PRZM_GPU_ERR_CAPACITY;
    if (c.ev.size() < 2) {
        for (int i = 0; i < 2; i++) { hipEvent_t e; HIPCHK(hipEventCreate(&e)); c.ev.push_back(e); }
    }
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    CHK(enqueue(g, c, (const uint8_t *)c.in.p, nb, (uint32_t)depth_of_level(level), (uint8_t *)c.out.p, false, true,
                false, g->lz_grid, nb, nullptr));
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    std::vector<przm_meta_t> mt(nb);
    HIPCHK(hipMemcpyAsync(mt.data(), c.meta.p, nb * sizeof(przm_meta_t), hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    uint8_t *dst[6] = {cmd, off, lit, op, meta, var};
    const uint64_t cap[6] = {cmd_cap, off_cap, lit_cap, op_cap, meta_cap, var_cap};
    const uint32_t wsoff[6] = {PRZM_WS_CMD, PRZM_WS_OFF, PRZM_WS_LIT, PRZM_WS_OP, PRZM_WS_META, PRZM_WS_VAR};
    uint64_t pos[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t b = 0; b < nb; b++) {
        for (int s = 0; s < 6; s++) {
            const uint32_t l = mt[b].len[s];
            if (block_lens) block_lens[(uint64_t)b * 7 + s] = l;
            if (!dst[s] || !l) continue;
            if (pos[s] + l > cap[s]) return PRZM_GPU_ERR_CAPACITY;
            HIPCHK(hipMemcpyAsync(dst[s] + pos[s], (uint8_t *)c.ws.p + (uint64_t)b * PRZM_WS_BYTES + wsoff[s], l,
                                  hipMemcpyDeviceToHost, c.st));
            pos[s] += l;
        }
        if (block_lens) block_lens[(uint64_t)b * 7 + 6] = mt[b].len[PRZM_S_TOK];
    }
    HIPCHK(hipStreamSynchronize(c.st));
    if (prof) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
        prof->kernel_ms = ms;
    }
    return 0;
}

/* przm_gpu_histogram, przm_gpu_vans_encode and przm_gpu_pack_flat: przm_stage_api.hip */

/* ---- host helpers ---- */

int32_t przm_normalize_freq(const uint32_t counts[256], uint32_t target_m, uint16_t freq[256]) {
    if (!counts || !freq || target_m != 4096) return PRZM_GPU_ERR_ARG;
    uint64_t total = 0;
    int unique = 0;
    for (int i = 0; i < 256; i++) {
        freq[i] = 0;
        if (counts[i]) { total += counts[i]; unique++; }
    }
    if (!unique) return 0;
    const uint64_t rem = 4096 - (uint64_t)unique;
    uint32_t sum = 0;
    for (int i = 0; i < 256; i++) {
        if (!counts[i]) continue;
        freq[i] = (uint16_t)(1 + (counts[i] * rem) / total);
        sum += freq[i];
    }
    /* deficit: +1 each to the largest counts, lower symbol first on ties, each at most once */
    int order[256];
    int m = 0;
    for (int i = 0; i < 256; i++) if (counts[i]) order[m++] = i;
    std::stable_sort(order, order + m, [&](int a, int b) { return counts[a] > counts[b]; });
    for (int r = 0; r < (int)(4096 - sum) && r < m; r++) freq[order[r]]++;
    return 0;
}

int32_t przm_build_cdf(const uint16_t freq[256], uint16_t cdf[257]) {
    if (!freq || !cdf) return PRZM_GPU_ERR_ARG;
    uint32_t c = 0;
    for (int i = 0; i < 256; i++) { cdf[i] = (uint16_t)c; c += freq[i]; }
    cdf[256] = (uint16_t)c;
    return 0;
}

int32_t przm_serialize_matrix(const uint16_t freq[256], uint8_t *out, uint64_t cap, uint64_t *len) {
    if (!freq || !out || !len) return PRZM_GPU_ERR_ARG;
    uint8_t tmp[600];
    uint32_t n = 0, rows = 0;
    for (int h = 0; h < 16; h++)
        for (int l = 0; l < 16; l++)
            if (freq[(h << 4) | l]) { rows |= 1u << h; break; }
    tmp[n++] = (uint8_t)rows; tmp[n++] = (uint8_t)(rows >> 8);
    for (int h = 0; h < 16; h++) {
        if (!(rows & (1u << h))) continue;
        uint32_t cols = 0;
        for (int l = 0; l < 16; l++) if (freq[(h << 4) | l]) cols |= 1u << l;
        tmp[n++] = (uint8_t)cols; tmp[n++] = (uint8_t)(cols >> 8);
        for (int l = 0; l < 16; l++) {
            if (!(cols & (1u << l))) continue;
            const uint32_t w = freq[(h << 4) | l];
            if (w < 128) tmp[n++] = (uint8_t)w;
            else { tmp[n++] = (uint8_t)((w >> 8) | 0x80); tmp[n++] = (uint8_t)w; }
        }
    }
    if (n > cap) return PRZM_GPU_ERR_CAPACITY;
    std::memcpy(out, tmp, n);
    *len = n;
    return 0;
}

/* ---- autotuner ---- */

int32_t przm_gpu_autotune(przm_gpu *g, const void *d_src, const przm_span *bufs, uint32_t nbufs, int32_t level,
                          int32_t tuned[4]) {
    if (!g || !bufs || !nbufs) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    /* sample: buffers from the front up to ~4 GiB, so that the sample spans several launches of
     * the smaller chunk shapes: each launch ends on its longest coder job (a stream's rANS steps
     * are sequential), so the number of launches, not only their size, sets the time */
    uint32_t ns = 0;
    uint64_t bytes = 0;
    while (ns < nbufs && bytes < (4ull << 30)) bytes += bufs[ns++].len;
    uint64_t bound = 0;
    for (uint32_t i = 0; i < ns; i++) bound += przm_gpu_frames_bound(bufs[i].len);
    void *d_out = nullptr;
    CHK(przm_gpu_malloc(g, bound + 64, &d_out));
    int32_t rc = 0;
    /* best of three passes of the sample (graph replays) in the current shape */
    auto time_shape = [&](double *t) -> int32_t {
        *t = 1e30;
        for (int rep = 0; rep < 3; rep++) {
            przm_gpu_profile p{};
            CHK(przm_gpu_encode_batch_device(g, d_src, bufs, ns, level, d_out, bound + 64, nullptr, nullptr, 0,
                                             nullptr, &p));
            if (rep > 0) *t = std::min(*t, p.kernel_ms);
        }
        return 0;
    };
    /* coordinate search, one stage at a time: the blocks per launch on two workgroups per CU,
     * then one workgroup per CU, then the LZ kernel's chain candidates per step, then the coder's
     * jobs per workgroup.  The LZ kernel holds a whole CU either way (its LDS), so the two grids
     * differ only in how the blocks are handed out and measure within about 1 % of each other:
     * one workgroup per CU replaces two only when more than 1 % faster, so that the choice does
     * not flip with run-to-run noise */
    const int chunks[] = {16384, 32768, 65536}; /* 65536 blocks: 32 GB of block workspace */
    int bg = 2 * g->cus, bc = (int)g->chunk, bsb = 4, brj = 16;
    przm_gpu_set_kernel_shape(g, bsb, brj);
    double best = 1e30, t = 0;
    for (int ci : chunks) {
        przm_gpu_set_tuning(g, bg, ci);
        if ((rc = time_shape(&t))) break;
        if (t < best) { best = t; bc = ci; }
    }
    if (!rc) {
        przm_gpu_set_tuning(g, g->cus, bc);
        rc = time_shape(&t);
        if (!rc && t < 0.99 * best) { best = t; bg = g->cus; }
    }
    przm_gpu_set_tuning(g, bg, bc);
    /* past the sample: the whole batch in one launch (one coder tail instead of one per chunk),
     * when its block workspace fits in free device memory with a quarter to spare; compared with
     * the best chunk on the whole batch */
    uint64_t nb_all = 0, bound_all = 0;
    for (uint32_t i = 0; i < nbufs; i++) {
        nb_all += bufs[i].len ? (bufs[i].len + PRZM_BLOCK - 1) / PRZM_BLOCK : 1;
        bound_all += przm_gpu_frames_bound(bufs[i].len);
    }
    size_t mfree = 0, mtotal = 0;
    if (!rc && nb_all > (uint64_t)bc && nb_all < (1ull << 31) && hipMemGetInfo(&mfree, &mtotal) == hipSuccess) {
        const uint64_t held = g->c.ws.cap + g->c.ftab.cap + g->c.perm.cap;
        const uint64_t need = nb_all * ((uint64_t)PRZM_WS_BYTES + 4 * sizeof(przm_ftab_t) + PRZM_J_N * 4);
        if (need + bound_all + (mtotal >> 2) < mfree + held) {
            (void)przm_gpu_free(g, d_out);
            d_out = nullptr;
            rc = przm_gpu_malloc(g, bound_all + 64, &d_out);
            const uint32_t saved_ns = ns;
            const uint64_t saved_bound = bound;
            const double best_sample = best;
            ns = nbufs;
            bound = bound_all;
            double t_best = 1e30, t_all = 1e30;
            if (!rc) rc = time_shape(&t_best);
            if (!rc) {
                przm_gpu_set_tuning(g, bg, (int)nb_all);
                rc = time_shape(&t_all);
            }
            if (!rc && t_all < t_best) bc = (int)nb_all;
            przm_gpu_set_tuning(g, bg, bc);
            if (!rc && bc == (int)nb_all) {
                best = t_all; /* the later stages are timed on the whole batch too */
            } else {
                ns = saved_ns; /* the later stages time the sample again */
                bound = saved_bound;
                best = best_sample;
            }
        }
    }
    for (int sb : {2, 8}) {
        if (rc) break;
        przm_gpu_set_kernel_shape(g, sb, brj);
        if ((rc = time_shape(&t))) break;
        if (t < best) { best = t; bsb = sb; }
    }
    przm_gpu_set_kernel_shape(g, bsb, brj);
    if (!rc) {
        przm_gpu_set_kernel_shape(g, bsb, 32);
        rc = time_shape(&t);
        if (!rc && t < best) { best = t; brj = 32; }
    }
    if (d_out) (void)przm_gpu_free(g, d_out);
    przm_gpu_set_tuning(g, bg, bc);
    przm_gpu_set_kernel_shape(g, bsb, brj);
    {
        std::lock_guard<std::mutex> lk(g->mu);
        if (g->c.gexec) { (void)hipGraphExecDestroy(g->c.gexec); g->c.gexec = nullptr; }
        if (g->c.graph) { (void)hipGraphDestroy(g->c.graph); g->c.graph = nullptr; }
    }
    if (tuned) { tuned[0] = bg; tuned[1] = bc; tuned[2] = brj; tuned[3] = bsb; }
    return rc;
}

}  // extern "C"

/* ======================= serving loop ======================= */

namespace {

struct Req {
    uint64_t id = 0;
    std::vector<uint8_t> src;
    std::vector<uint8_t> out;
    std::vector<uint32_t> fsz;
    int32_t status = 1; /* 1 pending, 0 done, <0 error */
    double submit_ms = 0, start_ms = 0, done_ms = 0;
};

struct Slot {
    Ctx c;
    bool busy = false;
    std::vector<std::shared_ptr<Req>> reqs;
    std::vector<przm_span> spans;
    std::vector<przm_block_t> blocks;
    std::vector<uint64_t> first;
    /* pinned host buffers: the staged input, the frames, the block table and the frame sizes.
     * All four copies of a launch are asynchronous, so the worker never waits for a launch's
     * kernels while it has other slots to retire or fill */
    uint8_t *h_in = nullptr, *h_out = nullptr, *h_blk = nullptr, *h_flen = nullptr;
    uint64_t h_in_cap = 0, h_out_cap = 0, h_blk_cap = 0, h_flen_cap = 0;
    uint64_t dev_bytes = 0;
    bool dead = false; /* its stream reported an error: never launched on again */
    hipEvent_t done = nullptr;
    void free_host() {
        for (uint8_t **p : {&h_in, &h_out, &h_blk, &h_flen}) {
            if (*p) (void)hipHostFree(*p);
            *p = nullptr;
        }
        h_in_cap = h_out_cap = h_blk_cap = h_flen_cap = 0;
    }
};

}  // namespace

struct przm_server {
    przm_gpu *g = nullptr;
    int level = 6;
    uint64_t budget = 0, max_batch = 0;
    uint32_t max_blocks = 0; /* blocks one launch may carry (device buffers are sized for it up front) */
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_in, cv_done;
    std::deque<std::shared_ptr<Req>> pending;
    std::map<uint64_t, std::shared_ptr<Req>> all;
    uint64_t next_id = 1;
    uint64_t admitted = 0;  /* bytes the launches in flight use (inside their slots) */
    uint64_t reserved = 0;  /* device bytes the slots hold: at most the budget */
    int32_t failed = 0;     /* a device error stopped the server: every later call fails with it */
    bool stop = false;
    std::thread worker;
    double t0 = 0;
    uint64_t stats[8] = {0};
};

namespace {

uint64_t slot_cost(uint64_t in_bytes, uint64_t nblocks) {
    return in_bytes + (in_bytes + 8 * nblocks) + nblocks * (uint64_t)PRZM_WS_BYTES + nblocks * 1024;
}

int32_t host_grow(uint8_t **p, uint64_t *cap, uint64_t need) {
    if (*cap >= need && *p) return 0;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    HIPCHK(hipHostMalloc((void **)p, need, hipHostMallocDefault));
    *cap = need;
    return 0;
}

/* device bytes a context holds */
uint64_t ctx_bytes(const Ctx &c) {
    uint64_t t = 0;
    for (const DevBuf *b : {&c.in, &c.out, &c.blocks, &c.ws, &c.meta, &c.ftab, &c.flen, &c.foff, &c.run, &c.scratch,
                            &c.dbg, &c.perm, &c.tok, &c.tmp, &c.ows, &c.sws})
        t += b->cap;
    return t;
}

/* device and pinned buffers of a slot for `nb` blocks over `arena` input bytes (sized for a full
 * launch when the server is created; a launch never needs more, since a request is at most
 * max_batch_bytes) */
int32_t slot_reserve(przm_server *s, Slot &sl, uint64_t arena, uint32_t nb) {
    przm_gpu *g = s->g;
    const uint64_t bound = arena + 8ull * nb + 64;
    CHK(host_grow(&sl.h_in, &sl.h_in_cap, arena + 64));
    CHK(host_grow(&sl.h_out, &sl.h_out_cap, bound));
    CHK(host_grow(&sl.h_blk, &sl.h_blk_cap, (uint64_t)std::max<uint32_t>(nb, 1) * sizeof(przm_block_t)));
    CHK(host_grow(&sl.h_flen, &sl.h_flen_cap, (uint64_t)std::max<uint32_t>(nb, 1) * 4));
    CHK(sl.c.in.ensure(arena + 64, &sl.c.gen));
    CHK(sl.c.out.ensure(bound, &sl.c.gen));
    CHK(ensure_ctx(sl.c, std::max<uint32_t>(nb, 1), std::max<uint32_t>(nb, 1), g->lz_grid, false, g->s3_prefix,
                   g->s4_sheaf));
    return 0;
}

/* host copies of a launch's requests (staging into pinned memory, splitting frames out) run on a
 * few threads: one thread's memcpy would cap the server near the offered byte rate */
void par_copy(const std::vector<std::pair<uint8_t *, const uint8_t *>> &dst_src, const std::vector<uint64_t> &len) {
    uint64_t total = 0;
    for (uint64_t l : len) total += l;
    /* one thread per MiB, at most 8: a launch's copies sit on its requests' latency path */
    const unsigned nt = (unsigned)std::min<uint64_t>(8, std::max<uint64_t>(1, total >> 20));
    auto work = [&](unsigned t) {
        /* thread t copies the t-th of nt equal parts of the concatenated byte range */
        const uint64_t lo = total * t / nt, hi = total * (t + 1) / nt;
        uint64_t pos = 0;
        for (size_t i = 0; i < len.size(); i++) {
            const uint64_t a = std::max(pos, lo), b = std::min(pos + len[i], hi);
            if (a < b) std::memcpy(dst_src[i].first + (a - pos), dst_src[i].second + (a - pos), b - a);
            pos += len[i];
        }
    };
    if (nt == 1) { work(0); return; }
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
}

/* stage and launch the batch of a slot (worker thread, server lock not held) */
int32_t launch_slot(przm_server *s, Slot &sl) {
    przm_gpu *g = s->g;
    uint64_t in_bytes = 0;
    sl.spans.clear();
    uint64_t arena = 0;
    for (auto &r : sl.reqs) {
        sl.spans.push_back({arena, (uint64_t)r->src.size()});
        arena += (r->src.size() + 255) & ~255ull;
        in_bytes += r->src.size();
    }
    build_blocks(sl.spans.data(), (uint32_t)sl.spans.size(), sl.blocks, sl.first);
    const uint32_t nb = (uint32_t)sl.blocks.size();
    const uint64_t bound = in_bytes + 8ull * nb + 64;
    CHK(slot_reserve(s, sl, arena, nb));
    {
        std::vector<std::pair<uint8_t *, const uint8_t *>> ds;
        std::vector<uint64_t> ln;
        for (size_t i = 0; i < sl.reqs.size(); i++) {
            ds.push_back({sl.h_in + sl.spans[i].offset, sl.reqs[i]->src.data()});
            ln.push_back(sl.reqs[i]->src.size());
        }
        par_copy(ds, ln);
    }
    Ctx &c = sl.c;
    const uint32_t chunk = std::max<uint32_t>(nb, 1);
    HIPCHK(hipMemcpyAsync(c.in.p, sl.h_in, arena, hipMemcpyHostToDevice, c.st));
    std::memcpy(sl.h_blk, sl.blocks.data(), nb * sizeof(przm_block_t));
    HIPCHK(hipMemcpyAsync(c.blocks.p, sl.h_blk, nb * sizeof(przm_block_t), hipMemcpyHostToDevice, c.st));
    CHK(enqueue(g, c, (const uint8_t *)c.in.p, nb, (uint32_t)depth_of_level(s->level), (uint8_t *)c.out.p, true, false,
                false, g->lz_grid, chunk, nullptr));
    HIPCHK(hipMemcpyAsync(sl.h_flen, c.flen.p, nb * 4ull, hipMemcpyDeviceToHost, c.st));
    /* frames: the device bound covers them; copy the bound and split on completion */
    HIPCHK(hipMemcpyAsync(sl.h_out, c.out.p, bound, hipMemcpyDeviceToHost, c.st));
    if (!sl.done) HIPCHK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    HIPCHK(hipEventRecord(sl.done, c.st));
    return 0;
}

/* split a finished launch's frames into its requests (worker thread, lock not held) */
void split_slot(Slot &sl) {
    const uint32_t *flen = (const uint32_t *)sl.h_flen;
    uint64_t pos = 0;
    std::vector<std::pair<uint8_t *, const uint8_t *>> ds;
    std::vector<uint64_t> ln;
    for (size_t i = 0; i < sl.reqs.size(); i++) {
        Req &r = *sl.reqs[i];
        uint64_t len = 0;
        r.fsz.clear();
        for (uint64_t b = sl.first[i]; b < sl.first[i + 1]; b++) { r.fsz.push_back(flen[b]); len += flen[b]; }
        r.out.resize(len);
        ds.push_back({r.out.data(), sl.h_out + pos});
        ln.push_back(len);
        pos += len;
    }
    par_copy(ds, ln);
}

/* publish a finished (rc == 0) or failed launch (lock held) */
void retire_slot(przm_server *s, Slot &sl, int32_t rc) {
    const double t = now_ms() - s->t0;
    for (auto &rp : sl.reqs) {
        Req &r = *rp;
        r.done_ms = t;
        if (rc) { r.status = rc; continue; }
        r.status = 0;
        s->stats[1]++;
        s->stats[2] += r.src.size();
        s->stats[3] += r.out.size();
        std::vector<uint8_t>().swap(r.src);
    }
    s->admitted -= sl.dev_bytes;
    sl.dev_bytes = 0;
    sl.reqs.clear();
    sl.busy = false;
}

/*
 * The worker: retire launches whose completion event fired, then admit queued
 * requests into a free slot (one launch per slot, up to `streams` in flight,
 * each bounded by max_batch_bytes / max_blocks and all by the HBM budget).
 * The lock only guards the queue and the request states: staging copies,
 * launches and frame splitting run unlocked, so submitters never wait on them.
 */
void serve_loop(przm_server *s) {
    (void)hipSetDevice(s->g->dev);
    std::vector<Slot *> done_slots;
    std::vector<std::pair<Slot *, int32_t>> failed;
    while (true) {
        done_slots.clear();
        bool any_busy = false;
        Slot *launch = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (auto &up : s->slots) {
                Slot &sl = *up;
                if (!sl.busy || sl.dead) continue;
                const hipError_t q = hipEventQuery(sl.done);
                if (q == hipSuccess) done_slots.push_back(&sl);
                else if (q != hipErrorNotReady) {
                    /* the slot's stream may still hold copies into its pinned buffers: the slot is
                     * never used again, and the server fails this launch, every queued request
                     * and every later submit with the error */
                    snprintf(g_err, sizeof g_err, "serving launch failed: %s", hipGetErrorString(q));
                    retire_slot(s, sl, PRZM_GPU_ERR_HIP);
                    sl.dead = true;
                    sl.busy = true;
                    s->failed = PRZM_GPU_ERR_HIP;
                    for (auto &r : s->pending) r->status = PRZM_GPU_ERR_HIP;
                    s->pending.clear();
                    s->cv_done.notify_all();
                } else any_busy = true;
            }
            if (s->stop && s->pending.empty() && !any_busy && done_slots.empty()) break;
            Slot *free_slot = nullptr;
            for (auto &up : s->slots)
                if (!up->busy) { free_slot = up.get(); break; }
            if (free_slot && !s->pending.empty()) {
                uint64_t bytes = 0, nb = 0, cost = 0;
                std::vector<std::shared_ptr<Req>> take;
                while (!s->pending.empty()) {
                    auto &r = s->pending.front();
                    const uint64_t n = r->src.size();
                    const uint64_t rb = n ? (n + PRZM_BLOCK - 1) / PRZM_BLOCK : 1;
                    const uint64_t c2 = slot_cost(bytes + n, nb + rb);
                    if (!take.empty() && (bytes + n > s->max_batch || nb + rb > s->max_blocks ||
                                          s->admitted + c2 > s->budget))
                        break;
                    if (take.empty() && s->admitted + c2 > s->budget && (any_busy || !done_slots.empty()))
                        break; /* wait for memory */
                    bytes += n;
                    nb += rb;
                    cost = c2;
                    take.push_back(r);
                    s->pending.pop_front();
                }
                if (!take.empty()) {
                    const double t = now_ms() - s->t0;
                    for (auto &r : take) r->start_ms = t;
                    free_slot->reqs = std::move(take);
                    free_slot->busy = true;
                    free_slot->dev_bytes = cost;
                    s->admitted += cost;
                    if (s->admitted > s->stats[4]) s->stats[4] = s->admitted;
                    if (free_slot->reqs.size() > s->stats[5]) s->stats[5] = free_slot->reqs.size();
                    s->stats[0]++;
                    if (any_busy) s->stats[6]++;
                    launch = free_slot;
                }
            }
            if (!launch && done_slots.empty()) {
                if (any_busy) {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                } else {
                    s->cv_in.wait_for(lk, std::chrono::milliseconds(2));
                }
                continue;
            }
        }
        if (launch) {
            const int32_t rc = launch_slot(s, *launch);
            if (rc) {
                std::lock_guard<std::mutex> lk(s->mu);
                retire_slot(s, *launch, rc);
                s->cv_done.notify_all();
            }
        }
        if (!done_slots.empty()) {
            for (Slot *sl : done_slots) split_slot(*sl);
            std::lock_guard<std::mutex> lk(s->mu);
            for (Slot *sl : done_slots) retire_slot(s, *sl, 0);
            s->cv_done.notify_all();
        }
    }
}

}  // namespace

extern "C" {

int32_t przm_serve_create(przm_gpu *g, int32_t level, uint64_t hbm_budget, uint64_t max_batch_bytes, uint32_t streams,
                          przm_server **out) {
    if (!g || !out || streams == 0 || streams > 8) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    przm_server *s = new przm_server();
    s->g = g;
    s->level = level;
    s->budget = hbm_budget ? hbm_budget : (64ull << 30);
    s->max_batch = max_batch_bytes ? max_batch_bytes : (256ull << 20);
    s->t0 = now_ms();
    /* a launch carries at most max_batch_bytes in up to max_blocks blocks (small requests take a
     * block each); every slot is sized for that now, so the loop never allocates.  The slots'
     * device memory counts against the HBM budget: as many slots as fit (at most `streams`) are
     * created, and a request waits for a free slot, i.e. for its memory */
    s->max_blocks = (uint32_t)std::min<uint64_t>(s->max_batch / PRZM_BLOCK + 1024, 65536);
    const uint64_t arena = s->max_batch + 256ull * s->max_blocks;
    auto cleanup = [&]() {
        for (auto &x : s->slots) {
            x->free_host();
            x->c.release();
        }
        delete s;
    };
    for (uint32_t i = 0; i < streams; i++) {
        auto sl = std::make_unique<Slot>();
        if (hipSetDevice(g->dev) != hipSuccess || hipStreamCreateWithFlags(&sl->c.st, hipStreamNonBlocking) != hipSuccess) {
            cleanup();
            return PRZM_GPU_ERR_HIP;
        }
        const int32_t rc = slot_reserve(s, *sl, arena, s->max_blocks);
        const uint64_t bytes = ctx_bytes(sl->c);
        if (rc == 0 && s->reserved + bytes > s->budget) {
            /* this slot would exceed the budget: keep the ones that fit */
            sl->free_host();
            sl->c.release();
            if (i == 0) {
                snprintf(g_err, sizeof g_err, "HBM budget %llu below one launch slot (%llu bytes for max_batch_bytes %llu)",
                         (unsigned long long)s->budget, (unsigned long long)bytes, (unsigned long long)s->max_batch);
                cleanup();
                return PRZM_GPU_ERR_CAPACITY;
            }
            break;
        }
        s->slots.push_back(std::move(sl));
        if (rc) {
            cleanup();
            return rc;
        }
        s->reserved += bytes;
    }
    s->stats[7] = s->reserved;
    s->worker = std::thread(serve_loop, s);
    *out = s;
    return 0;
}

int32_t przm_serve_submit(przm_server *s, const uint8_t *src, uint64_t n, uint64_t *id) {
    if (!s || !id || (!src && n)) return PRZM_GPU_ERR_ARG;
    if (n > s->max_batch) {
        snprintf(g_err, sizeof g_err, "request of %llu bytes above max_batch_bytes %llu", (unsigned long long)n,
                 (unsigned long long)s->max_batch);
        return PRZM_GPU_ERR_CAPACITY;
    }
    auto r = std::make_shared<Req>();
    r->src.assign(src, src + n);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->failed) return s->failed;
    if (s->pending.size() > 1000000) return PRZM_GPU_ERR_BUSY;
    r->id = s->next_id++;
    r->submit_ms = now_ms() - s->t0;
    s->pending.push_back(r);
    s->all[r->id] = r;
    *id = r->id;
    s->cv_in.notify_one();
    return 0;
}

int32_t przm_serve_wait(przm_server *s, uint64_t id, int64_t timeout_ms, uint8_t *out, uint64_t cap,
                        uint32_t *frame_sizes, uint64_t frames_cap, przm_result *res) {
    if (!s) return PRZM_GPU_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return PRZM_GPU_ERR_ARG;
    std::shared_ptr<Req> r = it->second;
    auto ready = [&] { return r->status != 1; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return PRZM_GPU_ERR_BUSY;
    if (res) {
        res->id = r->id;
        res->status = r->status;
        res->nframes = (uint32_t)r->fsz.size();
        res->out_len = r->out.size();
        res->submit_ms = r->submit_ms;
        res->start_ms = r->start_ms;
        res->done_ms = r->done_ms;
    }
    if (r->status) { s->all.erase(id); return r->status; }
    if (out) {
        if (r->out.size() > cap) return PRZM_GPU_ERR_CAPACITY;
        std::memcpy(out, r->out.data(), r->out.size());
    }
    if (frame_sizes) {
        if (r->fsz.size() > frames_cap) return PRZM_GPU_ERR_CAPACITY;
        std::memcpy(frame_sizes, r->fsz.data(), r->fsz.size() * 4);
    }
    s->all.erase(id);
    return 0;
}

int32_t przm_serve_stats(przm_server *s, uint64_t stats[8]) {
    if (!s || !stats) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::memcpy(stats, s->stats, sizeof s->stats);
    return 0;
}

int32_t przm_serve_destroy(przm_server *s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_in.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->g->dev);
    for (auto &sl : s->slots) {
        if (sl->c.st) (void)hipStreamSynchronize(sl->c.st);
        if (sl->done) (void)hipEventDestroy(sl->done);
        sl->free_host();
        sl->c.release();
    }
    delete s;
    return 0;
}

}  // extern "C"

#include "przm_stage_api.hip"
