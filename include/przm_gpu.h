/*
 * przm_gpu.h -- C ABI of the MI355X (gfx950) PRZM encode path.
 *
 * Plain pointers and sizes only.  Every function returns an int32 status:
 * 0 on success, a negative PRZM_GPU_ERR_* code otherwise; none aborts.
 * A bound caller (the Go `gpu` package over cgo, Python over ctypes) maps
 * each code to an error and keeps its own CPU path as the silent fallback.
 *
 * Output format ("mode G"): each 64 KiB block of a buffer becomes one
 * self-contained PRZM frame (the frame header's uncompressed size is 24-bit
 * and the match window is the block), frames are written back to back in
 * block order, and the frame sizes are returned beside them.  An empty
 * buffer yields one 8-byte uncompressed frame.  Frames decode with the
 * reference decoders (src/przm_dec.c przm_decompress, decompress.go
 * DecompressTo).
 */
#ifndef PRZM_GPU_H
#define PRZM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRZM_GPU_OK 0
#define PRZM_GPU_ERR_UNAVAILABLE -1   /* no HIP device / mock unavailable */
#define PRZM_GPU_ERR_ARG -2           /* bad argument */
#define PRZM_GPU_ERR_CAPACITY -3      /* caller buffer too small */
#define PRZM_GPU_ERR_HIP -4           /* HIP runtime error */
#define PRZM_GPU_ERR_MOCK -5          /* injected test failure */
#define PRZM_GPU_ERR_OOM -6           /* device allocation failed */
#define PRZM_GPU_ERR_BUSY -7          /* serving queue full */

typedef struct przm_gpu przm_gpu;
typedef struct przm_server przm_server;
typedef struct przm_pinned_pool przm_pinned_pool;

/* GPUCompressor device fields (gpu/bridge.go:283-287).  On AMD:
 * compute_major/minor = gfx major / minor (gfx950 -> 9, 5), sm_count = CUs. */
typedef struct {
    int32_t device;
    int32_t compute_major;
    int32_t compute_minor;
    int32_t sm_count;
    uint64_t total_vram_bytes;
    char device_name[256];
    char arch[64];
} przm_gpu_info;

/* GPUProfile (gpu/types.go:55-72) plus per-stage device times. */
typedef struct {
    uint64_t ingress_dma_bytes;
    double ingress_dma_ms;
    double ingress_dma_gbps;
    double kernel_ms;
    double internal_vram_gbps;
    uint64_t egress_dma_bytes;
    double egress_dma_ms;
    double egress_dma_gbps;
    double total_round_trip_ms;
    double end_to_end_gbps;
    double pcie_overhead_percent;
    int32_t pipelined_chunks;
    int32_t graph_launch;
    double stage_ms[8]; /* check, lz, hist, rans, frame-size, scan, frame-write, section options */
    /* filled by profiled (non-graph) runs: bytes of the blocks the LZ kernel
     * tokenized, token-stream bytes it wrote, frame bytes, uncompressed frames */
    uint64_t lz_in_bytes;
    uint64_t lz_stream_bytes;
    uint64_t frame_bytes;
    uint64_t raw_frames;
} przm_gpu_profile;

/* A buffer inside a device or host arena. */
typedef struct {
    uint64_t offset;
    uint64_t len;
} przm_span;

/* DeviceTokenStreams (gpu/types.go:90-100): token streams resident in device
 * memory.  The command, offset and literal streams of every 64 KiB block,
 * back to back in block order (d_block_lens: u32[num_chunks][3], each
 * block's cmd / off / lit bytes; d_totals: u32[3], the three totals).  The
 * pointers are borrowed from the handle and stay valid until its next call
 * that launches work (every entry point except przm_gpu_download and
 * przm_gpu_encode_streams); `generation` identifies that call. */
typedef struct {
    uint64_t d_cmds;
    uint64_t d_offs;
    uint64_t d_lits;
    uint64_t d_totals;
    uint32_t cmd_len;
    uint32_t off_len;
    uint32_t lit_len;
    uint32_t num_chunks;   /* blocks (frames) of the input */
    uint64_t stream;       /* hipStream_t the kernels ran on */
    uint64_t d_block_lens;
    uint64_t uncompressed_size;
    uint64_t generation;
} przm_device_token_streams;

/* ---- availability and lifecycle (gpu/bridge.go:45-123, 291-449, 513) ---- */

/* IsAvailable (bridge.go:61): 0 and *info filled when a gfx9 device is usable. */
int32_t przm_gpu_probe(przm_gpu_info *info);
/* NewGPUCompressor (bridge.go:291). */
int32_t przm_gpu_create(int32_t device, przm_gpu **out);
/* Close (bridge.go:513). */
int32_t przm_gpu_destroy(przm_gpu *g);
int32_t przm_gpu_get_info(przm_gpu *g, przm_gpu_info *info);
/* SetMockUnavailableForTest / SetMockMatchErrorForTest (bridge.go:45-56):
 * a nonzero code makes every encode entry point fail with that code before
 * any device work. */
void przm_gpu_set_mock_unavailable(int32_t on);
void przm_gpu_set_mock_error(int32_t code);

/* ---- stages (gpu/bridge.go MatchLZ77 1480, Tokenize 1895,
 *      ComputeHistogramHost 2961, VANSEncode4xHost 3233,
 *      PackStreamFlatHost 3079) ---- */

/* Per-position best match of every 64 KiB block of src: desc[p] =
 * (len << 16) | dist (block-relative), 0 where there is none.  n entries. */
int32_t przm_gpu_match(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level, uint32_t *desc,
                       przm_gpu_profile *prof);
/* MatchLZ77Streaming (bridge.go:1461, 1216-1458): the same descriptors,
 * computed chunk by chunk on two streams so that each chunk's upload and the
 * host's staging of the next overlap the device work on the previous one.
 * chunk_bytes 0 picks max(n / 2, 16 MiB), at most 64 MiB; rounded to 64 KiB. */
int32_t przm_gpu_match_streaming(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level, uint64_t chunk_bytes,
                                 uint32_t *desc, przm_gpu_profile *prof);
/* Token streams of every block, concatenated in block order; block_lens
 * gets (cmd, off, lit, op, meta, var, tokens) per block. */
int32_t przm_gpu_tokenize(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level, uint8_t *cmd, uint64_t cmd_cap,
                          uint8_t *off, uint64_t off_cap, uint8_t *lit, uint64_t lit_cap, uint8_t *op,
                          uint64_t op_cap, uint8_t *meta, uint64_t meta_cap, uint8_t *var, uint64_t var_cap,
                          uint32_t *block_lens, uint64_t nblocks_cap, przm_gpu_profile *prof);
/* 256-bin byte histogram of src. */
int32_t przm_gpu_histogram(przm_gpu *g, const uint8_t *src, uint64_t n, uint32_t counts[256],
                           przm_gpu_profile *prof);
/* Adaptive 4-lane V-ANS container of src ('V','A',1,0, u32 n, matrix,
 * 4 states, stream), core/literals/literals.go:368-384.  n <= 16 MiB. */
int32_t przm_gpu_vans_encode(przm_gpu *g, const uint8_t *src, uint64_t n, uint8_t *out, uint64_t cap,
                             uint64_t *out_len, przm_gpu_profile *prof);
/* ModeFlat section [0x00][u32 n][payload] (przm_ans.cu:202-234).  An empty
 * input gives five zero bytes (PackStreamFlatHost, bridge.go:3088-3090). */
int32_t przm_gpu_pack_flat(przm_gpu *g, const uint8_t *src, uint64_t n, uint8_t *out, uint64_t cap,
                           uint64_t *out_len, przm_gpu_profile *prof);

/* ---- the same stages on device buffers (ComputeHistogramGPU bridge.go:2890,
 *      PackStreamFlatGPU 3033, VANSEncode4xGPU 3161); kernel_ms may be NULL ---- */

/* 256-bin histogram of d_in[0, n). */
int32_t przm_gpu_histogram_device(przm_gpu *g, const void *d_in, uint64_t n, uint32_t counts[256], double *kernel_ms);
/* d_out = [0x00][u32 n][d_in[0, n)]; *out_len = n + 5 (0 for an empty input). */
int32_t przm_gpu_pack_flat_device(przm_gpu *g, const void *d_in, void *d_out, uint64_t n, uint64_t *out_len,
                                  double *kernel_ms);
/* The V-ANS container of d_in[0, n) with the caller's table: freq sums to 4096,
 * cdf[s] = sum of freq[< s] with cdf[256] = 4096, mat the serialized matrix
 * (przm_serialize_matrix).  A byte of the input with frequency 0 is
 * PRZM_GPU_ERR_ARG (EncodeVANS4xDirect's ErrCorruptedFrame, literals.go:312).
 * The container's flag byte is 0x00, as the CPU path writes it (literals.go:297). */
int32_t przm_gpu_vans_encode_device(przm_gpu *g, const void *d_in, void *d_out, uint64_t n, const uint16_t freq[256],
                                    const uint16_t cdf[257], const uint8_t *mat, uint32_t mat_len, uint64_t out_cap,
                                    uint64_t *out_len, double *kernel_ms);

/* ---- device-resident token streams (TokenizeDevice bridge.go:2139,
 *      TokenizeDeviceZeroCopy 2247, TokenizePointersDeviceOnly 2163,
 *      DownloadDeviceBytes 2253, EncodeD2D 2466) ---- */

/* The handle's device input buffer of at least n bytes (the reference's
 * slot.dIn), for callers that stage input before
 * przm_gpu_tokenize_device_resident. */
int32_t przm_gpu_input_buffer(przm_gpu *g, uint64_t n, void **d);
/* Host input: match search and tokenizer, streams left in device memory. */
int32_t przm_gpu_tokenize_device(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level,
                                 przm_device_token_streams *out, przm_gpu_profile *prof);
/* The same from input already in device memory (kept valid until the frames
 * are encoded). */
int32_t przm_gpu_tokenize_device_resident(przm_gpu *g, const void *d_src, uint64_t n, int32_t level,
                                          przm_device_token_streams *out, przm_gpu_profile *prof);
/* n bytes at device address d_ptr to h. */
int32_t przm_gpu_download(przm_gpu *g, void *h, uint64_t d_ptr, uint64_t n);
/* EncodeD2D: the frames of the handle's latest token streams (one per 64 KiB
 * block, the same bytes przm_gpu_encode gives).  usize is the tokenized
 * length; window_log must be 0 (each frame's window log follows its block
 * length, compress.go:183-202).  Streams of an earlier call are PRZM_GPU_ERR_ARG. */
int32_t przm_gpu_encode_streams(przm_gpu *g, const przm_device_token_streams *s, uint64_t usize, int32_t window_log,
                                uint8_t *out, uint64_t out_cap, uint64_t *out_len, uint32_t *frame_sizes,
                                uint64_t frames_cap, uint64_t *nframes, przm_gpu_profile *prof);

/* ---- probes (BenchmarkVRAMBandwidth bridge.go:2674, BenchmarkD2DEntropy
 *      2547, BenchmarkEntropyVRAM 3325); results in GB/s ---- */

/* 20 device copies of size_mb MiB: (read + write bytes) / time. */
int32_t przm_gpu_bench_vram(przm_gpu *g, uint32_t size_mb, double *gbps);
/* Per iteration, command and offset streams of size/4 bytes packed flat and a
 * literal stream of size/2 copied into one frame buffer: size bytes / time. */
int32_t przm_gpu_bench_d2d_entropy(przm_gpu *g, uint32_t size_mb, uint32_t iters, double *gbps);
/* The coder itself: size_mb of skewed bytes as 64 KiB literal streams, each
 * histogram, normalisation and 4-lane V-ANS per iteration: bytes / time. */
int32_t przm_gpu_bench_entropy_vram(przm_gpu *g, uint32_t size_mb, uint32_t iters, double *gbps);

/* ---- encode (CompressD2D bridge.go:2486, EncodeD2D 2466) ---- */

/* Host buffer in, frames out (host). */
int32_t przm_gpu_encode(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level, uint8_t *out, uint64_t out_cap,
                        uint64_t *out_len, uint32_t *frame_sizes, uint64_t frames_cap, uint64_t *nframes,
                        przm_gpu_profile *prof);
/* Device buffers in and out: one buffer, one launch per stage. */
int32_t przm_gpu_encode_device(przm_gpu *g, const void *d_src, uint64_t n, int32_t level, void *d_out,
                               uint64_t out_cap, uint64_t *out_len, uint32_t *frame_sizes, uint64_t frames_cap,
                               uint64_t *nframes, przm_gpu_profile *prof);
/* Same as przm_gpu_encode_device, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc); it has no counterpart in the reference.  Allocates
 * nothing: every intermediate buffer lives in one static device array of the
 * library (53.2 MB per GPU, held until the process ends), so n is at most
 * 1 MiB (16 frames; PRZM_GPU_ERR_CAPACITY above), out_cap must hold
 * przm_gpu_frames_bound(n) and frame_sizes is required.  All calls on one GPU
 * share that array: calls that overlap on one GPU (from two handles or two
 * threads) race. */
int32_t przm_gpu_encode_device_noalloc(przm_gpu *g, const void *d_src, uint64_t n, int32_t level, void *d_out,
                                       uint64_t out_cap, uint64_t *out_len, uint32_t *frame_sizes,
                                       uint64_t frames_cap, uint64_t *nframes);
/* Batched form: many buffers of a device arena through one hipGraph launch of
 * the persistent pipeline.  out_off[i] = byte offset of buffer i's first
 * frame in d_out (nbufs + 1 entries, the last = total); frame_sizes gets
 * every frame in buffer order. */
int32_t przm_gpu_encode_batch_device(przm_gpu *g, const void *d_src, const przm_span *bufs, uint32_t nbufs,
                                     int32_t level, void *d_out, uint64_t out_cap, uint64_t *out_off,
                                     uint32_t *frame_sizes, uint64_t frames_cap, uint64_t *nframes,
                                     przm_gpu_profile *prof);
/* Same, with host buffers (adds the DMA both ways). */
int32_t przm_gpu_encode_batch(przm_gpu *g, const uint8_t *src, const przm_span *bufs, uint32_t nbufs, int32_t level,
                              uint8_t *out, uint64_t out_cap, uint64_t *out_off, uint32_t *frame_sizes,
                              uint64_t frames_cap, uint64_t *nframes, przm_gpu_profile *prof);
/* Upper bound of the frame bytes of n input bytes (64 KiB blocks). */
uint64_t przm_gpu_frames_bound(uint64_t n);

/* ---- autotuner ---- */
/* Times candidate launch shapes on a sample (up to ~4 GiB from the front) of
 * the given device arena, one stage at a time, and keeps the fastest for this
 * handle: the blocks per launch (16384, 32768, 65536, or the whole batch when
 * its workspace fits) on two match workgroups per CU, then one per CU (taken
 * only when more than 1 % faster), then the match kernel's chain candidates
 * per search step, then the coder's jobs per workgroup.
 * tuned[0..3] = lz grid, chunk blocks, coder jobs per workgroup, search batch. */
int32_t przm_gpu_autotune(przm_gpu *g, const void *d_src, const przm_span *bufs, uint32_t nbufs, int32_t level,
                          int32_t tuned[4]);
/* lz_grid, chunk_blocks: 0 keeps the current value. */
int32_t przm_gpu_set_tuning(przm_gpu *g, int32_t lz_grid, int32_t chunk_blocks);
/* search_batch in {2, 4, 8}, rans_jobs in {16, 32}; 0 keeps the current value.
 * No shape changes an output byte. */
int32_t przm_gpu_set_kernel_shape(przm_gpu *g, int32_t search_batch, int32_t rans_jobs);
/* tuned[0..3] as przm_gpu_autotune reports them. */
int32_t przm_gpu_get_tuning(przm_gpu *g, int32_t tuned[4]);
/* use_graph = 0 runs the batched form as direct launches (per-stage profiling). */
int32_t przm_gpu_set_options(przm_gpu *g, int32_t use_graph);

/* ---- Section 3 of the frames ---- */
#define PRZM_OFFSETS_FLAT 0   /* 0x09 + the LEB128 varints: the reference GPU path's frames */
#define PRZM_OFFSETS_PREFIX 1 /* the distance-prefix container when shorter */
/* How every frame this handle writes from now on codes its match distances.
 * PRZM_OFFSETS_PREFIX writes Section 3 as the reference CPU path does
 * (packageSections, compress.go:956 -> PackageOffsetsSection,
 * core/distance/distance.go:825-849): distance codes and MSB-first extra bits,
 * the codes as a Mode-12 stream (distance.go:323-376) or, from 64 distances on
 * and when strictly shorter, an adaptive V-ANS container, in the smallest
 * header mode (0x0B / 0x0A / 0x06); taken only when strictly shorter than the
 * varints, and not when the code stream is exactly as long as the match count
 * (both reference decoders would read it as raw codes).  Every decoder of the
 * format reads both.  The option adds about 104 KB of device workspace per
 * block of a launch.  Set it while no call or server is running on the handle;
 * przm_gpu_encode_device_noalloc supports it within its fixed limit. */
int32_t przm_gpu_set_offsets_mode(przm_gpu *g, int32_t mode);

/* ---- Section 4 of the frames ---- */
#define PRZM_LITERALS_VANS 0  /* adaptive V-ANS when shorter, else raw: the reference GPU path */
#define PRZM_LITERALS_SHEAF 1 /* also the four sheaf containers (the reference CPU path) */
/* How every frame this handle writes from now on codes its literals.
 * PRZM_LITERALS_SHEAF writes Section 4 as the reference CPU path does
 * (PackageLiteralsSection, core/literals/literals.go:667-693): the shortest
 * of the adaptive V-ANS container and the four sheaf containers
 * (PackEtaleSheaf, literals.go:601-664: the literals split into four fibers by
 * the class of the one or two literals before each, core/literals/sheaf.go,
 * each fiber V-ANS-coded with its own table when shorter), tried in that order
 * with strict <, when shorter than the literals.  Every decoder of the format
 * reads both.  The option adds about 346 KB of device workspace per block of a
 * launch.  Set it while no call or server is running on the handle;
 * przm_gpu_encode_device_noalloc supports it within its fixed limit. */
int32_t przm_gpu_set_literals_mode(przm_gpu *g, int32_t mode);

/* ---- the PRZS multi-frame stream (decompress.go:842-866 ParseStreamHeader) ---- */
#define PRZM_CONTAINER_FRAMES 0 /* frames back to back, sizes beside them (the default) */
#define PRZM_CONTAINER_STREAM 1 /* a buffer of more than one block as one PRZS stream */
/* How the encode entry points of this handle lay out a buffer from now on.
 * PRZM_CONTAINER_STREAM writes every buffer longer than one 64 KiB block as the
 * reference's multi-chunk stream (what CompressWithOptions writes with
 * Threads > 1 and ChunkSize = 64 KiB): the 17-byte header 'P','R','Z','S',
 * flags 0, u32 chunk count, u64 total uncompressed size (decompress.go:849-866),
 * then per block its u32 frame length and the frame (the chunk walk of
 * cmd/przm-viz/analyzer.go:242-256); the frames are those of the default layout.
 * A buffer of one block or none stays a bare frame, as the reference writes a
 * single frame for input that fits one chunk.  frame_sizes then give the bytes
 * each block occupies (the first block of a stream includes the header, every
 * block its prefix) and out_off the start of each buffer's stream.  The
 * serving loop always writes frames.  Buffers above 2^24 - 1 blocks are
 * refused.  Set it while no call is running on the handle. */
int32_t przm_gpu_set_container(przm_gpu *g, int32_t mode);
/* Upper bound of the output bytes of an n-byte buffer in the stream layout. */
uint64_t przm_gpu_stream_bound(uint64_t n);

/* ---- device memory for callers without an allocator ---- */
int32_t przm_gpu_malloc(przm_gpu *g, uint64_t bytes, void **d);
int32_t przm_gpu_free(przm_gpu *g, void *d);
int32_t przm_gpu_h2d(przm_gpu *g, void *d, const void *h, uint64_t bytes);
int32_t przm_gpu_d2h(przm_gpu *g, void *h, const void *d, uint64_t bytes);
/* Device-to-device copy of `bytes` (the przm_device_copy_kernel launches of
 * gpu/bridge.go:382, 2628-2660 and gpu/kernel/przm_ans.cu:542-575). */
int32_t przm_gpu_copy_d2d(przm_gpu *g, void *dst, const void *src, uint64_t bytes);
int32_t przm_gpu_sync(przm_gpu *g);
/* AllocPinnedHost / FreePinnedHost (gpu/memory.go:51-90). */
int32_t przm_pinned_alloc(uint64_t bytes, void **p);
int32_t przm_pinned_free(void *p);
/* PinnedBufferPool (gpu/memory.go:99-165): NewPinnedBufferPool allocates
 * `initial` buffers of buf_size; Get reuses one or allocates; Put keeps up to
 * max(initial, 8) and frees the rest; Close (destroy) frees the pool. */
int32_t przm_pinned_pool_create(uint64_t buf_size, uint32_t initial, przm_pinned_pool **out);
int32_t przm_pinned_pool_get(przm_pinned_pool *p, void **buf);
int32_t przm_pinned_pool_put(przm_pinned_pool *p, void *buf);
int32_t przm_pinned_pool_destroy(przm_pinned_pool *p);
/* GetPinned (bridge.go:452): a buffer of at least size bytes from the handle's
 * pool of the next power of two (return it with przm_pinned_pool_put; the
 * handle's pools are freed by przm_gpu_destroy). */
int32_t przm_gpu_get_pinned(przm_gpu *g, uint64_t size, void **buf, przm_pinned_pool **pool);

/* ---- host helpers (gpu/bridge.go:2777-2886) ---- */
/* NormalizeFreq: the deterministic normaliser of core/literals/literals.go:78-140
 * (target_m must be 4096). */
int32_t przm_normalize_freq(const uint32_t counts[256], uint32_t target_m, uint16_t freq[256]);
/* BuildCDF: cdf[0] = 0, cdf[s+1] = cdf[s] + freq[s]. */
int32_t przm_build_cdf(const uint16_t freq[256], uint16_t cdf[257]);
/* SerializeMatrix: joint-nibble matrix of literals.go:143-184 (<= 546 bytes). */
int32_t przm_serialize_matrix(const uint16_t freq[256], uint8_t *out, uint64_t cap, uint64_t *len);

/* Text of the last HIP / capacity error on this thread (empty if none). */
const char *przm_gpu_last_error(void);

/* ---- serving loop (continuous batching under an HBM budget) ---- */
typedef struct {
    uint64_t id;
    int32_t status;
    uint32_t nframes;
    uint64_t out_len;
    double submit_ms, start_ms, done_ms; /* clock of the server, ms since creation */
} przm_result;

/* max_batch_bytes bounds one launch and one request.  Each launch slot holds
 * the device memory of a full launch (inputs, frames, block workspace, match
 * scratch), reserved at creation; hbm_budget bounds the slots' total, so as
 * many slots as fit (at most `streams`) are created and a request waits for a
 * free slot.  PRZM_GPU_ERR_CAPACITY when not even one slot fits. */
int32_t przm_serve_create(przm_gpu *g, int32_t level, uint64_t hbm_budget, uint64_t max_batch_bytes,
                          uint32_t streams, przm_server **out);
/* Copies src; returns a request id.  PRZM_GPU_ERR_BUSY when the host queue is
 * full, PRZM_GPU_ERR_CAPACITY when n > max_batch_bytes, and the error code of a
 * failed launch once a device error has stopped the server (that launch's and
 * every queued request's wait returns the same code). */
int32_t przm_serve_submit(przm_server *s, const uint8_t *src, uint64_t n, uint64_t *id);
/* Waits up to timeout_ms (-1 = forever) for request id; copies frames into
 * out (cap bytes) and sizes into frame_sizes when given. */
int32_t przm_serve_wait(przm_server *s, uint64_t id, int64_t timeout_ms, uint8_t *out, uint64_t cap,
                        uint32_t *frame_sizes, uint64_t frames_cap, przm_result *res);
/* counters: [0] launches, [1] requests done, [2] bytes in, [3] bytes out,
 * [4] peak device bytes used by the launches in flight, [5] max requests in
 * one launch, [6] launches that overlapped another in flight, [7] device
 * bytes reserved by the launch slots (<= hbm_budget) */
int32_t przm_serve_stats(przm_server *s, uint64_t stats[8]);
int32_t przm_serve_destroy(przm_server *s);

#ifdef __cplusplus
}
#endif
#endif
