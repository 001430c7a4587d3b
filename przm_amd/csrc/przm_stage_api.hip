/*
 * przm_stage_api.hip -- the stage-level entry points of include/przm_gpu.h:
 * histograms, ModeFlat packing and V-ANS coding over host or device buffers,
 * device-resident token streams and their frames (TokenizeDevice / EncodeD2D
 * of gpu/bridge.go), the bandwidth probes, and pinned host buffer pools
 * (gpu/memory.go).  Part of przm_runtime.hip's translation unit: it uses that
 * file's contexts, launch sequence and error helpers.
 */

/* PinnedBufferPool (gpu/memory.go:99-165): page-locked buffers of one size, reused */
struct przm_pinned_pool {
    std::mutex mu;
    uint64_t buf_size = 0;
    uint32_t max_cap = 8; /* defaultMaxPinnedPoolCap (memory.go:107) */
    std::vector<void *> free;
};

namespace {

int32_t ensure_events(Ctx &c, size_t k) {
    while (c.ev.size() < k) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c.ev.push_back(e);
    }
    return 0;
}

int32_t elapsed(hipEvent_t a, hipEvent_t b, double *ms) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    *ms = t;
    return 0;
}

uint32_t grid_for(const przm_gpu *g, uint64_t n16) {
    const uint64_t want = (n16 + 255) / 256;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)g->cus * 4));
}

/* counts (256 u32 in device memory) = histogram of d_in[0, n), on the handle's stream */
void launch_hist(przm_gpu *g, const uint8_t *d_in, uint64_t n, uint32_t *d_counts) {
    hipLaunchKernelGGL(przm_zero_kernel, dim3(1), dim3(64), 0, g->c.st, (uint64_t *)d_counts, 128u);
    if (n) hipLaunchKernelGGL(przm_hist_grid_kernel, dim3(grid_for(g, n >> 4)), dim3(256), 0, g->c.st, d_in, n, d_counts);
}

/* byte offsets inside the meta buffer used by the whole-buffer stages */
constexpr uint64_t SM_COUNTS = 0, SM_RES = 1024, SM_MATLEN = 1024 + 64, SM_FTAB = 2048;
constexpr uint64_t SM_BYTES = SM_FTAB + sizeof(przm_ftab_t) + 64;

/* The V-ANS container of d_in[0, n) with the table already at meta + SM_FTAB (its matrix length
 * at meta + SM_MATLEN when mat_len_host is null): coder, then the container into d_out.
 * Returns the container length in *out_len, or CAPACITY when it would exceed out_cap. */
int32_t vans_run(przm_gpu *g, const uint8_t *d_in, uint64_t n, uint8_t *d_out, uint64_t out_cap,
                 const uint32_t *mat_len_host, uint64_t *out_len, double *kernel_ms) {
    Ctx &c = g->c;
    uint8_t *mb = (uint8_t *)c.meta.p;
    const uint64_t slot = (2 * n + 64 + 15) & ~15ull; /* the coder's slot end is 16-byte aligned */
    CHK(c.tmp.ensure(slot, &c.gen));
    CHK(ensure_events(c, 2));
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    hipLaunchKernelGGL(przm_rans1_kernel, dim3(1), dim3(64), 0, c.st, d_in, (uint32_t)n,
                       (const przm_ftab_t *)(mb + SM_FTAB), (uint8_t *)c.tmp.p + slot, (uint32_t)(2 * n + 16),
                       (uint32_t *)(mb + SM_RES));
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    uint32_t res[6];
    HIPCHK(hipMemcpyAsync(res, mb + SM_RES, 20, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipMemcpyAsync(res + 5, mb + SM_MATLEN, 4, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    if (kernel_ms) CHK(elapsed(c.ev[0], c.ev[1], kernel_ms));
    if (res[0] == PRZM_ENC_OVERFLOW) {
        snprintf(g_err, sizeof g_err, "V-ANS stream of %llu bytes exceeds its slot", (unsigned long long)n);
        return PRZM_GPU_ERR_HIP;
    }
    const uint32_t ml = mat_len_host ? *mat_len_host : res[5];
    const uint64_t total = 8ull + ml + 16 + res[0];
    if (total > out_cap) {
        snprintf(g_err, sizeof g_err, "V-ANS container %llu > capacity %llu", (unsigned long long)total,
                 (unsigned long long)out_cap);
        return PRZM_GPU_ERR_CAPACITY;
    }
    hipLaunchKernelGGL(przm_vans_assemble_kernel, dim3(1), dim3(256), 0, c.st, (const uint8_t *)c.tmp.p + slot,
                       (const uint32_t *)(mb + SM_RES), (const przm_ftab_t *)(mb + SM_FTAB), ml, (uint32_t)n, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c.st));
    *out_len = total;
    return 0;
}

/* device token streams of the n input bytes at d_src (already on the device) */
int32_t tokenize_device_impl(przm_gpu *g, const uint8_t *d_src, uint64_t n, int level, przm_device_token_streams *out,
                             przm_gpu_profile *prof) {
    Ctx &c = g->c;
    przm_span sp{0, n};
    std::vector<przm_block_t> blocks;
    std::vector<uint64_t> first;
    build_blocks(&sp, 1, blocks, first);
    for (auto &b : blocks) b.flags = 0; /* every block is tokenized, high-entropy or not */
    const uint32_t nb = (uint32_t)blocks.size();
    CHK(ensure_ctx(c, nb, nb, g->lz_grid, false, g->s3_prefix, g->s4_sheaf));
    HIPCHK(hipMemcpyAsync(c.blocks.p, blocks.data(), nb * sizeof(przm_block_t), hipMemcpyHostToDevice, c.st));
    CHK(ensure_events(c, 4));
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    CHK(enqueue(g, c, d_src, nb, (uint32_t)depth_of_level(level), (uint8_t *)c.out.p, false, true, false, g->lz_grid, nb,
                nullptr));
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    std::vector<przm_meta_t> mt(nb);
    HIPCHK(hipMemcpyAsync(mt.data(), c.meta.p, nb * sizeof(przm_meta_t), hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    /* stream layout in c.tok: cmd | off | lit | totals u32[4] | block lengths u32[nb][3] | offsets u64[nb][3] */
    std::vector<uint64_t> off(3ull * nb);
    std::vector<uint32_t> lens(3ull * nb);
    uint64_t tot[3] = {0, 0, 0};
    const int sidx[3] = {PRZM_S_CMD, PRZM_S_OFF, PRZM_S_LIT};
    for (uint32_t b = 0; b < nb; b++)
        for (int k = 0; k < 3; k++) {
            off[3ull * b + k] = tot[k];
            lens[3ull * b + k] = mt[b].len[sidx[k]];
            tot[k] += mt[b].len[sidx[k]];
        }
    const uint64_t o_off = (tot[0] + 255) & ~255ull, o_lit = o_off + ((tot[1] + 255) & ~255ull);
    const uint64_t o_tot = o_lit + ((tot[2] + 255) & ~255ull), o_bl = o_tot + 256;
    const uint64_t o_do = o_bl + ((12ull * nb + 255) & ~255ull), bytes = o_do + 24ull * nb + 64;
    CHK(c.tok.ensure(bytes, &c.gen));
    uint8_t *t = (uint8_t *)c.tok.p;
    const uint32_t totals[4] = {(uint32_t)tot[0], (uint32_t)tot[1], (uint32_t)tot[2], 0};
    HIPCHK(hipMemcpyAsync(t + o_tot, totals, 16, hipMemcpyHostToDevice, c.st));
    HIPCHK(hipMemcpyAsync(t + o_bl, lens.data(), 12ull * nb, hipMemcpyHostToDevice, c.st));
    HIPCHK(hipMemcpyAsync(t + o_do, off.data(), 24ull * nb, hipMemcpyHostToDevice, c.st));
    hipLaunchKernelGGL(przm_gather_streams_kernel, dim3(nb), dim3(256), 0, c.st, (const uint8_t *)c.ws.p,
                       (const przm_meta_t *)c.meta.p, (const uint64_t *)(t + o_do), t, t + o_off, t + o_lit, nb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[2], c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    *out = przm_device_token_streams{};
    out->d_cmds = (uint64_t)(uintptr_t)t;
    out->d_offs = (uint64_t)(uintptr_t)(t + o_off);
    out->d_lits = (uint64_t)(uintptr_t)(t + o_lit);
    out->d_totals = (uint64_t)(uintptr_t)(t + o_tot);
    out->cmd_len = (uint32_t)tot[0];
    out->off_len = (uint32_t)tot[1];
    out->lit_len = (uint32_t)tot[2];
    out->num_chunks = nb;
    out->stream = (uint64_t)(uintptr_t)c.st;
    out->d_block_lens = (uint64_t)(uintptr_t)(t + o_bl);
    out->uncompressed_size = n;
    out->generation = g->epoch;
    g->tok.epoch = g->epoch;
    g->tok.n = n;
    g->tok.src = d_src;
    g->tok.nblocks = nb;
    g->tok.level = level;
    if (prof) {
        double lz = 0, ga = 0;
        CHK(elapsed(c.ev[0], c.ev[1], &lz));
        CHK(elapsed(c.ev[1], c.ev[2], &ga));
        prof->kernel_ms = lz + ga;
        prof->stage_ms[1] = lz;
        prof->internal_vram_gbps = lz + ga > 0 ? n / ((lz + ga) * 1e6) : 0;
        prof->pipelined_chunks = 1;
    }
    return 0;
}

int32_t pool_alloc(przm_pinned_pool *p, void **buf) {
    HIPCHK(hipHostMalloc(buf, p->buf_size ? p->buf_size : 1, hipHostMallocDefault));
    return 0;
}

}  // namespace

extern "C" {

/* ---- histogram ---- */

int32_t przm_gpu_histogram_device(przm_gpu *g, const void *d_in, uint64_t n, uint32_t counts[256], double *kernel_ms) {
    if (!g || !counts || (!d_in && n)) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    Ctx &c = g->c;
    g->epoch++;
    CHK(c.meta.ensure(SM_BYTES, &c.gen));
    CHK(ensure_events(c, 2));
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    launch_hist(g, (const uint8_t *)d_in, n, (uint32_t *)c.meta.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    HIPCHK(hipMemcpyAsync(counts, c.meta.p, 1024, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    if (kernel_ms) CHK(elapsed(c.ev[0], c.ev[1], kernel_ms));
    return 0;
}

int32_t przm_gpu_histogram(przm_gpu *g, const uint8_t *src, uint64_t n, uint32_t counts[256], przm_gpu_profile *prof) {
    if (!g || !counts || (!src && n)) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    void *d = nullptr;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        HIPCHK(hipSetDevice(g->dev));
        g->epoch++;
        CHK(g->c.in.ensure(n + 64, &g->c.gen));
        d = g->c.in.p;
        if (n) HIPCHK(hipMemcpy(d, src, n, hipMemcpyHostToDevice));
    }
    double ms = 0;
    CHK(przm_gpu_histogram_device(g, d, n, counts, &ms));
    if (prof) {
        prof->ingress_dma_bytes = n;
        prof->kernel_ms = ms;
        prof->internal_vram_gbps = ms > 0 ? n / (ms * 1e6) : 0;
        prof->egress_dma_bytes = 1024;
    }
    return 0;
}

/* ---- ModeFlat packing ---- */

int32_t przm_gpu_pack_flat_device(przm_gpu *g, const void *d_in, void *d_out, uint64_t n, uint64_t *out_len,
                                  double *kernel_ms) {
    if (!g || !out_len || (n && (!d_in || !d_out)) || n > 0xFFFFFFFFull) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    *out_len = 0;
    if (n == 0) return 0; /* PackStreamFlatGPU returns 0 for an empty input (bridge.go:3034-3036) */
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    Ctx &c = g->c;
    CHK(ensure_events(c, 2));
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    hipLaunchKernelGGL(przm_pack_flat_kernel, dim3(grid_for(g, n >> 4)), dim3(256), 0, c.st, (const uint8_t *)d_in, n,
                       (uint8_t *)d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    if (kernel_ms) CHK(elapsed(c.ev[0], c.ev[1], kernel_ms));
    *out_len = n + 5;
    return 0;
}

int32_t przm_gpu_pack_flat(przm_gpu *g, const uint8_t *src, uint64_t n, uint8_t *out, uint64_t cap, uint64_t *out_len,
                           przm_gpu_profile *prof) {
    if (!g || !out || !out_len || (!src && n) || n > 0xFFFFFFFFull) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    if (cap < n + 5) return PRZM_GPU_ERR_CAPACITY;
    if (n == 0) { /* PackStreamFlatHost: five zero bytes (bridge.go:3088-3090) */
        std::memset(out, 0, 5);
        *out_len = 5;
        return 0;
    }
    void *din = nullptr, *dout = nullptr;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        HIPCHK(hipSetDevice(g->dev));
        g->epoch++;
        CHK(g->c.in.ensure(n + 64, &g->c.gen));
        CHK(g->c.out.ensure(n + 64, &g->c.gen));
        din = g->c.in.p;
        dout = g->c.out.p;
        HIPCHK(hipMemcpy(din, src, n, hipMemcpyHostToDevice));
    }
    double ms = 0;
    CHK(przm_gpu_pack_flat_device(g, din, dout, n, out_len, &ms));
    {
        std::lock_guard<std::mutex> lk(g->mu);
        HIPCHK(hipMemcpy(out, dout, *out_len, hipMemcpyDeviceToHost));
    }
    if (prof) {
        prof->ingress_dma_bytes = n;
        prof->egress_dma_bytes = *out_len;
        prof->kernel_ms = ms;
        prof->internal_vram_gbps = ms > 0 ? 2.0 * n / (ms * 1e6) : 0;
    }
    return 0;
}

/* ---- V-ANS ---- */

int32_t przm_gpu_vans_encode_device(przm_gpu *g, const void *d_in, void *d_out, uint64_t n, const uint16_t freq[256],
                                    const uint16_t cdf[257], const uint8_t *mat, uint32_t mat_len, uint64_t out_cap,
                                    uint64_t *out_len, double *kernel_ms) {
    if (!g || !out_len || !freq || !cdf || (mat_len && !mat) || (n && (!d_in || !d_out)) || n > (16ull << 20) ||
        mat_len > sizeof(((przm_ftab_t *)nullptr)->mat))
        return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    *out_len = 0;
    if (n == 0) return 0;
    /* the tables must be a normalised distribution and its CDF (literals.go:258-268) */
    uint32_t sum = 0;
    for (int s = 0; s < 256; s++) {
        if (cdf[s] != sum) return PRZM_GPU_ERR_ARG;
        sum += freq[s];
    }
    if (sum != 4096 || cdf[256] != 4096) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    Ctx &c = g->c;
    g->epoch++;
    CHK(c.meta.ensure(SM_BYTES, &c.gen));
    /* a byte of the input with frequency 0 cannot be coded (EncodeVANS4xDirect returns
     * ErrCorruptedFrame, literals.go:312-314): the histogram finds it before the coder runs */
    uint32_t counts[256];
    launch_hist(g, (const uint8_t *)d_in, n, (uint32_t *)c.meta.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(counts, c.meta.p, 1024, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    for (int s = 0; s < 256; s++)
        if (counts[s] && !freq[s]) {
            snprintf(g_err, sizeof g_err, "byte 0x%02x occurs but has frequency 0", s);
            return PRZM_GPU_ERR_ARG;
        }
    przm_ftab_t ft;
    std::memset(&ft, 0, sizeof ft);
    for (int s = 0; s < 256; s++) ft.fc[s] = (uint32_t)freq[s] | ((uint32_t)cdf[s] << 16);
    if (mat_len) std::memcpy(ft.mat, mat, mat_len);
    HIPCHK(hipMemcpyAsync((uint8_t *)c.meta.p + SM_FTAB, &ft, sizeof ft, hipMemcpyHostToDevice, c.st));
    /* the coder reads its input as aligned 16-byte chunks up to the next 16-byte boundary: the
     * caller's buffer is first copied behind the coder's slot (aligned, with slack) */
    const uint64_t slot = (2 * n + 64 + 15) & ~15ull;
    CHK(c.tmp.ensure(slot + n + 64, &c.gen));
    uint8_t *stage = (uint8_t *)c.tmp.p + slot;
    hipLaunchKernelGGL(przm_copy_kernel, dim3(grid_for(g, n >> 4)), dim3(256), 0, c.st, (const uint8_t *)d_in, stage, n);
    HIPCHK(hipGetLastError());
    return vans_run(g, stage, n, (uint8_t *)d_out, out_cap, &mat_len, out_len, kernel_ms);
}

int32_t przm_gpu_vans_encode(przm_gpu *g, const uint8_t *src, uint64_t n, uint8_t *out, uint64_t cap,
                             uint64_t *out_len, przm_gpu_profile *prof) {
    if (!g || !out || !out_len || (!src && n)) return PRZM_GPU_ERR_ARG;
    if (n > (16ull << 20)) { /* MaxBlockSize, api.go:22 */
        snprintf(g_err, sizeof g_err, "przm_gpu_vans_encode: %llu bytes above 16 MiB", (unsigned long long)n);
        return PRZM_GPU_ERR_ARG;
    }
    CHK(mock_check());
    if (n == 0) { *out_len = 0; return 0; }
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    Ctx &c = g->c;
    g->epoch++;
    CHK(c.in.ensure(n + 64, &c.gen));
    CHK(c.out.ensure(8 + 600 + 16 + 2 * n + 64, &c.gen));
    CHK(c.meta.ensure(SM_BYTES, &c.gen));
    uint8_t *mb = (uint8_t *)c.meta.p;
    HIPCHK(hipMemcpyAsync(c.in.p, src, n, hipMemcpyHostToDevice, c.st));
    /* EncodeAdaptiveVANS4x (literals.go:368-384): histogram, NormalizeFrequencies, coder */
    launch_hist(g, (const uint8_t *)c.in.p, n, (uint32_t *)(mb + SM_COUNTS));
    hipLaunchKernelGGL(przm_normalize_kernel, dim3(1), dim3(64), 0, c.st, (const uint32_t *)(mb + SM_COUNTS),
                       (przm_ftab_t *)(mb + SM_FTAB), (uint32_t *)(mb + SM_MATLEN));
    HIPCHK(hipGetLastError());
    double ms = 0;
    uint64_t len = 0;
    CHK(vans_run(g, (const uint8_t *)c.in.p, n, (uint8_t *)c.out.p, cap, nullptr, &len, &ms));
    HIPCHK(hipMemcpy(out, c.out.p, len, hipMemcpyDeviceToHost));
    *out_len = len;
    if (prof) {
        prof->ingress_dma_bytes = n;
        prof->egress_dma_bytes = len;
        prof->kernel_ms = ms;
        prof->internal_vram_gbps = ms > 0 ? n / (ms * 1e6) : 0;
    }
    return 0;
}

/* ---- device-resident token streams ---- */

int32_t przm_gpu_input_buffer(przm_gpu *g, uint64_t n, void **d) {
    if (!g || !d) return PRZM_GPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    g->epoch++;
    CHK(g->c.in.ensure(n + 64, &g->c.gen));
    *d = g->c.in.p;
    return 0;
}

int32_t przm_gpu_tokenize_device(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level,
                                 przm_device_token_streams *out, przm_gpu_profile *prof) {
    if (!g || !out || (!src && n)) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    *out = przm_device_token_streams{};
    if (n == 0) return 0;
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    g->epoch++;
    CHK(g->c.in.ensure(n + 64, &g->c.gen));
    HIPCHK(hipMemcpyAsync(g->c.in.p, src, n, hipMemcpyHostToDevice, g->c.st));
    CHK(tokenize_device_impl(g, (const uint8_t *)g->c.in.p, n, level, out, prof));
    if (prof) prof->ingress_dma_bytes = n;
    return 0;
}

int32_t przm_gpu_tokenize_device_resident(przm_gpu *g, const void *d_src, uint64_t n, int32_t level,
                                          przm_device_token_streams *out, przm_gpu_profile *prof) {
    if (!g || !out || (!d_src && n)) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    *out = przm_device_token_streams{};
    if (n == 0) return 0;
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    g->epoch++;
    return tokenize_device_impl(g, (const uint8_t *)d_src, n, level, out, prof);
}

int32_t przm_gpu_download(przm_gpu *g, void *h, uint64_t d_ptr, uint64_t n) {
    if (!g || (n && (!h || !d_ptr))) return PRZM_GPU_ERR_ARG;
    if (!n) return 0;
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    HIPCHK(hipMemcpyAsync(h, (const void *)(uintptr_t)d_ptr, n, hipMemcpyDeviceToHost, g->c.st));
    HIPCHK(hipStreamSynchronize(g->c.st));
    return 0;
}

int32_t przm_gpu_encode_streams(przm_gpu *g, const przm_device_token_streams *s, uint64_t usize, int32_t window_log,
                                uint8_t *out, uint64_t out_cap, uint64_t *out_len, uint32_t *frame_sizes,
                                uint64_t frames_cap, uint64_t *nframes, przm_gpu_profile *prof) {
    if (!g || !s || !out || !out_len) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    *out_len = 0;
    if (usize == 0) return 0; /* EncodeD2D returns nothing for uSize <= 0 (bridge.go:2467-2469) */
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    if (window_log != 0) {
        snprintf(g_err, sizeof g_err, "window_log %d: each frame's window log follows its block length; pass 0",
                 window_log);
        return PRZM_GPU_ERR_ARG;
    }
    if (s->generation != g->tok.epoch || g->epoch != g->tok.epoch || usize != g->tok.n ||
        s->num_chunks != g->tok.nblocks || s->uncompressed_size != usize) {
        snprintf(g_err, sizeof g_err, "token streams are not this handle's latest (a later call reused its buffers)");
        return PRZM_GPU_ERR_ARG;
    }
    const uint32_t nb = g->tok.nblocks;
    if (frame_sizes && nb > frames_cap) return PRZM_GPU_ERR_CAPACITY;
    const uint64_t bound = przm_gpu_frames_bound(usize);
    Ctx &c = g->c;
    CHK(c.out.ensure(bound + 64, &c.gen));
    CHK(ensure_events(c, 2));
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    /* the stages after the match finder, over the streams still in the block workspace; the
     * high-entropy check marks the blocks that become uncompressed frames */
    CHK(enqueue(g, c, g->tok.src, nb, (uint32_t)depth_of_level(g->tok.level), (uint8_t *)c.out.p, true, false, false,
                g->lz_grid, nb, nullptr, true));
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    std::vector<uint32_t> fl(nb);
    HIPCHK(hipMemcpyAsync(fl.data(), c.flen.p, nb * 4ull, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    uint64_t total = 0;
    for (uint32_t v : fl) total += v;
    if (total > out_cap) return PRZM_GPU_ERR_CAPACITY;
    const double d0 = now_ms();
    HIPCHK(hipMemcpy(out, c.out.p, total, hipMemcpyDeviceToHost));
    const double d1 = now_ms();
    if (frame_sizes) std::memcpy(frame_sizes, fl.data(), nb * 4ull);
    if (nframes) *nframes = nb;
    *out_len = total;
    if (prof) {
        double ms = 0;
        CHK(elapsed(c.ev[0], c.ev[1], &ms));
        prof->kernel_ms = ms;
        prof->internal_vram_gbps = ms > 0 ? usize / (ms * 1e6) : 0;
        prof->egress_dma_bytes = total;
        prof->egress_dma_ms = d1 - d0;
        prof->egress_dma_gbps = d1 > d0 ? total / ((d1 - d0) * 1e6) : 0;
        prof->pipelined_chunks = 1;
    }
    return 0;
}

/* ---- probes ---- */

int32_t przm_gpu_bench_vram(przm_gpu *g, uint32_t size_mb, double *gbps) {
    if (!g || !gbps || size_mb == 0) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    const uint64_t bytes = (uint64_t)size_mb << 20;
    DevBuf a, b;
    int32_t rc = a.ensure(bytes, nullptr);
    if (!rc) rc = b.ensure(bytes, nullptr);
    Ctx &c = g->c;
    if (!rc) rc = ensure_events(c, 2);
    const uint32_t grid = std::min<uint32_t>((uint32_t)((bytes >> 4) / 256 + 1), (uint32_t)g->cus * 8);
    const int iters = 20; /* BenchmarkVRAMBandwidth: 20 passes, read + write bytes (bridge.go:2709, 2771) */
    if (!rc) {
        hipLaunchKernelGGL(przm_copy_kernel, dim3(grid), dim3(256), 0, c.st, (const uint8_t *)a.p, (uint8_t *)b.p, bytes);
        (void)hipEventRecord(c.ev[0], c.st);
        for (int i = 0; i < iters; i++)
            hipLaunchKernelGGL(przm_copy_kernel, dim3(grid), dim3(256), 0, c.st, (const uint8_t *)a.p, (uint8_t *)b.p,
                               bytes);
        (void)hipEventRecord(c.ev[1], c.st);
        const hipError_t e = hipStreamSynchronize(c.st);
        double ms = 0;
        if (e != hipSuccess) rc = PRZM_GPU_ERR_HIP;
        else rc = elapsed(c.ev[0], c.ev[1], &ms);
        if (!rc) *gbps = ms > 0 ? 2.0 * bytes * iters / (ms * 1e6) : 0;
    }
    a.release();
    b.release();
    return rc;
}

int32_t przm_gpu_bench_d2d_entropy(przm_gpu *g, uint32_t size_mb, uint32_t iters, double *gbps) {
    if (!g || !gbps) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    if (size_mb == 0) size_mb = 16;
    if (iters == 0) iters = 10;
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    Ctx &c = g->c;
    g->epoch++;
    /* BenchmarkD2DEntropy (bridge.go:2547-2670): command and offset streams of n/4 bytes packed
     * flat, a literal stream of n/2 bytes copied, into one frame buffer, per iteration */
    const uint64_t n = (uint64_t)size_mb << 20, lc = n / 4, lo = n / 4, ll = n / 2;
    CHK(c.in.ensure(n + 64, &c.gen));
    CHK(c.out.ensure(n + 64, &c.gen));
    CHK(ensure_events(c, 2));
    uint8_t *in = (uint8_t *)c.in.p, *fo = (uint8_t *)c.out.p;
    HIPCHK(hipMemsetAsync(in, 0x80, lc, c.st));
    HIPCHK(hipMemsetAsync(in + lc, 0x01, lo, c.st));
    HIPCHK(hipMemsetAsync(in + lc + lo, 0x41, ll, c.st));
    const uint64_t s2 = 22, s3 = s2 + lc + 5, s4 = s3 + lo + 5;
    auto pass = [&]() {
        hipLaunchKernelGGL(przm_pack_flat_kernel, dim3(grid_for(g, lc >> 4)), dim3(256), 0, c.st, in, lc, fo + s2);
        hipLaunchKernelGGL(przm_pack_flat_kernel, dim3(grid_for(g, lo >> 4)), dim3(256), 0, c.st, in + lc, lo, fo + s3);
        hipLaunchKernelGGL(przm_copy_kernel, dim3(grid_for(g, ll >> 4)), dim3(256), 0, c.st, in + lc + lo, fo + s4, ll);
    };
    pass();
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    for (uint32_t i = 0; i < iters; i++) pass();
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    double ms = 0;
    CHK(elapsed(c.ev[0], c.ev[1], &ms));
    *gbps = ms > 0 ? (double)n * iters / (ms * 1e6) : 0;
    return 0;
}

int32_t przm_gpu_bench_entropy_vram(przm_gpu *g, uint32_t size_mb, uint32_t iters, double *gbps) {
    if (!g || !gbps) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    if (size_mb == 0) size_mb = 16;
    if (iters == 0) iters = 10;
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    Ctx &c = g->c;
    g->epoch++;
    /* the coder itself: size_mb of bytes with a skewed distribution as 64 KiB literal streams,
     * one V-ANS job each (histogram, normalisation, table, 4-lane coder) per iteration */
    const uint64_t n = (uint64_t)size_mb << 20;
    const uint32_t nb = (uint32_t)((n + PRZM_BLOCK - 1) / PRZM_BLOCK);
    CHK(c.in.ensure(n + 64, &c.gen));
    CHK(ensure_ctx(c, nb, nb, g->lz_grid, false));
    CHK(ensure_events(c, 2));
    {
        std::vector<uint8_t> h(n);
        uint64_t x = 0x9E3779B97F4A7C15ull;
        for (uint64_t i = 0; i < n; i++) { /* xorshift64*, byte = number of trailing zero bits (geometric) */
            x ^= x >> 12; x ^= x << 25; x ^= x >> 27;
            const uint64_t r = x * 0x2545F4914F6CDD1Dull;
            h[i] = (uint8_t)(r ? __builtin_ctzll(r) : 63);
        }
        HIPCHK(hipMemcpy(c.in.p, h.data(), n, hipMemcpyHostToDevice));
    }
    przm_meta_t *meta = (przm_meta_t *)c.meta.p;
    hipLaunchKernelGGL(przm_probe_fill_kernel, dim3(nb), dim3(256), 0, c.st, (const uint8_t *)c.in.p, n,
                       (uint8_t *)c.ws.p, meta, nb);
    auto pass = [&]() {
        hipLaunchKernelGGL(przm_hist_kernel, dim3(nb), dim3(256), 0, c.st, (uint8_t *)c.ws.p, meta,
                           (przm_ftab_t *)c.ftab.p, nb);
        hipLaunchKernelGGL(przm_rans_order_kernel, dim3(PRZM_J_N), dim3(1024), 0, c.st, (const przm_meta_t *)meta, nb,
                           (uint32_t *)c.perm.p);
        hipLaunchKernelGGL(przm_rans_kernel, dim3(PRZM_J_N * ((nb + RANS_JOBS - 1) / RANS_JOBS)), dim3(RANS_NT), 0, c.st,
                           (uint8_t *)c.ws.p, meta, (const przm_ftab_t *)c.ftab.p, nb, (const uint32_t *)c.perm.p);
    };
    pass();
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[0], c.st));
    for (uint32_t i = 0; i < iters; i++) pass();
    HIPCHK(hipEventRecord(c.ev[1], c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    double ms = 0;
    CHK(elapsed(c.ev[0], c.ev[1], &ms));
    *gbps = ms > 0 ? (double)n * iters / (ms * 1e6) : 0;
    return 0;
}

/* ---- pinned host buffers ---- */

int32_t przm_pinned_pool_create(uint64_t buf_size, uint32_t initial, przm_pinned_pool **out) {
    if (!out || buf_size == 0) return PRZM_GPU_ERR_ARG;
    *out = nullptr;
    przm_pinned_pool *p = new (std::nothrow) przm_pinned_pool();
    if (!p) return PRZM_GPU_ERR_OOM;
    p->buf_size = buf_size;
    p->max_cap = std::max<uint32_t>(initial, 8);
    for (uint32_t i = 0; i < initial; i++) {
        void *b = nullptr;
        const int32_t rc = pool_alloc(p, &b);
        if (rc) {
            (void)przm_pinned_pool_destroy(p);
            return rc;
        }
        p->free.push_back(b);
    }
    *out = p;
    return 0;
}

int32_t przm_pinned_pool_get(przm_pinned_pool *p, void **buf) {
    if (!p || !buf) return PRZM_GPU_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(p->mu);
        if (!p->free.empty()) {
            *buf = p->free.back();
            p->free.pop_back();
            return 0;
        }
    }
    return pool_alloc(p, buf);
}

int32_t przm_pinned_pool_put(przm_pinned_pool *p, void *buf) {
    if (!p || !buf) return PRZM_GPU_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(p->mu);
        if (p->free.size() < p->max_cap) {
            p->free.push_back(buf);
            return 0;
        }
    }
    HIPCHK(hipHostFree(buf));
    return 0;
}

int32_t przm_pinned_pool_destroy(przm_pinned_pool *p) {
    if (!p) return 0;
    {
        std::lock_guard<std::mutex> lk(p->mu);
        for (void *b : p->free) (void)hipHostFree(b);
        p->free.clear();
    }
    delete p;
    return 0;
}

int32_t przm_gpu_get_pinned(przm_gpu *g, uint64_t size, void **buf, przm_pinned_pool **pool) {
    if (!g || !buf || !pool || size == 0) return PRZM_GPU_ERR_ARG;
    uint64_t p2 = 1;
    while (p2 < size) p2 <<= 1; /* roundUpPow2 (bridge.go:452-469) */
    przm_pinned_pool *p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        auto it = g->pools.find(p2);
        if (it == g->pools.end()) {
            CHK(przm_pinned_pool_create(p2, 0, &p));
            g->pools[p2] = p;
        } else {
            p = it->second;
        }
    }
    *pool = p;
    return przm_pinned_pool_get(p, buf);
}

}  // extern "C"

/* ---- double-buffered streaming match (matchLZ77StreamingLocked, bridge.go:1216-1458) ---- */

extern "C" int32_t przm_gpu_match_streaming(przm_gpu *g, const uint8_t *src, uint64_t n, int32_t level,
                                            uint64_t chunk_bytes, uint32_t *desc, przm_gpu_profile *prof) {
    if (!g || (!src && n) || (!desc && n)) return PRZM_GPU_ERR_ARG;
    CHK(mock_check());
    if (n == 0) return 0;
    /* chunk size as the reference picks it (max(n / 2, 16 MiB), at most 64 MiB), in whole blocks */
    uint64_t chunk = chunk_bytes ? chunk_bytes : std::max<uint64_t>(n / 2, 16ull << 20);
    chunk = std::min<uint64_t>(chunk, 64ull << 20);
    chunk = std::max<uint64_t>(chunk / PRZM_BLOCK, 1) * PRZM_BLOCK;
    const uint64_t nchunks = (n + chunk - 1) / chunk;
    if (nchunks <= 1) return przm_gpu_match(g, src, n, level, desc, prof);
    std::lock_guard<std::mutex> lk(g->mu);
    HIPCHK(hipSetDevice(g->dev));
    g->epoch++;
    const uint32_t cb = (uint32_t)(chunk / PRZM_BLOCK);
    const uint32_t depth = (uint32_t)depth_of_level(level);
    for (auto &pp : g->pipe) {
        if (!pp.c.st) HIPCHK(hipStreamCreateWithFlags(&pp.c.st, hipStreamNonBlocking));
        if (!pp.done) HIPCHK(hipEventCreateWithFlags(&pp.done, hipEventDisableTiming));
        CHK(ensure_ctx(pp.c, cb, cb, g->lz_grid, true));
        CHK(pp.c.in.ensure(chunk + 64, &pp.c.gen));
        CHK(pp.desc.ensure(chunk * 4, nullptr));
        if (pp.h_cap < chunk) {
            if (pp.h_in) (void)hipHostFree(pp.h_in);
            if (pp.h_desc) (void)hipHostFree(pp.h_desc);
            pp.h_in = nullptr;
            pp.h_desc = nullptr;
            pp.h_cap = 0;
            HIPCHK(hipHostMalloc((void **)&pp.h_in, chunk, hipHostMallocDefault));
            HIPCHK(hipHostMalloc((void **)&pp.h_desc, chunk * 4, hipHostMallocDefault));
            pp.h_cap = chunk;
        }
    }
    const double t0 = now_ms();
    /* chunk i runs on slot i & 1: while the device works on chunk i, the host stages chunk i + 1
     * and its upload overlaps chunk i's kernels; chunk i - 1's descriptors are copied out once
     * its slot's event fires */
    std::vector<przm_block_t> blocks[2];
    uint64_t pend_off[2] = {0, 0}, pend_n[2] = {0, 0};
    bool pend[2] = {false, false};
    auto drain = [&](int s) -> int32_t {
        if (!pend[s]) return 0;
        HIPCHK(hipEventSynchronize(g->pipe[s].done));
        std::memcpy(desc + pend_off[s], g->pipe[s].h_desc, pend_n[s] * 4);
        pend[s] = false;
        return 0;
    };
    for (uint64_t i = 0; i < nchunks; i++) {
        const int s = (int)(i & 1);
        auto &pp = g->pipe[s];
        CHK(drain(s));
        const uint64_t o = i * chunk, len = std::min<uint64_t>(chunk, n - o);
        std::memcpy(pp.h_in, src + o, len);
        przm_span sp{0, len};
        std::vector<uint64_t> first;
        build_blocks(&sp, 1, blocks[s], first);
        for (auto &b : blocks[s]) b.flags = 0;
        const uint32_t nb = (uint32_t)blocks[s].size();
        HIPCHK(hipMemcpyAsync(pp.c.in.p, pp.h_in, len, hipMemcpyHostToDevice, pp.c.st));
        HIPCHK(hipMemcpyAsync(pp.c.blocks.p, blocks[s].data(), nb * sizeof(przm_block_t), hipMemcpyHostToDevice, pp.c.st));
        CHK(enqueue(g, pp.c, (const uint8_t *)pp.c.in.p, nb, depth, (uint8_t *)pp.c.out.p, false, true, true,
                    g->lz_grid, nb, nullptr));
        hipLaunchKernelGGL(przm_desc_kernel, dim3(64, nb), dim3(256), 0, pp.c.st, (const uint8_t *)pp.c.in.p,
                           (const przm_block_t *)pp.c.blocks.p, nb, (const uint16_t *)pp.c.dbg.p, (uint32_t *)pp.desc.p,
                           0ull);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(pp.h_desc, pp.desc.p, len * 4, hipMemcpyDeviceToHost, pp.c.st));
        HIPCHK(hipEventRecord(pp.done, pp.c.st));
        pend[s] = true;
        pend_off[s] = o;
        pend_n[s] = len;
    }
    CHK(drain(0));
    CHK(drain(1));
    if (prof) {
        prof->total_round_trip_ms = now_ms() - t0;
        prof->pipelined_chunks = (int32_t)nchunks;
        prof->ingress_dma_bytes = n;
        prof->egress_dma_bytes = n * 4;
        prof->end_to_end_gbps = prof->total_round_trip_ms > 0 ? n / (prof->total_round_trip_ms * 1e6) : 0;
    }
    return 0;
}
