/*
 * The whole batched encode of one buffer on the CPU through the fiber emulation of
 * include/hip/hip_runtime.h, kernel by kernel in the runtime's launch order (przm_runtime.hip):
 * the high-entropy check, the LZ kernel, the histograms, the coder order, the rANS coder, the
 * section options' kernels when asked for (Section 3 prefix: offsets prep and code; Section 4
 * sheaf: split, class, place, count, encode), the frame sizes, their scan and the frame writer.
 * For tests/test_simt_lz.py; test infrastructure, built with g++ into a temporary directory by
 * that test.
 */
#include <hip/hip_runtime.h>

#include "../../przm_amd/csrc/przm_entropy.hip"
#include "lz_emul.cpp"
#include "../../przm_amd/csrc/przm_offsets.hip"
#include "../../przm_amd/csrc/przm_sheaf.hip"

extern "C" {

uint64_t simt_frames_bound(uint64_t n) {
    const uint64_t nb = (n + PRZM_BLOCK - 1) / PRZM_BLOCK;
    return n + 8 * (nb ? nb : 1) + 64;
}

/* frames of src[0, n) (64 KiB blocks) at chain depth `depth` into out (out_cap bytes), their
 * sizes into frame_len (one per block), with the Section-3 prefix option (prefix) and the
 * Section-4 sheaf option (sheaf); returns the launch status, or -2 when out_cap is short */
int simt_encode(const uint8_t *src, uint64_t n, uint32_t depth, int order, uint8_t *out, uint64_t out_cap,
                uint32_t *frame_len, int prefix, int sheaf) {
    const uint32_t nb = (uint32_t)((n + PRZM_BLOCK - 1) / PRZM_BLOCK);
    Guarded in(n + 64), blocks(nb * sizeof(przm_block_t)), ws((size_t)nb * PRZM_WS_BYTES), meta(nb * sizeof(przm_meta_t)),
        ftab((size_t)nb * 4 * sizeof(przm_ftab_t)), perm((size_t)nb * PRZM_J_N * 4), flen(nb * 4), foff(nb * 8),
        run(64 + 16), scratch(LZ_SCRATCH_STRIDE), dout(out_cap), ows(prefix ? (size_t)nb * PRZM_OWS_BYTES : 16),
        sws(sheaf ? (size_t)nb * PRZM_SWS_BYTES : 16), sperm((size_t)nb * 16 * 4), bins(PRZM_SHEAF_BINS * 4);
    memcpy(in.p, src, n);
    memset(in.p + n, 0, 64);
    przm_block_t *bt = (przm_block_t *)blocks.p;
    for (uint32_t b = 0; b < nb; b++) {
        bt[b].src_off = (uint64_t)b * PRZM_BLOCK;
        bt[b].len = (uint32_t)(n - bt[b].src_off < PRZM_BLOCK ? n - bt[b].src_off : PRZM_BLOCK);
        bt[b].flags = 0;
    }
    memset(ws.p, 0, (size_t)nb * PRZM_WS_BYTES);
    memset(meta.p, 0, nb * sizeof(przm_meta_t));
    memset(run.p, 0, 64 + 16);
    przm_meta_t *mt = (przm_meta_t *)meta.p;
    uint32_t *counter = (uint32_t *)(run.p + 64);
    const uint32_t per_kind = (nb + RANS_JOBS - 1) / RANS_JOBS;
    int rc = simt::launch(nb, 64, order, [&]() { przm_check_kernel(in.p, bt, nb); });
    if (!rc)
        rc = simt::launch(1, LZ_NT, order, [&]() {
            przm_lz_kernel_b8(in.p, bt, nb, depth, ws.p, mt, scratch.p, LZ_SCRATCH_STRIDE, counter);
        });
    if (!rc) rc = simt::launch(nb, 256, order, [&]() { przm_hist_kernel(ws.p, mt, (przm_ftab_t *)ftab.p, nb); });
    if (!rc) rc = simt::launch(PRZM_J_N, 1024, order, [&]() { przm_rans_order_kernel(mt, nb, (uint32_t *)perm.p); });
    if (!rc)
        rc = simt::launch(PRZM_J_N * per_kind, RANS_NT, order, [&]() {
            przm_rans_kernel(ws.p, mt, (const przm_ftab_t *)ftab.p, nb, (const uint32_t *)perm.p);
        });
    uint8_t *o = prefix ? ows.p : nullptr, *sh = sheaf ? sws.p : nullptr;
    const uint32_t m12_groups = (nb + 63) / 64, g16 = (nb + 15) / 16;
    if (!rc && prefix) rc = simt::launch(nb, OFFS_NT, order, [&]() { przm_offs_prep_kernel(ws.p, mt, o, nb); });
    if (!rc && prefix)
        rc = simt::launch(m12_groups + g16, 64, order, [&]() { przm_offs_code_kernel(o, nb, m12_groups); });
    if (!rc && sheaf) rc = simt::launch(nb, SHEAF_NT, order, [&]() { przm_sheaf_split_kernel(ws.p, mt, sh, nb); });
    if (sheaf) memset(bins.p, 0, PRZM_SHEAF_BINS * 4);
    if (!rc && sheaf) rc = simt::launch(g16, 256, order, [&]() { przm_sheaf_class_kernel(sh, nb, (uint32_t *)bins.p); });
    if (!rc && sheaf)
        rc = simt::launch(g16, 256, order,
                          [&]() { przm_sheaf_place_kernel(sh, nb, (uint32_t *)bins.p, (uint32_t *)sperm.p); });
    if (!rc && sheaf)
        rc = simt::launch(nb, 64, order, [&]() { przm_sheaf_count_kernel(sh, nb, (const uint32_t *)sperm.p); });
    if (!rc && sheaf) rc = simt::launch((nb + 3) / 4, 64, order, [&]() { przm_sheaf_encode_kernel(mt, sh, nb); });
    if (!rc)
        rc = simt::launch((nb + 255) / 256, 256, order, [&]() {
            przm_frame_size_kernel(bt, mt, ws.p, nb, (uint32_t *)flen.p, o, sh);
        });
    if (!rc)
        rc = simt::launch(1, 1024, order, [&]() {
            przm_scan_kernel((const uint32_t *)flen.p, nb, (uint64_t *)foff.p, (uint64_t *)run.p);
        });
    const uint64_t total = *(const uint64_t *)run.p;
    if (!rc && total > out_cap) rc = -2;
    if (!rc)
        rc = simt::launch(nb, 256, order, [&]() {
            przm_frame_write_kernel(in.p, bt, mt, ws.p, (const przm_ftab_t *)ftab.p, (const uint64_t *)foff.p, dout.p,
                                    nb, o, sh);
        });
    if (!rc) {
        memcpy(out, dout.p, total);
        memcpy(frame_len, flen.p, nb * 4);
    }
    return rc;
}

} // extern "C"
