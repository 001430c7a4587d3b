/*
 * The sheaf split kernel (przm_amd/csrc/przm_sheaf.hip, przm_sheaf_split_kernel), the histogram,
 * coder-order and rANS kernels (przm_amd/csrc/przm_entropy.hip: przm_hist_kernel,
 * przm_rans_order_kernel, przm_rans_kernel), the sheaf option's kernels after the coder
 * (przm_sheaf_split, _class, _place, _count and _encode kernels) and the offsets prep kernel
 * (przm_amd/csrc/przm_offsets.hip, przm_offs_prep_kernel) run on the CPU through the fiber
 * emulation of include/hip/hip_runtime.h, for tests/test_simt_kernels.py.  Test infrastructure:
 * built with g++ into a temporary directory by that test.
 */
#include <hip/hip_runtime.h>

#include "../../przm_amd/csrc/przm_entropy.hip"
#include "../../przm_amd/csrc/przm_offsets.hip"
#include "../../przm_amd/csrc/przm_sheaf.hip"

#include <cstddef>
#include <vector>

extern "C" {

/* layout constants the test reads the outputs with */
uint64_t simt_const(int which) {
    switch (which) {
    case 0: return PRZM_SWS_BYTES;
    case 1: return PRZM_SWS_HDR;
    case 2: return PRZM_SWS_FTAB;
    case 3: return PRZM_SWS_FIB;
    case 4: return PRZM_SWS_FIB_CAP;
    case 5: return sizeof(przm_ftab_t);
    case 6: return offsetof(przm_sheaf_t, n);
    case 7: return offsetof(przm_sheaf_t, l);
    case 8: return offsetof(przm_sheaf_t, mat);
    case 9: return offsetof(przm_ftab_t, mat);
    case 10: return sizeof(przm_meta_t);
    case 11: return offsetof(przm_meta_t, mat_len);
    case 12: return PRZM_OWS_BYTES;
    case 13: return PRZM_OWS_FTAB;
    case 14: return PRZM_OWS_CODES;
    case 15: return PRZM_OWS_EXTRA;
    case 16: return sizeof(przm_offs_t);
    case 17: return PRZM_J_N;
    case 18: return offsetof(przm_sheaf_t, k);
    case 19: return offsetof(przm_sheaf_t, choice);
    case 20: return offsetof(przm_sheaf_t, len);
    case 21: return PRZM_OWS_M12 + PRZM_OWS_M12_CAP;
    case 22: return PRZM_OWS_VANS + PRZM_OWS_VANS_CAP;
    default: return 0;
    }
}

/* przm_sheaf_split_kernel over nblocks blocks whose literal streams are lits[off[b], off[b + 1]);
 * sws: nblocks * PRZM_SWS_BYTES bytes, filled by the caller (the kernel writes only its outputs).
 * order: 0 ascending, 1 descending, 2 + seed random thread order between synchronisations. */
int simt_sheaf_split(const uint8_t *lits, const uint32_t *off, uint32_t nblocks, uint8_t *sws, int order) {
    std::vector<uint8_t> ws((size_t)nblocks * PRZM_WS_BYTES + 64, 0);
    std::vector<przm_meta_t> meta(nblocks);
    memset(meta.data(), 0, meta.size() * sizeof(przm_meta_t));
    for (uint32_t b = 0; b < nblocks; b++) {
        const uint32_t n = off[b + 1] - off[b];
        memcpy(ws.data() + (size_t)b * PRZM_WS_BYTES + PRZM_WS_LIT, lits + off[b], n);
        meta[b].len[PRZM_S_LIT] = n;
    }
    return simt::launch(nblocks, SHEAF_NT, order,
                        [&]() { przm_sheaf_split_kernel(ws.data(), meta.data(), sws, nblocks); });
}

/* przm_hist_kernel over nblocks blocks whose four streams (lit, op, meta, var) are
 * streams[off[4 b + k], off[4 b + k + 1]); ftab: nblocks * 4 tables out, mat_len: nblocks * 4 */
int simt_hist(const uint8_t *streams, const uint32_t *off, uint32_t nblocks, uint8_t *ftab, uint32_t *mat_len,
              int order) {
    std::vector<uint8_t> ws((size_t)nblocks * PRZM_WS_BYTES + 64, 0);
    std::vector<przm_meta_t> meta(nblocks);
    memset(meta.data(), 0, meta.size() * sizeof(przm_meta_t));
    const uint32_t slot[4] = {PRZM_WS_LIT, PRZM_WS_OP, PRZM_WS_META, PRZM_WS_VAR};
    const uint32_t sid[4] = {PRZM_S_LIT, PRZM_S_OP, PRZM_S_META, PRZM_S_VAR};
    for (uint32_t b = 0; b < nblocks; b++)
        for (int k = 0; k < 4; k++) {
            const uint32_t n = off[4 * b + k + 1] - off[4 * b + k];
            memcpy(ws.data() + (size_t)b * PRZM_WS_BYTES + slot[k], streams + off[4 * b + k], n);
            meta[b].len[sid[k]] = n;
        }
    const int rc = simt::launch(nblocks, 256, order, [&]() {
        przm_hist_kernel(ws.data(), meta.data(), (przm_ftab_t *)ftab, nblocks);
    });
    for (uint32_t b = 0; b < nblocks; b++)
        for (int k = 0; k < 4; k++) mat_len[4 * b + k] = meta[b].mat_len[k];
    return rc;
}

/* the coder stages of the pipeline: przm_hist_kernel, przm_rans_order_kernel, then
 * przm_rans_kernel (16 jobs a workgroup) over nblocks blocks whose four streams (lit, op, meta,
 * var) are streams[off[4 b + k], off[4 b + k + 1]).  Per block and job kind (PRZM_J_N of them):
 * klen the coded stream bytes (0xFFFFFFFF when the job did not run or overflowed), states the
 * four final states, and the stream bytes at out + (b * PRZM_J_N + kind) * stride. */
int simt_rans(const uint8_t *streams, const uint32_t *off, uint32_t nblocks, int order, uint8_t *out,
              uint32_t stride, uint32_t *klen, uint32_t *states) {
    std::vector<uint8_t> ws((size_t)nblocks * PRZM_WS_BYTES + 64, 0);
    std::vector<przm_meta_t> meta(nblocks);
    std::vector<przm_ftab_t> ftab((size_t)nblocks * 4);
    std::vector<uint32_t> perm((size_t)nblocks * PRZM_J_N);
    memset(meta.data(), 0, meta.size() * sizeof(przm_meta_t));
    const uint32_t slot[4] = {PRZM_WS_LIT, PRZM_WS_OP, PRZM_WS_META, PRZM_WS_VAR};
    const uint32_t sid[4] = {PRZM_S_LIT, PRZM_S_OP, PRZM_S_META, PRZM_S_VAR};
    for (uint32_t b = 0; b < nblocks; b++)
        for (int k = 0; k < 4; k++) {
            const uint32_t n = off[4 * b + k + 1] - off[4 * b + k];
            memcpy(ws.data() + (size_t)b * PRZM_WS_BYTES + slot[k], streams + off[4 * b + k], n);
            meta[b].len[sid[k]] = n;
        }
    int rc = simt::launch(nblocks, 256, order, [&]() { przm_hist_kernel(ws.data(), meta.data(), ftab.data(), nblocks); });
    if (!rc)
        rc = simt::launch(PRZM_J_N, 1024, order, [&]() { przm_rans_order_kernel(meta.data(), nblocks, perm.data()); });
    const uint32_t per_kind = (nblocks + RANS_JOBS - 1) / RANS_JOBS;
    if (!rc)
        rc = simt::launch(PRZM_J_N * per_kind, RANS_NT, order, [&]() {
            przm_rans_kernel(ws.data(), meta.data(), ftab.data(), nblocks, perm.data());
        });
    for (uint32_t b = 0; b < nblocks; b++)
        for (uint32_t k = 0; k < PRZM_J_N; k++) {
            const uint8_t *src;
            uint32_t n, tab;
            uint8_t *end;
            int m8;
            job_layout(k, ws.data() + (size_t)b * PRZM_WS_BYTES, &meta[b], &src, &n, &end, &m8, &tab);
            const uint32_t K = meta[b].enc_len[k];
            klen[b * PRZM_J_N + k] = K;
            for (int j = 0; j < 4; j++) states[(b * PRZM_J_N + k) * 4 + j] = meta[b].enc_state[k][j];
            if (K != PRZM_ENC_OVERFLOW && K <= stride) memcpy(out + (size_t)(b * PRZM_J_N + k) * stride, end - K, K);
        }
    return rc;
}

/* the coder stages and then the sheaf option's stages, in the order the runtime launches them,
 * over nblocks blocks of four streams as in simt_rans; sws: nblocks * PRZM_SWS_BYTES bytes out,
 * meta: nblocks przm_meta_t out */
int simt_sheaf_pipeline(const uint8_t *streams, const uint32_t *off, uint32_t nblocks, int order, uint8_t *sws,
                        uint8_t *meta_out) {
    std::vector<uint8_t> ws((size_t)nblocks * PRZM_WS_BYTES + 64, 0);
    std::vector<przm_meta_t> meta(nblocks);
    std::vector<przm_ftab_t> ftab((size_t)nblocks * 4);
    std::vector<uint32_t> perm((size_t)nblocks * PRZM_J_N), sperm((size_t)nblocks * 16), bins(PRZM_SHEAF_BINS, 0);
    memset(meta.data(), 0, meta.size() * sizeof(przm_meta_t));
    const uint32_t slot[4] = {PRZM_WS_LIT, PRZM_WS_OP, PRZM_WS_META, PRZM_WS_VAR};
    const uint32_t sid[4] = {PRZM_S_LIT, PRZM_S_OP, PRZM_S_META, PRZM_S_VAR};
    for (uint32_t b = 0; b < nblocks; b++)
        for (int k = 0; k < 4; k++) {
            const uint32_t n = off[4 * b + k + 1] - off[4 * b + k];
            memcpy(ws.data() + (size_t)b * PRZM_WS_BYTES + slot[k], streams + off[4 * b + k], n);
            meta[b].len[sid[k]] = n;
        }
    const uint32_t per_kind = (nblocks + RANS_JOBS - 1) / RANS_JOBS, g16 = (nblocks + 15) / 16;
    int rc = simt::launch(nblocks, 256, order, [&]() { przm_hist_kernel(ws.data(), meta.data(), ftab.data(), nblocks); });
    if (!rc) rc = simt::launch(PRZM_J_N, 1024, order, [&]() { przm_rans_order_kernel(meta.data(), nblocks, perm.data()); });
    if (!rc)
        rc = simt::launch(PRZM_J_N * per_kind, RANS_NT, order, [&]() {
            przm_rans_kernel(ws.data(), meta.data(), ftab.data(), nblocks, perm.data());
        });
    if (!rc)
        rc = simt::launch(nblocks, SHEAF_NT, order, [&]() { przm_sheaf_split_kernel(ws.data(), meta.data(), sws, nblocks); });
    if (!rc) rc = simt::launch(g16, 256, order, [&]() { przm_sheaf_class_kernel(sws, nblocks, bins.data()); });
    if (!rc) rc = simt::launch(g16, 256, order, [&]() { przm_sheaf_place_kernel(sws, nblocks, bins.data(), sperm.data()); });
    if (!rc) rc = simt::launch(nblocks, 64, order, [&]() { przm_sheaf_count_kernel(sws, nblocks, sperm.data()); });
    if (!rc)
        rc = simt::launch((nblocks + 3) / 4, 64, order, [&]() { przm_sheaf_encode_kernel(meta.data(), sws, nblocks); });
    memcpy(meta_out, meta.data(), nblocks * sizeof(przm_meta_t));
    return rc;
}

/* the V-ANS stage entry point's kernels (przm_stage_api.hip: przm_gpu_vans_encode): the
 * grid histogram (grid workgroups), the normalised table, the single-job coder and the
 * container; out gets the container (cap bytes), counts the histogram; returns the container
 * length, 0 when the stream overflows its slot, or -1 on a deadlock */
int64_t simt_vans_stage(const uint8_t *src, uint32_t n, uint32_t grid, int order, uint8_t *out, uint64_t cap,
                        uint32_t *counts) {
    std::vector<uint32_t> cnt(256, 0), res(8, 0);
    std::vector<przm_ftab_t> ft(1);
    uint32_t mat_len = 0;
    const uint64_t slot = (2ull * n + 64 + 15) & ~15ull;
    std::vector<uint8_t> tmp(slot + 64, 0);
    uint8_t *slot_end = (uint8_t *)(((uintptr_t)tmp.data() + slot + 15) & ~(uintptr_t)15) - 16;
    int rc = n ? simt::launch(grid, 256, order, [&]() { przm_hist_grid_kernel(src, n, cnt.data()); }) : 0;
    if (!rc) rc = simt::launch(1, 64, order, [&]() { przm_normalize_kernel(cnt.data(), ft.data(), &mat_len); });
    if (!rc)
        rc = simt::launch(1, 64, order, [&]() {
            przm_rans1_kernel(src, n, ft.data(), slot_end, 2u * n + 16u, res.data());
        });
    if (rc) return -1;
    memcpy(counts, cnt.data(), 1024);
    if (res[0] == PRZM_ENC_OVERFLOW) return 0;
    const uint64_t total = 8ull + mat_len + 16 + res[0];
    if (total > cap) return 0;
    rc = simt::launch(1, 256, order, [&]() {
        przm_vans_assemble_kernel(slot_end, res.data(), ft.data(), mat_len, n, out);
    });
    return rc ? -1 : (int64_t)total;
}

/* przm_offs_prep_kernel, then (code != 0) przm_offs_code_kernel, over nblocks blocks whose LEB128
 * offset streams are offs[off[b], off[b + 1]); ows: nblocks * PRZM_OWS_BYTES bytes, filled by the
 * caller */
int simt_offs_prep(const uint8_t *offs, const uint32_t *off, uint32_t nblocks, uint8_t *ows, int order, int code) {
    std::vector<uint8_t> ws((size_t)nblocks * PRZM_WS_BYTES + 64, 0);
    std::vector<przm_meta_t> meta(nblocks);
    memset(meta.data(), 0, meta.size() * sizeof(przm_meta_t));
    for (uint32_t b = 0; b < nblocks; b++) {
        const uint32_t n = off[b + 1] - off[b];
        memcpy(ws.data() + (size_t)b * PRZM_WS_BYTES + PRZM_WS_OFF, offs + off[b], n);
        meta[b].len[PRZM_S_OFF] = n;
    }
    int rc = simt::launch(nblocks, OFFS_NT, order,
                          [&]() { przm_offs_prep_kernel(ws.data(), meta.data(), ows, nblocks); });
    const uint32_t m12_groups = (nblocks + 63) / 64;
    if (!rc && code)
        rc = simt::launch(m12_groups + (nblocks + 15) / 16, 64, order,
                          [&]() { przm_offs_code_kernel(ows, nblocks, m12_groups); });
    return rc;
}

} // extern "C"
