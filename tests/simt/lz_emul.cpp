/*
 * The LZ kernel (przm_amd/csrc/przm_lz.hip, przm_lz_kernel_b8 and przm_lz_kernel_b2: match
 * search, tokenizer, stream emission) run on the CPU through the fiber emulation of
 * include/hip/hip_runtime.h, for tests/test_simt_lz.py.  Test infrastructure: built with g++
 * into a temporary directory by that test.  The buffers the kernel writes are allocated with an
 * inaccessible page right after them, so a write past their end stops the process.
 */
#include <hip/hip_runtime.h>

#include "../../przm_amd/csrc/przm_lz.hip"

#include <cstddef>
#include <sys/mman.h>
#include <unistd.h>
#include <vector>

namespace {
/* n bytes whose end is followed by a page that faults on any access */
struct Guarded {
    uint8_t *base = nullptr, *p = nullptr;
    size_t len = 0;
    explicit Guarded(size_t n) {
        const size_t pg = (size_t)sysconf(_SC_PAGESIZE), body = (n + pg - 1) / pg * pg;
        len = body + pg;
        base = (uint8_t *)mmap(nullptr, len, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        mprotect(base + body, pg, PROT_NONE);
        p = base + body - n;
    }
    ~Guarded() { munmap(base, len); }
};
} // namespace

extern "C" {

uint64_t simt_lz_const(int which) {
    switch (which) {
    case 0: return PRZM_WS_BYTES;
    case 1: return PRZM_WS_CMD;
    case 2: return PRZM_WS_OFF;
    case 3: return PRZM_WS_LIT;
    case 4: return PRZM_WS_OP;
    case 5: return PRZM_WS_META;
    case 6: return PRZM_WS_VAR;
    case 7: return sizeof(przm_meta_t);
    case 8: return offsetof(przm_meta_t, len);
    default: return 0;
    }
}

/* the LZ kernel with search batch B (2 or 8) on `grid` workgroups (run one after the other; the
 * persistent grid hands out blocks from its counter) over the 64 KiB blocks of src[0, n) at
 * chain depth `depth`; ws: nblocks * PRZM_WS_BYTES bytes out, meta: nblocks przm_meta_t out */
int simt_lz(const uint8_t *src, uint64_t n, uint32_t depth, uint32_t B, uint32_t grid, uint8_t *ws_out,
            uint8_t *meta_out, int order) {
    const uint32_t nblocks = (uint32_t)((n + PRZM_BLOCK - 1) / PRZM_BLOCK);
    Guarded in(n + 64), blocks(nblocks * sizeof(przm_block_t)), ws((size_t)nblocks * PRZM_WS_BYTES),
        meta(nblocks * sizeof(przm_meta_t)), scratch((size_t)grid * LZ_SCRATCH_STRIDE), counter(64);
    memcpy(in.p, src, n);
    memset(in.p + n, 0, 64);
    przm_block_t *bt = (przm_block_t *)blocks.p;
    for (uint32_t b = 0; b < nblocks; b++) {
        bt[b].src_off = (uint64_t)b * PRZM_BLOCK;
        bt[b].len = (uint32_t)(n - bt[b].src_off < PRZM_BLOCK ? n - bt[b].src_off : PRZM_BLOCK);
        bt[b].flags = 0;
    }
    memset(ws.p, 0, (size_t)nblocks * PRZM_WS_BYTES);
    memset(meta.p, 0, nblocks * sizeof(przm_meta_t));
    memset(counter.p, 0, 64);
    const int rc = simt::launch(grid, LZ_NT, order, [&]() {
        if (B == 2)
            przm_lz_kernel_b2(in.p, bt, nblocks, depth, ws.p, (przm_meta_t *)meta.p, scratch.p, LZ_SCRATCH_STRIDE,
                              (uint32_t *)counter.p);
        else
            przm_lz_kernel_b8(in.p, bt, nblocks, depth, ws.p, (przm_meta_t *)meta.p, scratch.p, LZ_SCRATCH_STRIDE,
                              (uint32_t *)counter.p);
    });
    memcpy(ws_out, ws.p, (size_t)nblocks * PRZM_WS_BYTES);
    memcpy(meta_out, meta.p, nblocks * sizeof(przm_meta_t));
    return rc;
}

} // extern "C"
