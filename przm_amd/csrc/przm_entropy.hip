/*
 * przm_entropy.hip -- block checks, histograms, frequency normalisation,
 * 4-lane rANS (V-ANS and Meta8) and frame assembly for gfx950.
 *
 *  przm_check_kernel    checkHighEntropy (compress.go:113-132) per block.
 *  przm_hist_kernel     one wave per (block, stream): 256-bin histogram in
 *                       LDS, NormalizeFrequencies (literals.go:78-140) with
 *                       the deficit handed out to the largest (count,
 *                       -symbol) keys, found by a bitwise search for the
 *                       threshold key, CDF, and the joint-nibble matrix
 *                       (literals.go:143-184).
 *  przm_rans_kernel     one lane quad per (block, coder job).  The format
 *                       fixes four interleaved states per stream
 *                       (literals.go:302-325, vans_meta.go:104-132): quad
 *                       lane j owns state j, the four lanes step together
 *                       through the reverse pass, and the bytes a step
 *                       emits (0-2 per lane) are ordered lane 3,2,1,0 by a
 *                       quad prefix sum and written backward from the end
 *                       of the job's slot, which leaves the container's
 *                       reversed stream in place.  x/f uses an fp32
 *                       reciprocal estimate (x < 2^24, q < 4096, so the
 *                       estimate is within one) with an exact integer
 *                       correction.
 *  przm_frame_size_kernel / przm_frame_write_kernel
 *                       section choices and frame bytes per block
 *                       (commands.go:371-444, 537-568; literals.go:368-384,
 *                       664-693; compress.go:962-1033), plus the mode's
 *                       guard against a raw Section 4 that decoders would
 *                       misread as a container.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "przm_common.h"
#include "przm_meta8_freq.h"

/* ---------------- block check ---------------- */

extern "C" __global__ void __launch_bounds__(64)
przm_check_kernel(const uint8_t *__restrict__ src, przm_block_t *__restrict__ blocks, uint32_t nblocks) {
    __shared__ uint32_t seen[8];
    const uint32_t blk = blockIdx.x, lane = threadIdx.x;
    if (blk >= nblocks) return;
    const przm_block_t bd = blocks[blk];
    const uint32_t n = bd.len;
    const uint32_t sample = n < 512 ? n : 512;
    if (lane < 8) seen[lane] = 0;
    __syncthreads();
    for (uint32_t i = lane; i < sample; i += 64) {
        const uint32_t b = src[bd.src_off + i];
        atomicOr(&seen[b >> 5], 1u << (b & 31));
    }
    __syncthreads();
    if (lane == 0) {
        uint32_t unique = 0;
        for (int k = 0; k < 8; k++) unique += (uint32_t)__popc(seen[k]);
        const uint32_t thr = sample >= 256 ? (256u * 88u) / 100u : (sample * 88u) / 100u;
        const bool high = sample >= 64 && unique > thr;
        blocks[blk].flags = (high || n == 0) ? (bd.flags | PRZM_BF_RAW) : (bd.flags & ~PRZM_BF_RAW);
    }
}

/* ---------------- histogram + normalisation ---------------- */

/* inclusive wave scan in DPP moves (no LDS round trips): row_shr 1, 2, 4, 8 inside each row of
 * 16 lanes, then row 0's last lane into row 1 and row 2's into row 3 (row_bcast:15), then
 * lane 31 into rows 2 and 3 (row_bcast:31); lanes whose source is out of range add 0 */
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t ent_dpp(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWMASK, 0xF, false);
}
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t x) {
    x += ent_dpp<0x111, 0xF>(x);
    x += ent_dpp<0x112, 0xF>(x);
    x += ent_dpp<0x114, 0xF>(x);
    x += ent_dpp<0x118, 0xF>(x);
    x += ent_dpp<0x142, 0xA>(x);
    x += ent_dpp<0x143, 0xC>(x);
    return x;
}
/* the sum over the wave, in every lane */
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_incl(v), 63);
}

/* stream of job kind k (0 lit, 1 op, 2 meta, 3 var) */
__device__ __forceinline__ void stream_of(uint32_t k, uint8_t *wsb, const przm_meta_t *mt, const uint8_t **p,
                                          uint32_t *n) {
    switch (k) {
    case 0: *p = wsb + PRZM_WS_LIT; *n = mt->len[PRZM_S_LIT]; break;
    case 1: *p = wsb + PRZM_WS_OP; *n = mt->len[PRZM_S_OP]; break;
    case 2: *p = wsb + PRZM_WS_META; *n = mt->len[PRZM_S_META]; break;
    default: *p = wsb + PRZM_WS_VAR; *n = mt->len[PRZM_S_VAR]; break;
    }
}

/*
 * NormalizeFrequencies (literals.go:78-140) + CDF + joint-nibble matrix
 * (literals.go:143-184) for the 256 counts in cnt[], by one wave.
 */
__device__ void normalize_wave(const uint32_t *cnt, przm_ftab_t *ft, uint32_t *mat_len) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t c4[4], f4[4], k4[4];
    uint32_t tot = 0, uni = 0;
    for (int i = 0; i < 4; i++) { c4[i] = cnt[lane * 4 + i]; tot += c4[i]; uni += c4[i] ? 1u : 0u; }
    tot = wave_sum(tot);
    uni = wave_sum(uni);
    const uint32_t rem = 4096u - uni;
    uint32_t sum = 0;
    for (int i = 0; i < 4; i++) {
        /* c * rem < 2^28 for c <= 65536; larger streams take the 64-bit product */
        f4[i] = c4[i] ? 1u + (uint32_t)(((uint64_t)c4[i] * rem) / (tot ? tot : 1u)) : 0u;
        sum += f4[i];
        const uint32_t s = lane * 4 + i;
        /* a count reaches 2^24 only in a one-symbol table of exactly 16 MiB (the largest input of
         * the stage entry points), whose deficit is 0 */
        k4[i] = c4[i] >= (1u << 24) ? 0xFFFFFF00u | (255u - s) : (c4[i] << 8) | (255u - s);
    }
    sum = wave_sum(sum);
    /* the deficit goes +1 each to the `diff` largest keys (count, then the lower symbol): the
     * diff-th largest key T, found bit by bit from the top (the keys are distinct, so exactly
     * diff keys are >= T); diff <= the symbols present, so only present symbols qualify */
    const uint32_t diff = 4096u - sum;
    if (diff) {
        uint32_t T = 0;
        for (int b = 31; b >= 0; b--) {
            const uint32_t cand = T | (1u << b);
            uint32_t c = 0;
            for (int i = 0; i < 4; i++) c += k4[i] >= cand ? 1u : 0u;
            if (wave_sum(c) >= diff) T = cand;
        }
        for (int i = 0; i < 4; i++) f4[i] += k4[i] >= T ? 1u : 0u;
    }
    uint32_t loc[4], acc = 0;
    for (int i = 0; i < 4; i++) { loc[i] = f4[i]; acc += loc[i]; }
    uint32_t cum = wave_scan_incl(acc) - acc;
    for (int i = 0; i < 4; i++) {
        ft->fc[lane * 4 + i] = loc[i] | (cum << 16);
        cum += loc[i];
    }
    /* the joint-nibble matrix (literals.go:143-184) by the whole wave: lane l holds symbols
     * 4l..4l+3, a quarter of row l >> 2 (high nibble) with low nibbles 4q..4q+3, q = l & 3.  The
     * matrix is the row mask, then per present row its column mask and the weights of its
     * present symbols in order (one byte below 128, else two); each lane's bytes go at an
     * exclusive scan of the lane byte counts, in row-major order. */
    const uint32_t q = lane & 3u;
    uint32_t colbits = 0, nb = 0;
    for (int i = 0; i < 4; i++)
        if (f4[i]) {
            colbits |= 1u << (4u * q + (uint32_t)i);
            nb += f4[i] < 128u ? 1u : 2u;
        }
    uint32_t cols = colbits;
    cols |= (uint32_t)__shfl_xor((int)cols, 1);
    cols |= (uint32_t)__shfl_xor((int)cols, 2);
    const bool head = q == 0 && cols != 0; /* the row's first lane writes its column mask */
    const uint64_t rb = __ballot(head);
    uint32_t rows = 0;
    for (int h = 0; h < 16; h++) rows |= (uint32_t)((rb >> (4 * h)) & 1ull) << h;
    const uint32_t c = nb + (head ? 2u : 0u);
    const uint32_t incl = wave_scan_incl(c);
    uint32_t off = 2u + incl - c;
    uint8_t *o = ft->mat;
    if (head) { o[off] = (uint8_t)cols; o[off + 1] = (uint8_t)(cols >> 8); off += 2; }
    for (int i = 0; i < 4; i++) {
        const uint32_t wgt = f4[i];
        if (!wgt) continue;
        if (wgt < 128u) o[off++] = (uint8_t)wgt;
        else { o[off] = (uint8_t)((wgt >> 8) | 0x80u); o[off + 1] = (uint8_t)wgt; off += 2; }
    }
    if (lane == 0) { o[0] = (uint8_t)rows; o[1] = (uint8_t)(rows >> 8); }
    if (lane == 63) *mat_len = 2u + incl;
}

/* one wave per (block, stream) job: lit, op, meta, var */
extern "C" __global__ void __launch_bounds__(256)
przm_hist_kernel(uint8_t *__restrict__ ws, przm_meta_t *__restrict__ meta, przm_ftab_t *__restrict__ ftab,
                 uint32_t nblocks) {
    __shared__ uint32_t cnt[4][256];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t job = blockIdx.x * 4 + w;
    const bool alive = job < nblocks * 4;
    const uint32_t blk = alive ? job >> 2 : 0, k = job & 3;
    const uint8_t *p = nullptr;
    uint32_t n = 0;
    if (alive) stream_of(k, ws + (uint64_t)blk * PRZM_WS_BYTES, meta + blk, &p, &n);
    const uint32_t minlen = k == 0 ? 1u : 16u;
    const bool run = alive && n >= minlen;
    for (int i = 0; i < 4; i++) cnt[w][lane * 4 + i] = 0;
    __builtin_amdgcn_wave_barrier();
    if (run) {
        /* streams start 16-byte aligned: 16 bytes per lane and load, two loads in flight */
        const uint4 *p4 = (const uint4 *)p;
        const uint32_t n16 = n >> 4;
        uint32_t i = lane;
        for (; i + 64 < n16; i += 128) {
            const uint4 a = p4[i], b = p4[i + 64];
            const uint32_t wd[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
            for (int k = 0; k < 8; k++) {
                atomicAdd(&cnt[w][wd[k] & 0xFFu], 1u);
                atomicAdd(&cnt[w][(wd[k] >> 8) & 0xFFu], 1u);
                atomicAdd(&cnt[w][(wd[k] >> 16) & 0xFFu], 1u);
                atomicAdd(&cnt[w][wd[k] >> 24], 1u);
            }
        }
        for (; i < n16; i += 64) {
            const uint4 a = p4[i];
            const uint32_t wd[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                atomicAdd(&cnt[w][wd[k] & 0xFFu], 1u);
                atomicAdd(&cnt[w][(wd[k] >> 8) & 0xFFu], 1u);
                atomicAdd(&cnt[w][(wd[k] >> 16) & 0xFFu], 1u);
                atomicAdd(&cnt[w][wd[k] >> 24], 1u);
            }
        }
        for (uint32_t t = (n16 << 4) + lane; t < n; t += 64) atomicAdd(&cnt[w][p[t]], 1u);
    }
    __syncthreads();
    if (!run) return;
    normalize_wave(cnt[w], ftab + (uint64_t)blk * 4 + k, &meta[blk].mat_len[k]);
}

/* whole-buffer histogram (one workgroup): counts_out, and the normalised table */
extern "C" __global__ void __launch_bounds__(256)
przm_hist1_kernel(const uint8_t *__restrict__ src, uint64_t n, uint32_t *__restrict__ counts_out,
                  przm_ftab_t *__restrict__ ft, uint32_t *__restrict__ mat_len) {
    __shared__ uint32_t cnt[256];
    const uint32_t tid = threadIdx.x;
    cnt[tid] = 0;
    __syncthreads();
    uint64_t i = tid;
    for (; i + 3 * 256 < n; i += 4 * 256) {
        atomicAdd(&cnt[src[i]], 1u); atomicAdd(&cnt[src[i + 256]], 1u);
        atomicAdd(&cnt[src[i + 512]], 1u); atomicAdd(&cnt[src[i + 768]], 1u);
    }
    for (; i < n; i += 256) atomicAdd(&cnt[src[i]], 1u);
    __syncthreads();
    counts_out[tid] = cnt[tid];
    if (tid < 64 && ft) normalize_wave(cnt, ft, mat_len);
}

/* ---------------- rANS ---------------- */

/* coder workgroups of RANS_NT threads (one lane quad per job, RANS_JOBS jobs): the default;
 * przm_rans_kernel_128 has twice as many per workgroup (the autotuner picks one) */
#define RANS_NT 64
#define RANS_JOBS (RANS_NT / 4)

/* vans_meta.go:45-68: 0 for 0, 1 for 1..3, 2 for 4..15, 3 for 16..31, 4 for 32..63, 5 for
 * 64..127, 6 for 128..191, 7 for 192..255.  Branch-free from the bit length b of p (b = 0 for
 * p = 0): max((b + 1) / 2, b - 2), plus one from 192 on. */
__device__ __forceinline__ uint32_t ctx_meta(uint32_t p) {
    const int32_t b = 31 - (int32_t)__builtin_clz(2u * p + 1u);
    const int32_t lo = (b + 1) >> 1, hi = b - 2;
    return (uint32_t)(lo > hi ? lo : hi) + (p >= 192u ? 1u : 0u);
}

/*
 * x / f and x % f for f in [1, 4096] and a quotient below 4096 (x <= 4096 f - 1, which the
 * renormalisation guarantees; f = 1 with any x < 2^24 is exact too).  x * rcp(f) is within
 * 1.5 * 2^-23 relative, i.e. 2^-10 absolute, of x / f; adding 2^-9 before truncating gives
 * floor(x / f) or one more, so a single correction, without branches, finishes it.
 */
__device__ __forceinline__ void divmod(uint32_t x, uint32_t f, uint32_t *q, uint32_t *r) {
    uint32_t qq = (uint32_t)__builtin_fmaf((float)x, __builtin_amdgcn_rcpf((float)f), 0x1p-9f);
    int32_t rr = (int32_t)(x - __umul24(qq, f)); /* qq and f both fit in 24 bits */
    const bool lo = rr < 0;
    qq = lo ? qq - 1u : qq;
    rr = lo ? rr + (int32_t)f : rr;
    *q = qq;
    *r = (uint32_t)rr;
}

__device__ __forceinline__ void job_layout(uint32_t kind, uint8_t *wsb, const przm_meta_t *mt, const uint8_t **src,
                                           uint32_t *n, uint8_t **slot_end, int *m8, uint32_t *tab) {
    switch (kind) {
    case PRZM_J_LIT: *src = wsb + PRZM_WS_LIT; *n = mt->len[PRZM_S_LIT]; *slot_end = wsb + PRZM_WS_ENC_LIT + PRZM_ENC_LIT_CAP; *m8 = 0; *tab = 0; break;
    case PRZM_J_OP_M8: *src = wsb + PRZM_WS_OP; *n = mt->len[PRZM_S_OP]; *slot_end = wsb + PRZM_WS_ENC_OP_M8 + PRZM_ENC_OP_CAP; *m8 = 1; *tab = 1; break;
    case PRZM_J_OP_V: *src = wsb + PRZM_WS_OP; *n = mt->len[PRZM_S_OP]; *slot_end = wsb + PRZM_WS_ENC_OP_V + PRZM_ENC_OP_CAP; *m8 = 0; *tab = 1; break;
    case PRZM_J_META_M8: *src = wsb + PRZM_WS_META; *n = mt->len[PRZM_S_META]; *slot_end = wsb + PRZM_WS_ENC_META_M8 + PRZM_ENC_META_CAP; *m8 = 1; *tab = 2; break;
    case PRZM_J_META_V: *src = wsb + PRZM_WS_META; *n = mt->len[PRZM_S_META]; *slot_end = wsb + PRZM_WS_ENC_META_V + PRZM_ENC_META_CAP; *m8 = 0; *tab = 2; break;
    case PRZM_J_VAR_M8: *src = wsb + PRZM_WS_VAR; *n = mt->len[PRZM_S_VAR]; *slot_end = wsb + PRZM_WS_ENC_VAR_M8 + PRZM_ENC_VAR_CAP; *m8 = 1; *tab = 3; break;
    default: *src = wsb + PRZM_WS_VAR; *n = mt->len[PRZM_S_VAR]; *slot_end = wsb + PRZM_WS_ENC_VAR_V + PRZM_ENC_VAR_CAP; *m8 = 0; *tab = 3; break;
    }
}

struct M8Table {
    uint32_t fc[8][256];
    constexpr M8Table() : fc() {
        for (int c = 0; c < 8; c++) {
            uint32_t cum = 0;
            for (int s = 0; s < 256; s++) {
                fc[c][s] = (uint32_t)przm_meta8_freq[c][s] | (cum << 16);
                cum += przm_meta8_freq[c][s];
            }
        }
    }
};
__constant__ M8Table przm_m8 = M8Table();

/* value of lane sel[j] of the quad for quad lane j; ctrl = sel0 | sel1 << 2 | sel2 << 4 | sel3 << 6 */
#define QPERM(v, ctrl) ((uint32_t)__builtin_amdgcn_update_dpp(0, (int)(v), (ctrl), 0xF, 0xF, false))


#define RANS_STAGE 256u /* per-job LDS staging of emitted bytes: two 128-byte halves */

/* reverse the byte order of a 16-byte word */
__device__ __forceinline__ uint4 rev16(uint4 v) {
    return make_uint4(__builtin_bswap32(v.w), __builtin_bswap32(v.z), __builtin_bswap32(v.y), __builtin_bswap32(v.x));
}

#define RANS_PF 8u /* 16-byte symbol chunks in flight per quad (4 steps each) */

template <bool B> struct RansLive { static constexpr bool value = B; };

/* a readable zero chunk: the load address of chunks outside a job's stream */
__device__ __align__(16) uint4 przm_zero16;

/* The table entry of quad lane j's symbol for step g (symbols 4g .. 4g + 3; this lane's is byte
 * j of dw, lower is the word of step g - 1): it depends on the symbols alone, so a chunk's
 * entries are read while the chunk before it is coded. */
template <bool M8>
__device__ __forceinline__ uint32_t rans_entry(uint32_t g, uint32_t dw, uint32_t lower, uint32_t j,
                                               const uint32_t *__restrict__ tab) {
    uint32_t idx = (dw >> (8 * j)) & 0xFFu;
    if (M8) {
        const uint32_t prev = j ? (dw >> (8 * j - 8)) & 0xFFu : lower >> 24;
        idx += (4 * g + j > 0 ? ctx_meta(prev) : 0u) << 8;
    }
    return tab[idx];
}

/* One coder step of quad lane j for step g with its table entry e (freq | cdf << 16), live for
 * symbols below nl; a lane that is not live has some entry of its table and changes nothing.
 * Updates the state x and the emitted byte count K.  LIVE = false: the caller knows the step is
 * live in every lane of the wave. */
template <bool LIVE = true>
__device__ __forceinline__ void rans_step(uint32_t g, uint32_t e, uint32_t j, uint32_t nl,
                                          uint8_t *__restrict__ stg, uint8_t *dummy,
                                          uint8_t *__restrict__ slot_end, uint32_t &x, uint32_t &K) {
    const uint32_t i = 4 * g + j;
    const bool v = !LIVE || i < nl;
    const uint32_t f = e & 0xFFFFu, cdf = e >> 16;
    const uint32_t t = 4096u * f - 1u;
    const bool r1 = v && x > t;
    const uint32_t x1 = r1 ? x >> 8 : x;
    const bool r2 = r1 && x1 > t;
    const uint32_t xn = r2 ? x1 >> 8 : x1;
    const uint32_t ne = (uint32_t)r1 + (uint32_t)r2;
    /* bytes of higher lanes come first: lane 3 emits before lane 2, ...  The quad's counts
     * (0..2 each) as bytes, gathered over the quad by two quad permutes; this lane's offset
     * is the sum of the bytes above its own (v_sad_u8 against zero adds a word's four bytes),
     * the step's total the sum of all four. */
    const uint32_t pn = ne << (8u * j);
    const uint32_t qn = pn + QPERM(pn, 0xB1);  /* lanes [1, 0, 3, 2] */
    const uint32_t rn = qn + QPERM(qn, 0x4E);  /* lanes [2, 3, 0, 1] */
    const uint32_t above = (uint32_t)(0xFFFFFF00ull << (8u * j)); /* bytes of lanes j + 1 .. 3 */
    const uint32_t after = __builtin_amdgcn_sad_u8(rn & above, 0u, 0u);
    const uint32_t tot = __builtin_amdgcn_sad_u8(rn, 0u, 0u);
    const uint32_t k0 = K + after;
    *(r1 ? stg + (k0 & (RANS_STAGE - 1)) : dummy) = (uint8_t)x;
    *(r2 ? stg + ((k0 + 1) & (RANS_STAGE - 1)) : dummy) = (uint8_t)x1;
    const uint32_t K2 = K + tot;
    __builtin_amdgcn_wave_barrier();
    if ((K2 >> 7) != (K >> 7)) { /* the run [128m, 128m + 128) is complete: m = K >> 7 */
        const uint32_t m = K >> 7;
        /* lane j stores bytes k = 128m + 96 - 32j .. 128m + 127 - 32j, which sit at
         * slot_end - 128(m + 1) + 32j + (127 - (k - 128m) - 32j) ascending: the upper 16 of
         * them, reversed, first */
        const uint32_t k_lo = 128u * m + 96u - 32u * j;
        const uint4 w1 = *(const uint4 *)(stg + ((k_lo + 16u) & (RANS_STAGE - 1)));
        const uint4 w0 = *(const uint4 *)(stg + (k_lo & (RANS_STAGE - 1)));
        uint8_t *d = slot_end - 128u * (m + 1u) + 32u * j;
        *(uint4 *)d = rev16(w1);
        *(uint4 *)(d + 16) = rev16(w0);
    }
    __builtin_amdgcn_wave_barrier();
    K = K2;
    uint32_t qq, rr;
    divmod(xn, f, &qq, &rr);
    x = v ? qq * 4096u + cdf + rr : x;
}

/*
 * The reverse 4-state pass of one stream by one lane quad (lane j = state j):
 * literals.go:305-325 / vans_meta.go:109-132.  tab is the job's table
 * (freq | cdf << 16): V-ANS uses tab[sym], Meta8 tab[ctx * 256 + sym].  src
 * must be 16-byte aligned.
 *
 * Symbols arrive 16 at a time: chunk c holds the symbols of steps 4c .. 4c + 3.
 * RANS_PF chunks are in flight in a register ring whose slots are named
 * statically (the loop over one turn of the ring is unrolled), and every load
 * is issued unconditionally (chunks outside the stream read przm_zero16), so
 * the wait for a chunk waits for that load only, RANS_PF chunks after it was
 * issued, not for the loads and stores issued since.
 *
 * Every quad of the wave takes the same number of steps, 4 * RANS_PF * the
 * wave's longest job in ring turns (gmax: the longest job in 4-symbol groups),
 * so the steps run in uniform control flow; a lane is live only for its own
 * symbols.  The output bound is checked once per chunk: a job past cap
 * (never chosen) stops there, having written at most 32 bytes past cap, which
 * every slot has room for.
 *
 * Output byte k (k = 0, 1, ... in emission order) belongs at slot_end[-1-k].
 * Bytes are first written to the job's LDS staging ring stg[k % 256]; each
 * time a 128-byte run [128m, 128m + 128) is complete the quad stores it as
 * eight aligned 16-byte words, two per lane (slot_end is 16-byte aligned), so
 * a wave takes the store branch half as often as with 64-byte runs.  At most cap bytes;
 * returns the byte count (or overflow) and the lane's final state.
 */
template <bool M8>
__device__ __forceinline__ uint32_t rans_quad(const uint8_t *__restrict__ src, uint32_t n,
                                              const uint32_t *__restrict__ tab, uint8_t *__restrict__ stg,
                                              uint8_t *__restrict__ slot_end, uint32_t cap, uint32_t gmax,
                                              uint32_t *state) {
    const uint32_t j = threadIdx.x & 3;
    const int32_t own = (int32_t)((n + 15u) >> 4); /* 16-byte chunks of this job's stream */
    const int32_t NC = (int32_t)((((gmax + 3u) >> 2) + RANS_PF - 1u) / RANS_PF * RANS_PF);
    const uint4 *src4 = (const uint4 *)src;
    auto chunk = [&](int32_t c) -> const uint4 * { return c >= 0 && c < own ? src4 + c : &przm_zero16; };
    uint32_t x = 65536u, K = 0, nl = n;
    bool over = false;
    uint8_t *const dummy = stg + RANS_STAGE + j; /* the target of stores that do not happen */
    uint4 buf[RANS_PF];
#pragma unroll
    for (uint32_t d = 0; d < RANS_PF; d++) buf[d] = *chunk(NC - 1 - (int32_t)d);
    /* entries of the chunk being coded; slot s + 1 of the ring holds the chunk below slot s's */
    uint32_t e[4];
    {
        const uint32_t g0 = 4u * (uint32_t)(NC - 1);
        e[0] = rans_entry<M8>(g0 + 3, buf[0].w, buf[0].z, j, tab);
        e[1] = rans_entry<M8>(g0 + 2, buf[0].z, buf[0].y, j, tab);
        e[2] = rans_entry<M8>(g0 + 1, buf[0].y, buf[0].x, j, tab);
        e[3] = rans_entry<M8>(g0, buf[0].x, buf[1].w, j, tab);
    }
    /* one turn of the ring from chunk cb down; LIVE = false: every step of the turn is live in
     * every lane and no quad reaches its cap in it, so neither is tested */
    auto turn = [&](int32_t cb, auto live) {
        constexpr bool LIVE = decltype(live)::value;
#pragma unroll
        for (uint32_t d = 0; d < RANS_PF; d++) {
            const int32_t c = cb - (int32_t)d;
            /* the entries of chunk c - 1 (its slot was reloaded earlier in this turn when it wraps) */
            const uint4 nb = buf[(d + 1) % RANS_PF];
            const uint32_t nlw = buf[(d + 2) % RANS_PF].w;
            const uint32_t gn = 4u * (uint32_t)(c - 1);
            uint32_t en[4];
            en[0] = rans_entry<M8>(gn + 3, nb.w, nb.z, j, tab);
            en[1] = rans_entry<M8>(gn + 2, nb.z, nb.y, j, tab);
            en[2] = rans_entry<M8>(gn + 1, nb.y, nb.x, j, tab);
            en[3] = rans_entry<M8>(gn, nb.x, nlw, j, tab);
            const uint32_t g0 = 4u * (uint32_t)c;
            rans_step<LIVE>(g0 + 3, e[0], j, nl, stg, dummy, slot_end, x, K);
            rans_step<LIVE>(g0 + 2, e[1], j, nl, stg, dummy, slot_end, x, K);
            rans_step<LIVE>(g0 + 1, e[2], j, nl, stg, dummy, slot_end, x, K);
            rans_step<LIVE>(g0, e[3], j, nl, stg, dummy, slot_end, x, K);
#pragma unroll
            for (int u = 0; u < 4; u++) e[u] = en[u];
            buf[d] = *chunk(c - (int32_t)RANS_PF);
            if (LIVE && K > cap) { /* the job is not chosen any more: stop it */
                over = true;
                nl = 0;
            }
        }
    };
    /* the wave's shortest job: a turn below it is live in every lane (idle quads have n = 0) */
    uint32_t nmin = n;
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t y = (uint32_t)__shfl_xor((int)nmin, o);
        nmin = y < nmin ? y : nmin;
    }
    for (int32_t cb = NC - 1; cb >= 0; cb -= (int32_t)RANS_PF) {
        /* a turn emits at most 2 bytes per lane and step: 256 per quad */
        if (16u * (uint32_t)cb + 16u <= nmin && __all(K + 256u <= cap))
            turn(cb, RansLive<false>{});
        else
            turn(cb, RansLive<true>{});
    }
    if (!over) { /* the last, partial run [128 (K >> 7), K) */
        const uint32_t base = K & ~127u;
        for (uint32_t k = base + j; k < K; k += 4) slot_end[-1 - (int64_t)k] = stg[k & (RANS_STAGE - 1)];
    }
    *state = x;
    return over ? PRZM_ENC_OVERFLOW : K;
}

/* the stream a coder job reads */
__device__ __forceinline__ uint32_t job_stream(uint32_t kind) {
    return kind == PRZM_J_LIT ? PRZM_S_LIT : kind <= PRZM_J_OP_V ? PRZM_S_OP : kind <= PRZM_J_META_V ? PRZM_S_META : PRZM_S_VAR;
}

/*
 * Order of the coder jobs of one kind (one workgroup per kind): the blocks by
 * the length of the job's stream in 256-byte classes, longest first, into
 * perm[kind * nblocks ..].  A wave of przm_rans_kernel steps as often as its
 * longest job, so jobs of like length share a wave, and the longest waves
 * start first.  The order inside a class is whatever the atomics give: every
 * job writes only its own outputs, so the results do not depend on it.
 */
extern "C" __global__ void __launch_bounds__(1024)
przm_rans_order_kernel(const przm_meta_t *__restrict__ meta, uint32_t nblocks, uint32_t *__restrict__ perm) {
    __shared__ uint32_t cnt[256];
    const uint32_t kind = blockIdx.x, tid = threadIdx.x, lane = tid & 63u;
    const uint32_t sx = job_stream(kind);
    if (tid < 256) cnt[tid] = 0;
    __syncthreads();
    for (uint32_t b = tid; b < nblocks; b += blockDim.x) {
        const uint32_t c = meta[b].len[sx] >> 8;
        atomicAdd(&cnt[255u - (c < 255u ? c : 255u)], 1u);
    }
    __syncthreads();
    if (tid < 64) { /* exclusive scan of the 256 counts, four per lane */
        uint32_t c4[4], t = 0;
        for (int k = 0; k < 4; k++) { c4[k] = cnt[lane * 4 + k]; t += c4[k]; }
        uint32_t x = t;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= (uint32_t)o) x += y;
        }
        uint32_t run = x - t;
        for (int k = 0; k < 4; k++) { cnt[lane * 4 + k] = run; run += c4[k]; }
    }
    __syncthreads();
    for (uint32_t b = tid; b < nblocks; b += blockDim.x) {
        const uint32_t c = meta[b].len[sx] >> 8;
        const uint32_t pos = atomicAdd(&cnt[255u - (c < 255u ? c : 255u)], 1u);
        perm[(uint64_t)kind * nblocks + pos] = b;
    }
}

/*
 * One lane quad per (block, coder job).  Workgroups are kind-aligned: the
 * first ceil(nblocks / 16) code the literal streams of 16 blocks each, the
 * next the op streams with Meta8, and so on, so a workgroup's table is
 * either sixteen per-job V-ANS tables or the one shared Meta8 table, both in
 * LDS.  Inside a kind the blocks come in the order perm gives
 * (przm_rans_order_kernel).  grid = PRZM_J_N * ceil(nblocks / RANS_JOBS).
 */
template <uint32_t NT>
struct RansShared {
    union {
        uint32_t vt[NT / 4][256];
        uint32_t m8[8][256];
    } T;
    __align__(16) uint8_t stage[NT / 4][RANS_STAGE + 16]; /* + a dummy byte per lane */
};

template <uint32_t NT>
__device__ __forceinline__ void rans_body(RansShared<NT> &SH, uint8_t *__restrict__ ws, przm_meta_t *__restrict__ meta,
                                          const przm_ftab_t *__restrict__ ftab, uint32_t nblocks,
                                          const uint32_t *__restrict__ perm) {
    constexpr uint32_t RANS_JOBS_T = NT / 4;
    auto &T = SH.T;
    auto &stage = SH.stage;
    const uint32_t tid = threadIdx.x, q = tid >> 2, j = tid & 3;
    const uint32_t per_kind = (nblocks + RANS_JOBS_T - 1) / RANS_JOBS_T;
    const uint32_t kind = blockIdx.x / per_kind;
    if (kind >= PRZM_J_N) return;
    const uint32_t slot = (blockIdx.x - kind * per_kind) * RANS_JOBS_T + q;
    const uint32_t blk = slot < nblocks ? perm[(uint64_t)kind * nblocks + slot] : nblocks;
    const bool m8 = kind == PRZM_J_OP_M8 || kind == PRZM_J_META_M8 || kind == PRZM_J_VAR_M8;
    bool run = blk < nblocks;
    const uint8_t *src = nullptr;
    uint8_t *slot_end = nullptr;
    uint32_t n = 0, tsel = 0;
    int jm8 = 0;
    uint8_t *wsb = ws + (uint64_t)blk * PRZM_WS_BYTES;
    if (run) {
        job_layout(kind, wsb, meta + blk, &src, &n, &slot_end, &jm8, &tsel);
        const uint32_t minlen = kind == PRZM_J_LIT ? 1u : (m8 ? 24u : 16u);
        run = n >= minlen;
    }
    if (m8) {
        for (uint32_t e = tid; e < 8u * 256u; e += NT) (&T.m8[0][0])[e] = przm_m8.fc[e >> 8][e & 255u];
    } else if (run) {
        const przm_ftab_t *ft = ftab + (uint64_t)blk * 4 + tsel;
        for (uint32_t s = j; s < 256; s += 4) T.vt[q][s] = ft->fc[s];
    }
    __syncthreads();
    /* the wave's longest job sets the step count of every quad (idle quads have none) */
    uint32_t gmax = run ? (n + 3) >> 2 : 0u;
    for (int o = 32; o >= 4; o >>= 1) {
        const uint32_t y = __shfl_xor(gmax, o);
        gmax = y > gmax ? y : gmax;
    }
    uint32_t x;
    const uint32_t K = m8 ? rans_quad<true>(src, run ? n : 0u, &T.m8[0][0], stage[q], slot_end, n, gmax, &x)
                          : rans_quad<false>(src, run ? n : 0u, T.vt[q], stage[q], slot_end, n, gmax, &x);
    if (!run) {
        if (blk < nblocks && j == 0) meta[blk].enc_len[kind] = PRZM_ENC_OVERFLOW;
        return;
    }
    przm_meta_t *mt = meta + blk;
    if (j == 0) mt->enc_len[kind] = K;
    mt->enc_state[kind][j] = x;
}

/* grid = PRZM_J_N * ceil(nblocks / (threads / 4)) */
extern "C" __global__ void __launch_bounds__(64)
przm_rans_kernel(uint8_t *__restrict__ ws, przm_meta_t *__restrict__ meta, const przm_ftab_t *__restrict__ ftab,
                 uint32_t nblocks, const uint32_t *__restrict__ perm) {
    __shared__ RansShared<64> SH;
    rans_body<64>(SH, ws, meta, ftab, nblocks, perm);
}
extern "C" __global__ void __launch_bounds__(128)
przm_rans_kernel_128(uint8_t *__restrict__ ws, przm_meta_t *__restrict__ meta, const przm_ftab_t *__restrict__ ftab,
                     uint32_t nblocks, const uint32_t *__restrict__ perm) {
    __shared__ RansShared<128> SH;
    rans_body<128>(SH, ws, meta, ftab, nblocks, perm);
}

/* a single V-ANS job over an arbitrary buffer (one quad); slot_end 16-byte aligned */
extern "C" __global__ void __launch_bounds__(64)
przm_rans1_kernel(const uint8_t *__restrict__ src, uint32_t n, const przm_ftab_t *__restrict__ ft,
                  uint8_t *__restrict__ slot_end, uint32_t cap, uint32_t *__restrict__ out) {
    __shared__ uint32_t vt[256];
    __shared__ __align__(16) uint8_t stg[RANS_STAGE + 16];
    for (uint32_t s = threadIdx.x; s < 256; s += 64) vt[s] = ft->fc[s];
    __syncthreads();
    /* the whole wave steps (rans_quad's wave-wide exchanges read every lane): quads 1..15 have
     * an empty job, which emits nothing and keeps the coder on its tested turns */
    const bool job = threadIdx.x < 4;
    uint32_t x;
    const uint32_t K = rans_quad<false>(src, job ? n : 0u, vt, stg, slot_end, cap, (n + 3) >> 2, &x);
    if (threadIdx.x == 0) out[0] = K;
    if (job) out[1 + threadIdx.x] = x;
}

/* ---------------- frame assembly ---------------- */

struct Plan {
    uint32_t kind, flags, l2, l3, l4, ext, sub_len[3], sub_ch[3], s4v;
    uint32_t s3m, s3c, s3cl; /* Section 3: mode byte (0 none), prefix code stream kind and bytes */
};

/* SelectPrefixMode (distance.go:750-758): the container's mode byte and header bytes */
__device__ __forceinline__ uint32_t prefix_mode(uint32_t orig, uint32_t n, uint32_t cl, uint32_t *h) {
    if (orig <= 255u && n <= 255u && cl <= 255u) { *h = 4; return 0x0Bu; }
    if (orig <= 65535u && n <= 65535u && cl <= 65535u) { *h = 7; return 0x0Au; }
    *h = 13;
    return 0x06u;
}

__device__ __forceinline__ uint32_t rd32g(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

__device__ Plan plan_frame(const przm_block_t &bd, const przm_meta_t *mt, const uint8_t *wsb,
                           const przm_offs_t *of, const przm_sheaf_t *sh) {
    Plan P = {};
    const uint32_t n = bd.len;
    if ((bd.flags & PRZM_BF_RAW) || n == 0) { P.kind = 1; return P; }
    /* Section 2: commands.go:371-406 per substream, 409-444 + 537-568 for the section */
    const uint32_t ncmd = mt->len[PRZM_S_CMD];
    const uint32_t slen[3] = {mt->len[PRZM_S_OP], mt->len[PRZM_S_META], mt->len[PRZM_S_VAR]};
    uint32_t pack = 13;
    for (int s = 0; s < 3; s++) {
        uint32_t best = 1 + slen[s], ch = 0;
        const uint32_t em8 = mt->enc_len[1 + 2 * s], ev = mt->enc_len[2 + 2 * s];
        if (slen[s] >= 24 && em8 != PRZM_ENC_OVERFLOW && 24 + em8 < best) { best = 24 + em8; ch = 1; }
        if (slen[s] >= 16 && ev != PRZM_ENC_OVERFLOW) {
            const uint32_t tv = 8 + mt->mat_len[1 + s] + 16 + ev;
            if (tv < best) { best = tv; ch = 2; }
        }
        P.sub_len[s] = best;
        P.sub_ch[s] = ch;
        pack += best;
    }
    if (ncmd == 0) P.l2 = 0;
    else if (pack < ncmd) P.l2 = pack;
    else { P.l2 = 1 + ncmd; P.sub_ch[0] = P.sub_ch[1] = P.sub_ch[2] = 0xFFFFFFFFu; }
    /* Section 3: flat LEB128, or with the prefix option PackageOffsetsSection's choice
     * (distance.go:692-849): the shorter code stream (V-ANS only when strictly shorter than
     * Mode 12, from 64 distances on), and the container only when strictly shorter than the
     * varints and its code stream is not exactly the match count long */
    const uint32_t noff = mt->len[PRZM_S_OFF];
    P.l3 = noff ? 1 + noff : 0;
    P.s3m = noff ? 0x09u : 0u;
    P.s3c = 0;
    P.s3cl = 0;
    if (of && noff) {
        const uint32_t nd = of->n;
        uint32_t cl = of->m12_len == PRZM_ENC_OVERFLOW ? 0xFFFFFFFFu : 5u + of->m12_len, ck = 0;
        if (nd >= 64u && of->v_len != PRZM_ENC_OVERFLOW) {
            const uint32_t tv = 8u + of->mat_len + 16u + of->v_len;
            if (tv < cl) { cl = tv; ck = 1; }
        }
        if (cl != 0xFFFFFFFFu) {
            uint32_t h;
            const uint32_t m = prefix_mode(noff, nd, cl, &h);
            const uint32_t l3 = h + cl + of->extra;
            if (l3 < noff && cl != nd) { P.l3 = l3; P.s3m = m; P.s3c = ck; P.s3cl = cl; }
        }
    }
    /* Section 4: literals.go:368-384 / 664-693 without the sheaf trials, or with the sheaf option
     * the choice of PackageLiteralsSection (przm_sheaf_encode_kernel) */
    const uint32_t nlit = mt->len[PRZM_S_LIT];
    P.s4v = 0;
    P.l4 = nlit;
    if (sh && nlit) {
        P.l4 = sh->len;
        P.s4v = sh->choice != 0 ? 1u : 0u;
    } else if (nlit) {
        const uint32_t ev = mt->enc_len[PRZM_J_LIT];
        if (ev != PRZM_ENC_OVERFLOW) {
            const uint32_t tv = 8 + mt->mat_len[0] + 16 + ev;
            if (tv < nlit) { P.l4 = tv; P.s4v = 1; }
        }
    }
    P.ext = (P.l2 > 65535 || P.l3 > 65535) ? 1 : 0;
    const uint32_t comp = 8 + (P.ext ? 6 : 4) + 8 + P.l2 + P.l3 + P.l4;
    bool hazard = false;
    if (!P.s4v && P.l4 >= 8) {
        const uint8_t *lit = wsb + PRZM_WS_LIT;
        hazard = (lit[0] == 'V' && lit[1] == 'A' && lit[2] == 0x01) || rd32g(lit + 4) == (rd32g(lit) + 1) / 2;
    }
    if (comp >= n) P.kind = 2;
    else if (hazard) P.kind = 3;
    else P.kind = 0;
    P.flags = (P.s4v ? 0x80u : 0u) | (P.ext ? 0x08u : 0u);
    return P;
}

__device__ __forceinline__ uint32_t frame_bytes(const przm_block_t &bd, const Plan &P) {
    if (P.kind != 0) return 8 + bd.len;
    return 8 + (P.ext ? 6 : 4) + 8 + P.l2 + P.l3 + P.l4;
}

/* ows / sws: the blocks' offsets / sheaf workspaces with those options, nullptr without */
extern "C" __global__ void __launch_bounds__(256)
przm_frame_size_kernel(const przm_block_t *__restrict__ blocks, przm_meta_t *__restrict__ meta,
                       const uint8_t *__restrict__ ws, uint32_t nblocks, uint32_t *__restrict__ frame_len,
                       const uint8_t *__restrict__ ows, const uint8_t *__restrict__ sws) {
    const uint32_t blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nblocks) return;
    const przm_block_t bd = blocks[blk];
    przm_meta_t *mt = meta + blk;
    const przm_offs_t *of = ows ? (const przm_offs_t *)(ows + (uint64_t)blk * PRZM_OWS_BYTES + PRZM_OWS_HDR) : nullptr;
    const przm_sheaf_t *sh = sws ? (const przm_sheaf_t *)(sws + (uint64_t)blk * PRZM_SWS_BYTES + PRZM_SWS_HDR) : nullptr;
    const Plan P = plan_frame(bd, mt, ws + (uint64_t)blk * PRZM_WS_BYTES, of, sh);
    const uint32_t fl = frame_bytes(bd, P);
    mt->frame_len = fl;
    mt->frame_kind = P.kind;
    for (int s = 0; s < 3; s++) mt->choice[s] = P.sub_ch[s];
    mt->s_len[0] = P.l2; mt->s_len[1] = P.l3; mt->s_len[2] = P.l4;
    mt->s4_vans = P.s4v;
    mt->s3_mode = P.s3m;
    mt->s3_code = P.s3c;
    mt->s3_code_len = P.s3cl;
    /* a PRZS chunk also takes its length prefix, and the stream's first chunk the header */
    frame_len[blk] = fl + ((bd.flags & PRZM_BF_CHUNK) ? PRZM_PRZS_PREFIX : 0u) +
                     ((bd.flags & PRZM_BF_HEAD) ? PRZM_PRZS_HDR : 0u);
}

/* Exclusive scan of frame lengths into 64-bit offsets; base/total carried in *run
 * so consecutive chunks continue the numbering.  One workgroup of 1024 threads. */
extern "C" __global__ void __launch_bounds__(1024)
przm_scan_kernel(const uint32_t *__restrict__ len, uint32_t count, uint64_t *__restrict__ off,
                 uint64_t *__restrict__ run) {
    __shared__ uint64_t part[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (count + 1023) / 1024;
    const uint32_t b = tid * per, e = b + per < count ? b + per : count;
    uint64_t s = 0;
    for (uint32_t i = b; i < e; i++) s += len[i];
    part[tid] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint64_t y = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += y;
        __syncthreads();
    }
    const uint64_t base = *run;
    uint64_t acc = base + part[tid] - s;
    for (uint32_t i = b; i < e; i++) { off[i] = acc; acc += len[i]; }
    __syncthreads();
    if (tid == 1023) *run = base + part[1023];
}

/* dst[0, n) = src[0, n) by the workgroup: bytes up to dst's next 16-byte boundary and after
 * the last full 16 bytes one by one, the rest as aligned 16-byte stores whose bytes come from
 * aligned source words shifted into place (a source word read is one that holds a byte of
 * the range, so it lies inside the source's allocation) */
__device__ __forceinline__ void wr_bytes(uint8_t *dst, const uint8_t *src, uint32_t n) {
    const uint32_t t = threadIdx.x, nt = blockDim.x;
    const uint32_t lead = (16u - (uint32_t)((uintptr_t)dst & 15u)) & 15u;
    const uint32_t head = lead < n ? lead : n;
    const uint32_t body = (n - head) & ~15u;
    for (uint32_t i = t; i < head; i += nt) dst[i] = src[i];
    for (uint32_t i = head + body + t; i < n; i += nt) dst[i] = src[i];
    if (!body) return;
    uint4 *d = (uint4 *)(dst + head);
    const uint8_t *s = src + head;
    const uint32_t sh = (uint32_t)((uintptr_t)s & 3u);
    const uint32_t *sw = (const uint32_t *)(s - sh);
    if (sh == 0) {
        for (uint32_t i = t; i < (body >> 4); i += nt) {
            const uint32_t *w = sw + 4 * i;
            d[i] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    } else {
        for (uint32_t i = t; i < (body >> 4); i += nt) {
            const uint32_t *w = sw + 4 * i;
            const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4];
            d[i] = make_uint4(__builtin_amdgcn_alignbyte(w1, w0, sh), __builtin_amdgcn_alignbyte(w2, w1, sh),
                              __builtin_amdgcn_alignbyte(w3, w2, sh), __builtin_amdgcn_alignbyte(w4, w3, sh));
        }
    }
}
__device__ __forceinline__ void wr_le32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24);
}

/* the V-ANS container of job `jk` with matrix of ftab entry `t`: 8 + mat + 16 + stream */
__device__ uint32_t put_vans(uint8_t *o, const uint8_t *wsb, const przm_meta_t *mt, const przm_ftab_t *ft, uint32_t jk,
                             uint32_t n, uint32_t t, const uint8_t *slot_end) {
    const uint32_t ml = mt->mat_len[t], sl = mt->enc_len[jk];
    if (threadIdx.x == 0) {
        o[0] = 'V'; o[1] = 'A'; o[2] = 0x01; o[3] = 0x00;
        wr_le32(o + 4, n);
        for (int j = 0; j < 4; j++) wr_le32(o + 8 + ml + 4 * j, mt->enc_state[jk][j]);
    }
    wr_bytes(o + 8, ft[t].mat, ml);
    wr_bytes(o + 8 + ml + 16, slot_end - sl, sl);
    return 8 + ml + 16 + sl;
}

__device__ uint32_t put_meta8(uint8_t *o, const przm_meta_t *mt, uint32_t jk, uint32_t n, const uint8_t *slot_end) {
    const uint32_t sl = mt->enc_len[jk];
    if (threadIdx.x == 0) {
        o[0] = 'V'; o[1] = 'A'; o[2] = 0x01; o[3] = 0x41;
        wr_le32(o + 4, n);
        for (int j = 0; j < 4; j++) wr_le32(o + 8 + 4 * j, mt->enc_state[jk][j]);
    }
    wr_bytes(o + 24, slot_end - sl, sl);
    return 24 + sl;
}

extern "C" __global__ void __launch_bounds__(256)
przm_frame_write_kernel(const uint8_t *__restrict__ src, const przm_block_t *__restrict__ blocks,
                        const przm_meta_t *__restrict__ meta, const uint8_t *__restrict__ ws,
                        const przm_ftab_t *__restrict__ ftab, const uint64_t *__restrict__ frame_off,
                        uint8_t *__restrict__ out, uint32_t nblocks, const uint8_t *__restrict__ ows,
                        const uint8_t *__restrict__ sws) {
    const uint32_t blk = blockIdx.x;
    if (blk >= nblocks) return;
    const przm_block_t bd = blocks[blk];
    const przm_meta_t *mt = meta + blk;
    const uint8_t *wsb = ws + (uint64_t)blk * PRZM_WS_BYTES;
    const przm_ftab_t *ft = ftab + (uint64_t)blk * 4;
    uint8_t *o = out + frame_off[blk];
    const uint32_t n = bd.len;
    if (bd.flags & PRZM_BF_HEAD) { /* the PRZS stream header (decompress.go:849-866) */
        if (threadIdx.x == 0) {
            /* the buffer's blocks follow this one in the table, all full but the last */
            const uint32_t cnt = bd.flags >> PRZM_BF_COUNT_SHIFT;
            const uint64_t total = (uint64_t)(cnt - 1u) * PRZM_BLOCK + blocks[blk + cnt - 1u].len;
            o[0] = 'P'; o[1] = 'R'; o[2] = 'Z'; o[3] = 'S'; o[4] = 0;
            wr_le32(o + 5, cnt);
            wr_le32(o + 9, (uint32_t)total);
            wr_le32(o + 13, (uint32_t)(total >> 32));
        }
        o += PRZM_PRZS_HDR;
    }
    if (bd.flags & PRZM_BF_CHUNK) { /* the chunk's length (cmd/przm-viz/analyzer.go:243-246) */
        if (threadIdx.x == 0) wr_le32(o, mt->frame_len);
        o += PRZM_PRZS_PREFIX;
    }
    if (mt->frame_kind != 0) { /* compress.go:973-981 */
        if (threadIdx.x == 0) {
            o[0] = 'P'; o[1] = 'R'; o[2] = 'Z'; o[3] = 'M'; o[4] = 0x04;
            o[5] = (uint8_t)n; o[6] = (uint8_t)(n >> 8); o[7] = (uint8_t)(n >> 16);
        }
        wr_bytes(o + 8, src + bd.src_off, n);
        return;
    }
    const uint32_t l2 = mt->s_len[0], l3 = mt->s_len[1], l4 = mt->s_len[2];
    const bool ext = l2 > 65535 || l3 > 65535;
    uint32_t p = 0;
    if (threadIdx.x == 0) { /* compress.go:988-1033 */
        o[0] = 'P'; o[1] = 'R'; o[2] = 'Z'; o[3] = 'M';
        o[4] = (uint8_t)((mt->s4_vans ? 0x80u : 0u) | (ext ? 0x08u : 0u));
        o[5] = (uint8_t)n; o[6] = (uint8_t)(n >> 8); o[7] = (uint8_t)(n >> 16);
        if (ext) {
            o[8] = (uint8_t)l2; o[9] = (uint8_t)(l2 >> 8); o[10] = (uint8_t)(l2 >> 16);
            o[11] = (uint8_t)l3; o[12] = (uint8_t)(l3 >> 8); o[13] = (uint8_t)(l3 >> 16);
        } else {
            o[8] = (uint8_t)l2; o[9] = (uint8_t)(l2 >> 8); o[10] = (uint8_t)l3; o[11] = (uint8_t)(l3 >> 8);
        }
    }
    p = ext ? 14 : 12;
    if (threadIdx.x < 8) {
        uint32_t wl = n <= 1024 ? 10u : n <= 4096 ? 12u : n <= 16384 ? 14u : 16u; /* compress.go:183-202 */
        o[p + threadIdx.x] = threadIdx.x == 1 ? (uint8_t)wl : 0;
    }
    p += 8;
    /* Section 2 */
    const uint32_t ncmd = mt->len[PRZM_S_CMD];
    if (l2) {
        if (mt->choice[0] == 0xFFFFFFFFu) {
            if (threadIdx.x == 0) o[p] = 0x09;
            wr_bytes(o + p + 1, wsb + PRZM_WS_CMD, ncmd);
        } else {
            uint8_t *s2 = o + p;
            uint32_t q = 13;
            const uint32_t srcoff[3] = {PRZM_WS_OP, PRZM_WS_META, PRZM_WS_VAR};
            const uint32_t slens[3] = {mt->len[PRZM_S_OP], mt->len[PRZM_S_META], mt->len[PRZM_S_VAR]};
            const uint32_t m8end[3] = {PRZM_WS_ENC_OP_M8 + PRZM_ENC_OP_CAP, PRZM_WS_ENC_META_M8 + PRZM_ENC_META_CAP,
                                       PRZM_WS_ENC_VAR_M8 + PRZM_ENC_VAR_CAP};
            const uint32_t vend[3] = {PRZM_WS_ENC_OP_V + PRZM_ENC_OP_CAP, PRZM_WS_ENC_META_V + PRZM_ENC_META_CAP,
                                      PRZM_WS_ENC_VAR_V + PRZM_ENC_VAR_CAP};
            uint32_t lop = 0, lm = 0;
            for (int s = 0; s < 3; s++) {
                uint32_t l;
                if (mt->choice[s] == 1) l = put_meta8(s2 + q, mt, 1 + 2 * s, slens[s], wsb + m8end[s]);
                else if (mt->choice[s] == 2) l = put_vans(s2 + q, wsb, mt, ft, 2 + 2 * s, slens[s], 1 + s, wsb + vend[s]);
                else {
                    if (threadIdx.x == 0) s2[q] = 0x09;
                    wr_bytes(s2 + q + 1, wsb + srcoff[s], slens[s]);
                    l = 1 + slens[s];
                }
                if (s == 0) lop = l;
                if (s == 1) lm = l;
                q += l;
            }
            if (threadIdx.x == 0) {
                s2[0] = 0x08;
                wr_le32(s2 + 1, ncmd);
                wr_le32(s2 + 5, lop);
                wr_le32(s2 + 9, lm);
            }
        }
        p += l2;
    }
    /* Section 3 */
    if (l3 && mt->s3_mode == 0x09u) {
        if (threadIdx.x == 0) o[p] = 0x09;
        wr_bytes(o + p + 1, wsb + PRZM_WS_OFF, l3 - 1);
        p += l3;
    } else if (l3) { /* assemblePrefixContainer (distance.go:777-823) */
        const uint8_t *ob = ows + (uint64_t)blk * PRZM_OWS_BYTES;
        const przm_offs_t *of = (const przm_offs_t *)(ob + PRZM_OWS_HDR);
        const uint32_t m = mt->s3_mode, cl = mt->s3_code_len, orig = mt->len[PRZM_S_OFF], nd = of->n;
        const uint32_t h = m == 0x0Bu ? 4u : m == 0x0Au ? 7u : 13u;
        uint8_t *s3 = o + p;
        if (threadIdx.x == 0) {
            s3[0] = (uint8_t)m;
            if (m == 0x0Bu) {
                s3[1] = (uint8_t)orig; s3[2] = (uint8_t)nd; s3[3] = (uint8_t)cl;
            } else if (m == 0x0Au) {
                s3[1] = (uint8_t)orig; s3[2] = (uint8_t)(orig >> 8); s3[3] = (uint8_t)nd; s3[4] = (uint8_t)(nd >> 8);
                s3[5] = (uint8_t)cl; s3[6] = (uint8_t)(cl >> 8);
            } else {
                wr_le32(s3 + 1, orig); wr_le32(s3 + 5, nd); wr_le32(s3 + 9, cl);
            }
        }
        uint8_t *cs = s3 + h;
        if (mt->s3_code == 0) { /* [0x0C][u32 state][stream] (distance.go:364-375) */
            if (threadIdx.x == 0) { cs[0] = 0x0C; wr_le32(cs + 1, of->m12_state); }
            wr_bytes(cs + 5, ob + PRZM_OWS_M12 + PRZM_OWS_M12_CAP - of->m12_len, of->m12_len);
        } else { /* 'V','A',1,0, u32 n, matrix, four states, stream */
            const uint32_t ml = of->mat_len, sl = of->v_len;
            if (threadIdx.x == 0) {
                cs[0] = 'V'; cs[1] = 'A'; cs[2] = 0x01; cs[3] = 0x00;
                wr_le32(cs + 4, nd);
                for (int j = 0; j < 4; j++) wr_le32(cs + 8 + ml + 4 * j, of->v_state[j]);
            }
            wr_bytes(cs + 8, ((const przm_ftab_t *)(ob + PRZM_OWS_FTAB))->mat, ml);
            wr_bytes(cs + 8 + ml + 16, ob + PRZM_OWS_VANS + PRZM_OWS_VANS_CAP - sl, sl);
        }
        wr_bytes(cs + cl, ob + PRZM_OWS_EXTRA, of->extra);
        p += l3;
    }
    /* Section 4 */
    const uint8_t *sb = sws ? sws + (uint64_t)blk * PRZM_SWS_BYTES : nullptr;
    const uint32_t ch = sb ? ((const przm_sheaf_t *)(sb + PRZM_SWS_HDR))->choice : (mt->s4_vans ? 1u : 0u);
    if (l4 && ch == 1) {
        put_vans(o + p, wsb, mt, ft, PRZM_J_LIT, mt->len[PRZM_S_LIT], 0, wsb + PRZM_WS_ENC_LIT + PRZM_ENC_LIT_CAP);
    } else if (l4 && ch >= 2) { /* assembleSheafContainer (literals.go:645-664) */
        const przm_sheaf_t *H = (const przm_sheaf_t *)(sb + PRZM_SWS_HDR);
        const uint32_t m = ch - 2;
        uint8_t *s4 = o + p;
        if (threadIdx.x == 0) {
            s4[0] = 'V'; s4[1] = 'A'; s4[2] = 0x01; s4[3] = (uint8_t)(0x05u | (m << 4));
            wr_le32(s4 + 4, H->n);
            for (uint32_t k = 0; k < 4; k++) {
                const bool coded = H->l[m][k] && H->k[m][k] != PRZM_ENC_OVERFLOW;
                wr_le32(s4 + 8 + 4 * k, H->l[m][k]);
                wr_le32(s4 + 24 + 4 * k, coded ? H->mat[m][k] + 16u + H->k[m][k] : H->l[m][k]);
            }
        }
        const uint8_t *fib = sb + PRZM_SWS_FIB + (uint64_t)m * PRZM_SWS_FIB_CAP;
        const przm_ftab_t *sft = (const przm_ftab_t *)(sb + PRZM_SWS_FTAB) + 4u * m;
        uint32_t at = 40, fs = 0, end = 0;
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t lk = H->l[m][k];
            end += (lk + 64u + 15u) & ~15u;
            if (!lk) continue;
            if (H->k[m][k] != PRZM_ENC_OVERFLOW) { /* matrix, four states, stream (EncodeVANS4x) */
                /* the counted length, which the frame's size was planned with (the final coding
                 * of the fiber emits the same bytes: the tests compare them) */
                const uint32_t ml = H->mat[m][k], ek = H->k[m][k];
                if (threadIdx.x == 0)
                    for (int j = 0; j < 4; j++) wr_le32(s4 + at + ml + 4 * j, H->st[k][j]);
                wr_bytes(s4 + at, sft[k].mat, ml);
                wr_bytes(s4 + at + ml + 16, sb + PRZM_SWS_OUT + end - ek, ek);
                at += ml + 16 + ek;
            } else {
                wr_bytes(s4 + at, fib + fs, lk);
                at += lk;
            }
            fs += (lk + 15u) & ~15u;
        }
    } else if (l4) {
        wr_bytes(o + p, wsb + PRZM_WS_LIT, l4);
    }
}

/* ---------------- stage kernels over arbitrary device buffers ---------------- */

/* dst[0, n) = src[0, n) by the whole grid (thread t of nt): as wr_bytes, with aligned 16-byte
 * stores built from aligned source words */
__device__ __forceinline__ void wr_bytes_grid(uint8_t *dst, const uint8_t *src, uint64_t n, uint64_t t, uint64_t nt) {
    const uint64_t lead = (16u - (uint32_t)((uintptr_t)dst & 15u)) & 15u;
    const uint64_t head = lead < n ? lead : n;
    const uint64_t body = (n - head) & ~15ull;
    for (uint64_t i = t; i < head; i += nt) dst[i] = src[i];
    for (uint64_t i = head + body + t; i < n; i += nt) dst[i] = src[i];
    if (!body) return;
    uint4 *d = (uint4 *)(dst + head);
    const uint8_t *s = src + head;
    const uint32_t sh = (uint32_t)((uintptr_t)s & 3u);
    const uint32_t *sw = (const uint32_t *)(s - sh);
    const uint64_t n16 = body >> 4;
    if (sh == 0) {
        for (uint64_t i = t; i < n16; i += nt) {
            const uint32_t *w = sw + 4 * i;
            d[i] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    } else {
        for (uint64_t i = t; i < n16; i += nt) {
            const uint32_t *w = sw + 4 * i;
            const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4];
            d[i] = make_uint4(__builtin_amdgcn_alignbyte(w1, w0, sh), __builtin_amdgcn_alignbyte(w2, w1, sh),
                              __builtin_amdgcn_alignbyte(w3, w2, sh), __builtin_amdgcn_alignbyte(w4, w3, sh));
        }
    }
}

/* 256-bin byte histogram of src[0, n) over the whole grid: one LDS copy per wave, 16-byte
 * loads from the first 16-byte boundary on, the copies summed and added to counts (which the
 * caller cleared) with one atomic per nonzero bin and workgroup */
extern "C" __global__ void __launch_bounds__(256)
przm_hist_grid_kernel(const uint8_t *__restrict__ src, uint64_t n, uint32_t *__restrict__ counts) {
    __shared__ uint32_t cnt[4][256];
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    for (uint32_t i = tid; i < 4 * 256; i += 256) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const uint64_t gt = (uint64_t)blockIdx.x * 256 + tid, gn = (uint64_t)gridDim.x * 256;
    const uint64_t lead = (16u - (uint32_t)((uintptr_t)src & 15u)) & 15u;
    const uint64_t head = lead < n ? lead : n;
    const uint64_t n16 = (n - head) >> 4;
    if (gt < head) atomicAdd(&cnt[w][src[gt]], 1u);
    for (uint64_t i = head + (n16 << 4) + gt; i < n; i += gn) atomicAdd(&cnt[w][src[i]], 1u);
    const uint4 *p4 = (const uint4 *)(src + head);
    for (uint64_t i = gt; i < n16; i += gn) {
        const uint4 a = p4[i];
        const uint32_t wd[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            atomicAdd(&cnt[w][wd[k] & 0xFFu], 1u);
            atomicAdd(&cnt[w][(wd[k] >> 8) & 0xFFu], 1u);
            atomicAdd(&cnt[w][(wd[k] >> 16) & 0xFFu], 1u);
            atomicAdd(&cnt[w][wd[k] >> 24], 1u);
        }
    }
    __syncthreads();
    const uint32_t c = cnt[0][tid] + cnt[1][tid] + cnt[2][tid] + cnt[3][tid];
    if (c) atomicAdd(&counts[tid], c);
}

/* the normalised table (freq | cdf << 16) and joint-nibble matrix of a whole-buffer histogram */
extern "C" __global__ void __launch_bounds__(64)
przm_normalize_kernel(const uint32_t *__restrict__ counts, przm_ftab_t *__restrict__ ft, uint32_t *__restrict__ mat_len) {
    __shared__ uint32_t cnt[256];
    for (uint32_t s = threadIdx.x; s < 256; s += 64) cnt[s] = counts[s];
    __syncthreads();
    normalize_wave(cnt, ft, mat_len);
}

/* ModeFlat section [0x00][u32 n][payload] (PackStreamFlat, gpu/kernel/przm_ans.cu:202-234) */
extern "C" __global__ void __launch_bounds__(256)
przm_pack_flat_kernel(const uint8_t *__restrict__ src, uint64_t n, uint8_t *__restrict__ out) {
    const uint64_t gt = (uint64_t)blockIdx.x * 256 + threadIdx.x, gn = (uint64_t)gridDim.x * 256;
    if (gt == 0) {
        out[0] = 0x00;
        wr_le32(out + 1, (uint32_t)n);
    }
    wr_bytes_grid(out + 5, src, n, gt, gn);
}

/* the V-ANS container ('V','A',1,0, u32 n, matrix, four states, stream) of a coder run whose
 * stream ends at slot_end: res = {stream bytes, x0, x1, x2, x3} */
extern "C" __global__ void __launch_bounds__(256)
przm_vans_assemble_kernel(const uint8_t *__restrict__ slot_end, const uint32_t *__restrict__ res,
                          const przm_ftab_t *__restrict__ ft, uint32_t mat_len, uint32_t n, uint8_t *__restrict__ out) {
    const uint32_t sl = res[0];
    if (threadIdx.x == 0) {
        out[0] = 'V'; out[1] = 'A'; out[2] = 0x01; out[3] = 0x00;
        wr_le32(out + 4, n);
        for (int j = 0; j < 4; j++) wr_le32(out + 8 + mat_len + 4 * j, res[1 + j]);
    }
    wr_bytes(out + 8, ft->mat, mat_len);
    wr_bytes(out + 8 + mat_len + 16, slot_end - sl, sl);
}

/* every block's cmd, off and lit streams to three device arrays at the given offsets
 * (dst_off[3 * b + k]): one workgroup per block */
extern "C" __global__ void __launch_bounds__(256)
przm_gather_streams_kernel(const uint8_t *__restrict__ ws, const przm_meta_t *__restrict__ meta,
                           const uint64_t *__restrict__ dst_off, uint8_t *__restrict__ d_cmd,
                           uint8_t *__restrict__ d_off, uint8_t *__restrict__ d_lit, uint32_t nblocks) {
    const uint32_t b = blockIdx.x;
    if (b >= nblocks) return;
    const uint8_t *wsb = ws + (uint64_t)b * PRZM_WS_BYTES;
    const przm_meta_t *mt = meta + b;
    wr_bytes(d_cmd + dst_off[3 * b], wsb + PRZM_WS_CMD, mt->len[PRZM_S_CMD]);
    wr_bytes(d_off + dst_off[3 * b + 1], wsb + PRZM_WS_OFF, mt->len[PRZM_S_OFF]);
    wr_bytes(d_lit + dst_off[3 * b + 2], wsb + PRZM_WS_LIT, mt->len[PRZM_S_LIT]);
}

/* the coder probe's input: block b's literal stream = src[65536 b, ...), its other streams
 * empty, so the histogram and coder kernels run one V-ANS job per block */
extern "C" __global__ void __launch_bounds__(256)
przm_probe_fill_kernel(const uint8_t *__restrict__ src, uint64_t n, uint8_t *__restrict__ ws,
                       przm_meta_t *__restrict__ meta, uint32_t nblocks) {
    const uint32_t b = blockIdx.x;
    if (b >= nblocks) return;
    const uint64_t o = (uint64_t)b * PRZM_BLOCK;
    const uint32_t len = (uint32_t)(n - o < PRZM_BLOCK ? n - o : PRZM_BLOCK);
    if (threadIdx.x < PRZM_S_N) meta[b].len[threadIdx.x] = threadIdx.x == PRZM_S_LIT ? len : 0u;
    wr_bytes(ws + (uint64_t)b * PRZM_WS_BYTES + PRZM_WS_LIT, src + o, len);
}
