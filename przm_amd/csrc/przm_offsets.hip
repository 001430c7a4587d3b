/*
 * przm_offsets.hip -- Section 3 as a distance-prefix container on gfx950: the
 * encoder option that writes each frame's offsets as PackageOffsetsSection
 * chooses them (core/distance/distance.go:692-849) instead of flat LEB128.
 *
 *  przm_offs_prep_kernel  one workgroup per block: the offset stream's LEB128
 *                         varints are found 256 bytes at a time (a byte without
 *                         its top bit ends a varint; its thread gathers the
 *                         continuation bytes before it), numbered by a
 *                         workgroup scan, and turned into distance codes and
 *                         extra bits (DistanceToCode, distance.go:165-180).  The
 *                         extra bits go MSB first into an LDS bit array at the
 *                         positions a second scan gives (BitWriter,
 *                         distance.go:217-247); the codes' histogram feeds
 *                         NormalizeFrequencies for the V-ANS alternative.
 *  przm_offs_code_kernel  both code streams of the codes, in one launch:
 *                         the Mode-12 stream (distance.go:323-376), one rANS
 *                         state whose table is picked by the class of the
 *                         previous code, one lane per block, written backward
 *                         from the end of the block's slot; and, for blocks
 *                         with 64 or more distances, the adaptive V-ANS stream
 *                         (the 4-state coder of the literal sections,
 *                         literals.go:368-384), one lane quad per block.
 *
 * The frame kernels then compare the code streams, pick the container mode
 * (SelectPrefixMode, distance.go:750-758) and take the container only when it
 * is shorter than the varints (distance.go:837), and not when its code stream
 * is exactly as long as the match count (both reference decoders read such a
 * code stream as raw codes, distance.go:601-604, src/przm_dec.c:1097-1098).
 */
#include "przm_dist_freq.h"

/* freq | cdf << 16 per (context, code) of the Mode-12 tables */
struct DistTable {
    uint32_t fc[4][64];
    constexpr DistTable() : fc() {
        for (int c = 0; c < 4; c++) {
            uint32_t cum = 0;
            for (int s = 0; s < 64; s++) {
                fc[c][s] = (uint32_t)przm_dist_freq[c][s] | (cum << 16);
                cum += przm_dist_freq[c][s];
            }
        }
    }
};
__constant__ DistTable przm_dt = DistTable();

/* the context class of the previous code (distance.go:190-202) */
__device__ __forceinline__ uint32_t dist_ctx(uint32_t prev) {
    return prev < 6u ? 0u : prev < 14u ? 1u : prev < 26u ? 2u : 3u;
}

#define OFFS_NT 256u
#define OFFS_EXT_WORDS (PRZM_OWS_EXTRA_CAP / 4u)

/* exclusive scan of v over the workgroup of OFFS_NT threads; *total gets the sum.
 * part: LDS scratch of OFFS_NT / 64 words */
__device__ __forceinline__ uint32_t wg_scan_excl(uint32_t v, uint32_t *part, uint32_t *total) {
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t inc = wave_scan_incl(v);
    if (lane == 63u) part[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t k = 0; k < OFFS_NT / 64u; k++) {
        const uint32_t p = part[k];
        before += k < w ? p : 0u;
        all += p;
    }
    __syncthreads(); /* part is reused by the next call */
    *total = all;
    return before + inc - v;
}

/* bytes of a 32-bit word, little-endian: byte j of w */
#define WBYTE(w, j) (((w) >> (8u * (j))) & 0xFFu)

/*
 * One workgroup per block; each thread takes 4 bytes of the offset stream per round (1024
 * bytes a round).  A byte without its top bit ends a varint: its thread reads the varint's
 * continuation bytes from the 8-byte window of its own word and the word before it (a varint
 * of a distance below 2^28 has at most 4 bytes).  One workgroup scan per round numbers the
 * varints and places their extra bits (count << 20 | bits: a round has at most 1024 varints
 * and 1024 * 14 extra bits).
 */
extern "C" __global__ void __launch_bounds__(OFFS_NT)
przm_offs_prep_kernel(const uint8_t *__restrict__ ws, const przm_meta_t *__restrict__ meta, uint8_t *__restrict__ ows,
                      uint32_t nblocks) {
    __shared__ __align__(16) uint32_t ext[OFFS_EXT_WORDS];
    __shared__ uint32_t cnt[256];
    __shared__ uint32_t part[OFFS_NT / 64u];
    const uint32_t blk = blockIdx.x, tid = threadIdx.x;
    if (blk >= nblocks) return;
    const uint8_t *off = ws + (uint64_t)blk * PRZM_WS_BYTES + PRZM_WS_OFF;
    const uint32_t *offw = (const uint32_t *)off;
    const uint32_t noff = meta[blk].len[PRZM_S_OFF];
    uint8_t *ob = ows + (uint64_t)blk * PRZM_OWS_BYTES;
    przm_offs_t *H = (przm_offs_t *)(ob + PRZM_OWS_HDR);
    uint8_t *codes = ob + PRZM_OWS_CODES;
    for (uint32_t i = tid; i < OFFS_EXT_WORDS / 4u; i += OFFS_NT) ((uint4 *)ext)[i] = make_uint4(0, 0, 0, 0);
    for (uint32_t i = tid; i < 256; i += OFFS_NT) cnt[i] = 0;
    __syncthreads();
    uint32_t nd = 0, nbits = 0;
    /* own word (bytes at or past noff read as 0x80: no varint ends there) and the word before it
     * (none before the stream: 0, so a walk back stops); the next round's are loaded a round ahead */
    auto load_cur = [&](uint32_t p0) -> uint32_t { return p0 < noff ? offw[p0 >> 2] : 0x80808080u; };
    auto load_prev = [&](uint32_t p0) -> uint32_t { return p0 >= 4u && p0 < noff ? offw[(p0 >> 2) - 1u] : 0u; };
    uint32_t ncur = load_cur(4u * tid), nprev = load_prev(4u * tid);
    for (uint32_t base = 0; base < noff; base += 4u * OFFS_NT) {
        const uint32_t p0 = base + 4u * tid;
        uint32_t cur = ncur;
        const uint32_t prevw = nprev;
        ncur = load_cur(p0 + 4u * OFFS_NT);
        nprev = load_prev(p0 + 4u * OFFS_NT);
        if (p0 < noff && noff - p0 < 4u) {
            const uint32_t valid = noff - p0;
            cur = (cur & ((1u << (8u * valid)) - 1u)) | (0x80808080u << (8u * valid));
        }
        const uint64_t win = ((uint64_t)cur << 32) | prevw; /* bytes p0 - 4 .. p0 + 3 */
        uint32_t code[4], extra[4], nbk[4], term = 0, nt = 0, nbt = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            code[j] = 0; extra[j] = 0; nbk[j] = 0;
            const uint32_t b = WBYTE(cur, j);
            if (b & 0x80u) continue;
            uint32_t d = b;
#pragma unroll
            for (uint32_t k = 1; k <= 4; k++) { /* window byte 4 + j - k */
                const uint32_t c = (uint32_t)(win >> (8u * (4u + j - k))) & 0xFFu;
                if (!(c & 0x80u)) break;
                d = (d << 7) | (c & 0x7Fu);
            }
            if (d <= 4u) {
                code[j] = d - 1u;
            } else {
                const uint32_t u = d - 1u;
                const uint32_t nb = 30u - (uint32_t)__builtin_clz(u);
                const uint32_t half = (u >> nb) & 1u;
                code[j] = 2u * nb + 2u + half;
                extra[j] = d - (((2u + half) << nb) + 1u);
                nbk[j] = nb;
            }
            term |= 1u << j;
            nt++;
            nbt += nbk[j];
        }
        uint32_t tot;
        const uint32_t ex = wg_scan_excl((nt << 20) | nbt, part, &tot);
        uint32_t idx = nd + (ex >> 20), bit = nbits + (ex & 0xFFFFFu);
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            if (!(term >> j & 1u)) continue;
            const uint32_t nb = nbk[j];
            /* a block holds at most PRZM_OFFS_MAX distances; a longer stream keeps its flat
             * section (the coder kernels skip it), so nothing past the slots is written */
            if (idx < PRZM_OFFS_MAX && bit + nb <= 32u * (OFFS_EXT_WORDS - 1u)) {
                codes[idx] = (uint8_t)code[j];
                atomicAdd(&cnt[code[j] & 255u], 1u); /* codes are <= 48 for the distances of a block */
                if (nb) {
                    /* bit k of the stream is bit 31 - (k & 31) of word k >> 5 (bytes big-endian) */
                    const uint32_t w = bit >> 5, o = bit & 31u;
                    if (o + nb <= 32u) {
                        atomicOr(&ext[w], extra[j] << (32u - o - nb));
                    } else {
                        atomicOr(&ext[w], extra[j] >> (o + nb - 32u));
                        atomicOr(&ext[w + 1], extra[j] << (64u - o - nb));
                    }
                }
            }
            idx++;
            bit += nb;
        }
        nd += tot >> 20;
        nbits += tot & 0xFFFFFu;
    }
    __syncthreads();
    const uint32_t eb = nd <= PRZM_OFFS_MAX ? (nbits + 7u) >> 3 : 0u;
    uint32_t *ew = (uint32_t *)(ob + PRZM_OWS_EXTRA);
    for (uint32_t w = tid; w < (eb + 3u) >> 2; w += OFFS_NT) ew[w] = __builtin_bswap32(ext[w]);
    if (tid == 0) {
        H->n = nd;
        H->orig = noff;
        H->extra = eb;
        H->m12_len = PRZM_ENC_OVERFLOW;
        H->v_len = PRZM_ENC_OVERFLOW;
        H->mat_len = 0;
    }
    if (nd >= 64u && nd <= PRZM_OFFS_MAX && tid < 64u)
        normalize_wave(cnt, (przm_ftab_t *)(ob + PRZM_OWS_FTAB), &H->mat_len);
}

/*
 * The Mode-12 stream of block blockIdx.x * 64 + lane's codes (distance.go:323-376), one lane
 * per block.  The codes are read 16 at a time with the next two chunks in flight, so a step
 * waits on its table entry (LDS) and its own arithmetic only.  Emitted byte k belongs at
 * end[-1 - k]: four bytes are gathered into one aligned 32-bit store (end is 16-byte aligned).
 * The renormalisation emits at most two bytes per step (x < 2^24 and 4096 f - 1 >= 4095).
 */
__device__ void offs_m12(uint8_t *__restrict__ ows, uint32_t nblocks, uint32_t group, uint32_t *tab) {
    for (uint32_t i = threadIdx.x; i < 256; i += 64) tab[i] = (&przm_dt.fc[0][0])[i];
    __syncthreads();
    const uint32_t blk = group * 64u + threadIdx.x;
    if (blk >= nblocks) return;
    uint8_t *ob = ows + (uint64_t)blk * PRZM_OWS_BYTES;
    przm_offs_t *H = (przm_offs_t *)(ob + PRZM_OWS_HDR);
    if (H->n > PRZM_OFFS_MAX) return; /* m12_len stays PRZM_ENC_OVERFLOW: the section stays flat */
    const uint32_t n = H->n;
    const uint4 *c4 = (const uint4 *)(ob + PRZM_OWS_CODES);
    uint32_t *const endw = (uint32_t *)(ob + PRZM_OWS_M12 + PRZM_OWS_M12_CAP);
    /* a chunk emits at most 32 bytes, so stopping once past cap writes at most cap + 32 */
    const uint32_t cap = PRZM_OWS_M12_CAP - 48u;
    const uint4 z = make_uint4(0, 0, 0, 0);
    const int32_t nc = (int32_t)((n + 15u) >> 4);
    uint4 cur = nc > 0 ? c4[nc - 1] : z, nxt = nc > 1 ? c4[nc - 2] : z;
    uint32_t x = 65536u, K = 0, acc = 0;
    bool over = false;
#define M12_EMIT(b)                                              \
    do {                                                         \
        acc |= (b) << (24u - 8u * (K & 3u));                     \
        if ((K & 3u) == 3u) { endw[-1 - (int32_t)(K >> 2)] = acc; acc = 0; } \
        K++;                                                     \
    } while (0)
    for (int32_t c = nc - 1; c >= 0; c--) {
        const uint4 pre = c > 1 ? c4[c - 2] : z;
        const uint32_t w[4] = {cur.x, cur.y, cur.z, cur.w};
        const uint32_t below = c > 0 ? nxt.w >> 24 : 0u; /* code 16 c - 1 (0 before the first) */
#pragma unroll
        for (int32_t k = 15; k >= 0; k--) {
            if ((uint32_t)(16 * c + k) >= n) continue;
            const uint32_t sym = WBYTE(w[k >> 2], (uint32_t)k & 3u);
            const uint32_t prev = k ? WBYTE(w[(k - 1) >> 2], (uint32_t)(k - 1) & 3u) : below;
            const uint32_t e = tab[dist_ctx(prev) * 64u + sym];
            const uint32_t f = e & 0xFFFFu, lim = 4096u * f - 1u;
            if (x > lim) {
                M12_EMIT(x & 0xFFu);
                x >>= 8;
                if (x > lim) {
                    M12_EMIT(x & 0xFFu);
                    x >>= 8;
                }
            }
            uint32_t q, r;
            divmod(x, f, &q, &r);
            x = (q << 12) + (e >> 16) + r;
        }
        cur = nxt;
        nxt = pre;
        if (K > cap) { over = true; break; }
    }
#undef M12_EMIT
    /* the last, partial word: its bytes sit at the word's high addresses; the rest of the word
     * lies below the stream and is not part of it */
    if (!over && (K & 3u)) endw[-1 - (int32_t)(K >> 2)] = acc;
    H->m12_len = over ? PRZM_ENC_OVERFLOW : K;
    H->m12_state = x;
}

/* the adaptive V-ANS stream of the codes of block group * 16 + quad (64+ distances) */
__device__ void offs_vans(uint8_t *__restrict__ ows, uint32_t nblocks, uint32_t group, uint32_t (*vt)[256],
                          uint8_t (*stg)[RANS_STAGE + 16]) {
    const uint32_t q = threadIdx.x >> 2, j = threadIdx.x & 3u;
    const uint32_t blk = group * 16u + q;
    const bool has = blk < nblocks;
    uint8_t *ob = ows + (uint64_t)(has ? blk : 0u) * PRZM_OWS_BYTES;
    przm_offs_t *H = (przm_offs_t *)(ob + PRZM_OWS_HDR);
    const uint32_t n0 = has ? H->n : 0u;
    const bool live = n0 >= 64u && n0 <= PRZM_OFFS_MAX;
    const uint32_t n = live ? n0 : 0u;
    if (live) {
        const przm_ftab_t *ft = (const przm_ftab_t *)(ob + PRZM_OWS_FTAB);
        for (uint32_t s = j; s < 256; s += 4) vt[q][s] = ft->fc[s];
    }
    /* every quad of the wave steps as often as the wave's longest job */
    uint32_t gmax = (n + 3u) >> 2;
    for (uint32_t o = 32; o >= 4; o >>= 1) {
        const uint32_t v = (uint32_t)__shfl_xor((int)gmax, (int)o, 64);
        gmax = v > gmax ? v : gmax;
    }
    __syncthreads();
    uint32_t x;
    const uint32_t K = rans_quad<false>(ob + PRZM_OWS_CODES, n, vt[q], stg[q], ob + PRZM_OWS_VANS + PRZM_OWS_VANS_CAP,
                                        PRZM_OWS_VANS_CAP - 48u, gmax, &x);
    if (!live) return;
    if (j == 0) H->v_len = K;
    H->v_state[j] = x;
}

/* both code streams in one launch, so that the few long-running Mode-12 workgroups (one lane
 * per block) and the V-ANS workgroups (one quad per block) overlap: workgroups below m12_groups
 * run the Mode-12 coder over 64 blocks each, the rest the V-ANS coder over 16 blocks each */
struct OffsCoderShared {
    uint32_t vt[16][256];
    __align__(16) uint8_t stg[16][RANS_STAGE + 16];
};
extern "C" __global__ void __launch_bounds__(64)
przm_offs_code_kernel(uint8_t *__restrict__ ows, uint32_t nblocks, uint32_t m12_groups) {
    __shared__ OffsCoderShared S;
    if (blockIdx.x < m12_groups) offs_m12(ows, nblocks, blockIdx.x, S.vt[0]);
    else offs_vans(ows, nblocks, blockIdx.x - m12_groups, S.vt, S.stg);
}
