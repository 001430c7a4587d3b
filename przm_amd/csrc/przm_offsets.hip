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
 *  przm_offs_m12_kernel   one lane per block: the codes' Mode-12 stream
 *                         (distance.go:323-376), one rANS state whose table is
 *                         picked by the class of the previous code, written
 *                         backward from the end of the block's slot.
 *  przm_offs_vans_kernel  one lane quad per block with 64 or more distances:
 *                         the codes' adaptive V-ANS stream (the 4-state coder of
 *                         the literal sections, literals.go:368-384).
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

extern "C" __global__ void __launch_bounds__(OFFS_NT)
przm_offs_prep_kernel(const uint8_t *__restrict__ ws, const przm_meta_t *__restrict__ meta, uint8_t *__restrict__ ows,
                      uint32_t nblocks) {
    __shared__ uint32_t ext[OFFS_EXT_WORDS];
    __shared__ uint32_t cnt[256];
    __shared__ uint32_t key[256];
    __shared__ uint32_t frq[256];
    __shared__ uint32_t part[OFFS_NT / 64u];
    const uint32_t blk = blockIdx.x, tid = threadIdx.x;
    if (blk >= nblocks) return;
    const uint8_t *off = ws + (uint64_t)blk * PRZM_WS_BYTES + PRZM_WS_OFF;
    const uint32_t noff = meta[blk].len[PRZM_S_OFF];
    uint8_t *ob = ows + (uint64_t)blk * PRZM_OWS_BYTES;
    przm_offs_t *H = (przm_offs_t *)(ob + PRZM_OWS_HDR);
    uint8_t *codes = ob + PRZM_OWS_CODES;
    for (uint32_t i = tid; i < OFFS_EXT_WORDS; i += OFFS_NT) ext[i] = 0;
    for (uint32_t i = tid; i < 256; i += OFFS_NT) cnt[i] = 0;
    __syncthreads();
    uint32_t nd = 0, nbits = 0;
    for (uint32_t base = 0; base < noff; base += OFFS_NT) {
        const uint32_t i = base + tid;
        const uint32_t b = i < noff ? off[i] : 0x80u;
        const bool term = (b & 0x80u) == 0;
        uint32_t d = 0, code = 0, extra = 0, nb = 0;
        if (term) {
            /* this byte ends a varint: its continuation bytes are the ones just before it */
            d = b;
            for (uint32_t k = 1; k <= 4u && k <= i; k++) {
                const uint32_t c = off[i - k];
                if (!(c & 0x80u)) break;
                d = (d << 7) | (c & 0x7Fu);
            }
            if (d <= 4u) {
                code = d - 1u;
            } else {
                const uint32_t u = d - 1u;
                nb = 30u - (uint32_t)__builtin_clz(u);
                const uint32_t half = (u >> nb) & 1u;
                code = 2u * nb + 2u + half;
                extra = d - (((2u + half) << nb) + 1u);
            }
        }
        uint32_t nt, bt;
        const uint32_t idx = nd + wg_scan_excl(term ? 1u : 0u, part, &nt);
        const uint32_t bit = nbits + wg_scan_excl(term ? nb : 0u, part, &bt);
        /* a block holds at most PRZM_OFFS_MAX distances; a longer stream keeps its flat section
         * (the coder kernels skip it), so nothing past the slots is written */
        if (term && idx < PRZM_OFFS_MAX && (bit + nb) <= 32u * (OFFS_EXT_WORDS - 1u)) {
            codes[idx] = (uint8_t)code;
            atomicAdd(&cnt[code], 1u);
            if (nb) {
                /* bit k of the stream is bit 31 - (k & 31) of word k >> 5 (bytes big-endian) */
                const uint32_t w = bit >> 5, o = bit & 31u;
                if (o + nb <= 32u) {
                    atomicOr(&ext[w], extra << (32u - o - nb));
                } else {
                    atomicOr(&ext[w], extra >> (o + nb - 32u));
                    atomicOr(&ext[w + 1], extra << (64u - o - nb));
                }
            }
        }
        nd += nt;
        nbits += bt;
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
        normalize_wave(cnt, key, frq, (przm_ftab_t *)(ob + PRZM_OWS_FTAB), &H->mat_len);
}

/* the Mode-12 stream of block blockIdx.x * 64 + lane's codes (distance.go:323-376) */
extern "C" __global__ void __launch_bounds__(64)
przm_offs_m12_kernel(uint8_t *__restrict__ ows, uint32_t nblocks) {
    __shared__ uint32_t tab[256];
    for (uint32_t i = threadIdx.x; i < 256; i += 64) tab[i] = (&przm_dt.fc[0][0])[i];
    __syncthreads();
    const uint32_t blk = blockIdx.x * 64u + threadIdx.x;
    if (blk >= nblocks) return;
    uint8_t *ob = ows + (uint64_t)blk * PRZM_OWS_BYTES;
    przm_offs_t *H = (przm_offs_t *)(ob + PRZM_OWS_HDR);
    if (H->n > PRZM_OFFS_MAX) return; /* m12_len stays PRZM_ENC_OVERFLOW: the section stays flat */
    const uint32_t n = H->n;
    const uint8_t *codes = ob + PRZM_OWS_CODES;
    uint8_t *const end = ob + PRZM_OWS_M12 + PRZM_OWS_M12_CAP;
    const uint32_t cap = PRZM_OWS_M12_CAP - 16u;
    uint32_t x = 65536u, K = 0;
    bool over = false;
    uint32_t sym = n ? codes[n - 1] : 0u;
    for (uint32_t i = n; i-- > 0;) {
        const uint32_t prev = i ? codes[i - 1] : 0u;
        const uint32_t e = tab[dist_ctx(prev) * 64u + sym];
        const uint32_t f = e & 0xFFFFu;
        while (x > 4096u * f - 1u) {
            end[-1 - (int64_t)K] = (uint8_t)x;
            K++;
            x >>= 8;
        }
        uint32_t q, r;
        divmod(x, f, &q, &r);
        x = (q << 12) + (e >> 16) + r;
        sym = prev;
        if (K > cap) { over = true; break; }
    }
    H->m12_len = over ? PRZM_ENC_OVERFLOW : K;
    H->m12_state = x;
}

/* the adaptive V-ANS stream of the codes of block blockIdx.x * 16 + quad (64+ distances) */
extern "C" __global__ void __launch_bounds__(64)
przm_offs_vans_kernel(uint8_t *__restrict__ ows, uint32_t nblocks) {
    __shared__ uint32_t vt[16][256];
    __shared__ __align__(16) uint8_t stg[16][RANS_STAGE + 16];
    const uint32_t q = threadIdx.x >> 2, j = threadIdx.x & 3u;
    const uint32_t blk = blockIdx.x * 16u + q;
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
