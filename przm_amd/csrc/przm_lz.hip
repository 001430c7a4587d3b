/*
 * przm_lz.hip -- per-block match finder + tokenizer for gfx950 (CDNA4).
 *
 * One workgroup (8 wave64s) owns one 64 KiB block at a time and loops over
 * blocks (persistent grid).  The block's bytes live in LDS for the whole
 * kernel.  Result: the block's command / offset / literal streams and the
 * Section-2 substreams (op, meta, var), byte-identical to the oracle.
 *
 * Match finding (compress.go:681-707, 857-862, 374-387).  The chain of p is
 * "every earlier position with the same 16-bit hash4, newest first".  With
 * 160 KiB of LDS there is no room for a 65536-entry head table next to the
 * block, so chains are built by grouping instead of by sequential insertion:
 *   A. positions are bucketed by the hash's high byte with a stable LDS
 *      scatter (per-wave digit cursors + wave-ballot peer ranks), one half
 *      block at a time, and spilled to global scratch (coalesced);
 *   C. each bucket (position ordered) is walked by one wave: a 256-entry
 *      head table on the hash's low byte gives each element its in-bucket
 *      predecessor (== same full hash), and the search walks that chain
 *      right away against the LDS copy of the block.  Results (pos, dist)
 *      go to global scratch in bucket order;
 *   E. per half block, results are scattered back into an LDS dist array.
 *
 * Tokenizing (przm_tokenizer.cuh:115-466, one tile == the block).  The
 * reference walks tokens with one lane.  Here every 64-byte segment of a
 * half block is walked speculatively by its own thread from the segment
 * start; one wave then follows the true path across segments (a true entry
 * that lies on the speculative path inherits its exit; otherwise that
 * segment is re-walked), resolving the tokens that cross segment ends with
 * wave-wide compares.  Emission is two scans over segments (literal and
 * token counts, then byte counts) and a write pass; the ring-register index
 * of each match comes from a backward scan for the last three distinct
 * distances.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "przm_common.h"

#define LZ_NT 512
#define LZ_NW 8
#define SEG 64
#define NSEG (PRZM_HALF / SEG)
#define WS_CAP 2048
#define NONE16 0xFFFFu
#define NONE32 0xFFFFFFFFu
#define XCROSS 0x80000000u
#define LEN_CROSS 0xFFFFFFFFu

enum { T_LIT = 0, T_RLE = 1, T_MATCH = 2 };

/* uniform per-block state words */
enum {
    ST_ENTRY = 0, ST_LITSTART, ST_R0, ST_R1, ST_R2,
    ST_CMD, ST_OFF, ST_LIT, ST_OP, ST_META, ST_VAR, ST_TOK,
    ST_BUCKET, ST_DISTB, ST_CNT0, ST_CNT1, ST_NMATCH, ST_N
};

struct LzShared {
    uint8_t in[PRZM_BLOCK + 64];
    union {
        uint16_t stage[PRZM_HALF];
        uint16_t ws[LZ_NW][2][WS_CAP];
        uint16_t dist[PRZM_HALF + 4];
    } big;
    union {
        uint32_t tab[LZ_NW][256];
        struct {
            uint64_t bmp[NSEG];
            uint32_t xpos[NSEG];
            uint32_t tstart[NSEG];
            uint32_t xlen[NSEG];
            uint16_t tdist[PRZM_HALF / 4];
        } ps;
    } sm;
    uint32_t bstart[2][257];
    uint32_t wsum[LZ_NW];
    uint32_t st[ST_N + 3];
};

static_assert(sizeof(LzShared) <= 163840, "LDS budget");

/* ---------------- LDS byte helpers ---------------- */

__device__ __forceinline__ uint32_t ld32(const uint8_t *in, uint32_t off) {
    const uint32_t a = off & ~3u;
    const uint32_t lo = *(const uint32_t *)(in + a);
    const uint32_t hi = *(const uint32_t *)(in + a + 4);
    return __builtin_amdgcn_alignbyte(hi, lo, off & 3u);
}

__device__ __forceinline__ uint32_t hash4(const uint8_t *in, uint32_t p) { /* compress.go:204-206 */
    return (ld32(in, p) * 0x1E35A7BDu) >> 16;
}

/* length of the common prefix of in[p..] and in[q..], at most cap */
__device__ __forceinline__ uint32_t lcp(const uint8_t *in, uint32_t p, uint32_t q, uint32_t cap) {
    uint32_t l = 0;
    while (l < cap) {
        const uint32_t x = ld32(in, p + l) ^ ld32(in, q + l);
        if (x) {
            l += (uint32_t)__builtin_ctz(x) >> 3;
            return l < cap ? l : cap;
        }
        l += 4;
    }
    return cap;
}

/* bytes equal to b from p on, at most cap */
__device__ __forceinline__ uint32_t runlen(const uint8_t *in, uint32_t p, uint32_t b, uint32_t cap) {
    const uint32_t rep = b * 0x01010101u;
    uint32_t l = 0;
    while (l < cap) {
        const uint32_t x = ld32(in, p + l) ^ rep;
        if (x) {
            l += (uint32_t)__builtin_ctz(x) >> 3;
            return l < cap ? l : cap;
        }
        l += 4;
    }
    return cap;
}

/* wave-cooperative forms: every lane calls with the same arguments */
__device__ uint32_t lcp_wave(const uint8_t *in, uint32_t p, uint32_t q, uint32_t cap) {
    const uint32_t lane = __lane_id();
    for (uint32_t base = 0; base < cap; base += 256) {
        const uint32_t o = base + lane * 4;
        const uint32_t x = o < cap ? (ld32(in, p + o) ^ ld32(in, q + o)) : 0u;
        const uint64_t mm = __ballot(x != 0);
        if (mm) {
            const int f = __ffsll((unsigned long long)mm) - 1;
            const uint32_t xf = __shfl(x, f);
            const uint32_t l = base + (uint32_t)f * 4 + ((uint32_t)__builtin_ctz(xf) >> 3);
            return l < cap ? l : cap;
        }
    }
    return cap;
}

__device__ uint32_t runlen_wave(const uint8_t *in, uint32_t p, uint32_t b, uint32_t cap) {
    const uint32_t lane = __lane_id(), rep = b * 0x01010101u;
    for (uint32_t base = 0; base < cap; base += 256) {
        const uint32_t o = base + lane * 4;
        const uint32_t x = o < cap ? (ld32(in, p + o) ^ rep) : 0u;
        const uint64_t mm = __ballot(x != 0);
        if (mm) {
            const int f = __ffsll((unsigned long long)mm) - 1;
            const uint32_t xf = __shfl(x, f);
            const uint32_t l = base + (uint32_t)f * 4 + ((uint32_t)__builtin_ctz(xf) >> 3);
            return l < cap ? l : cap;
        }
    }
    return cap;
}

template <bool COOP>
__device__ __forceinline__ uint32_t LCP(const uint8_t *in, uint32_t p, uint32_t q, uint32_t cap) {
    return COOP ? lcp_wave(in, p, q, cap) : lcp(in, p, q, cap);
}
template <bool COOP>
__device__ __forceinline__ uint32_t RUN(const uint8_t *in, uint32_t p, uint32_t b, uint32_t cap) {
    return COOP ? runlen_wave(in, p, b, cap) : runlen(in, p, b, cap);
}

/* ---------------- wave / workgroup helpers ---------------- */

__device__ __forceinline__ uint64_t peers8(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

/* exclusive prefix sum over the workgroup; every thread calls it */
__device__ uint32_t wg_scan(LzShared &S, uint32_t v, uint32_t *total) {
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) S.wsum[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < LZ_NW; i++) {
        const uint32_t t = S.wsum[i];
        pre += (uint32_t)i < w ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

/* exclusive running maximum over the workgroup */
__device__ uint32_t wg_scan_max(LzShared &S, uint32_t v) {
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x = x > y ? x : y;
    }
    uint32_t ex = __shfl_up(x, 1);
    if (lane == 0) ex = 0;
    if (lane == 63) S.wsum[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t i = 0; i < w; i++) pre = pre > S.wsum[i] ? pre : S.wsum[i];
    __syncthreads();
    return pre > ex ? pre : ex;
}

__device__ __forceinline__ uint32_t leb_len(uint32_t v) {
    return v < 0x80u ? 1u : v < 0x4000u ? 2u : v < 0x200000u ? 3u : v < 0x10000000u ? 4u : 5u;
}
__device__ __forceinline__ uint32_t put_leb(uint8_t *b, uint32_t n, uint32_t v) {
    while (v >= 0x80u) { b[n++] = (uint8_t)(v | 0x80u); v >>= 7; }
    b[n++] = (uint8_t)v;
    return n;
}

/* ---------------- token evaluation ---------------- */

struct Tok { uint32_t kind, len; };

/*
 * The token the reference tokenizer emits at position p (p in [a, b) with
 * dist[] holding [a, b]), restated from przm_tokenizer.cuh:171-320 for a
 * walk that only needs exact token ends below `lim`.  A match / run whose
 * end is at or beyond lim comes back with len == LEN_CROSS.
 */
template <bool COOP>
__device__ Tok eval_tok(const LzShared &S, uint32_t p, uint32_t a, uint32_t n, uint32_t lim) {
    const uint8_t *in = S.in;
    const uint32_t d = S.big.dist[p - a];
    if (!d) return {T_LIT, 1};
    const uint32_t cap = n - p < PRZM_MAX_MATCH ? n - p : PRZM_MAX_MATCH;
    const uint32_t L = LCP<COOP>(in, p, p - d, cap); /* descriptor length (>= 4) */
    if (p + 1 < n) { /* lookahead, :196-216 */
        const uint32_t d1 = S.big.dist[p + 1 - a];
        if (d1) {
            uint32_t c1 = n - p - 1 < PRZM_MAX_MATCH ? n - p - 1 : PRZM_MAX_MATCH;
            if (c1 > L + 2) c1 = L + 2;
            if (LCP<COOP>(in, p + 1, p + 1 - d1, c1) > L + 1) return {T_LIT, 1};
        }
    }
    const uint32_t room = lim - p;
    if (p + 12 <= n) { /* run-length, :218-286 */
        const uint32_t w4 = ld32(in, p), b0 = w4 & 0xFFu;
        if (w4 == b0 * 0x01010101u) {
            const uint32_t need = L > 12 ? L : 12;
            const uint32_t maxr = n - p < PRZM_RLE_CAP ? n - p : PRZM_RLE_CAP;
            if (need <= maxr && RUN<COOP>(in, p, b0, need) >= need) {
                const uint32_t c = maxr < room ? maxr : room;
                const uint32_t r = RUN<COOP>(in, p, b0, c);
                if (r < c || maxr <= room) return {T_RLE, r};
                return {T_RLE, LEN_CROSS};
            }
        }
    }
    /* match: validated by construction (d <= p, L >= 4, 4 bytes equal); extend, :288-320 */
    const uint32_t capx = n - p < PRZM_EXT_CAP ? n - p : PRZM_EXT_CAP;
    const uint32_t c = capx < room ? capx : room;
    if (L >= c) {
        if (capx <= room) return {T_MATCH, capx < L ? capx : L};
        return {T_MATCH, LEN_CROSS};
    }
    const uint32_t e = L + LCP<COOP>(in, p + L, p - d + L, c - L);
    if (e < c || capx <= room) return {T_MATCH, e};
    return {T_MATCH, LEN_CROSS};
}

/* ring-register index of match number i (0..2) or -1 for an explicit offset */
__device__ int lru_index(const LzShared &S, uint32_t i, uint32_t d) {
    uint32_t seen0 = 0, seen1 = 0;
    int ns = 0;
    for (uint32_t j = i; j-- > 0;) {
        const uint32_t v = S.sm.ps.tdist[j];
        if (v == d) return ns;
        if (ns == 0) { seen0 = v; ns = 1; }
        else if (v != seen0 && (ns == 1 || v != seen1)) {
            if (ns == 1) { seen1 = v; ns = 2; }
            else return -1;
        }
    }
    const uint32_t r[3] = {S.st[ST_R0], S.st[ST_R1], S.st[ST_R2]};
    for (int j = 0; j < 3; j++) {
        const uint32_t v = r[j];
        if (v == d) return ns;
        if (ns == 0) { seen0 = v; ns = 1; }
        else if (v != seen0 && (ns == 1 || v != seen1)) {
            if (ns == 1) { seen1 = v; ns = 2; }
            else return -1;
        }
    }
    return -1;
}

/* ---------------- phase C: one bucket ---------------- */

template <bool LDSWS>
__device__ void bucket_search(LzShared &S, uint32_t w, uint32_t dgt, uint32_t n, uint32_t depth,
                              const uint16_t *__restrict__ T0, const uint16_t *__restrict__ T1,
                              uint32_t *__restrict__ R0, uint32_t *__restrict__ R1, uint16_t *gwpos,
                              uint16_t *gwprev) {
    const uint32_t lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t s0 = S.bstart[0][dgt], m0 = S.bstart[0][dgt + 1] - s0;
    const uint32_t s1 = S.bstart[1][dgt], m1 = S.bstart[1][dgt + 1] - s1;
    const uint32_t m = m0 + m1, tag = dgt + 1;
    uint16_t *wpos = LDSWS ? S.big.ws[w][0] : gwpos;
    uint16_t *wprev = LDSWS ? S.big.ws[w][1] : gwprev;
    const uint8_t *in = S.in;
    for (uint32_t c0 = 0; c0 < m; c0 += 64) {
        const uint32_t j = c0 + lane;
        const bool valid = j < m;
        uint32_t pos = 0;
        if (valid) pos = j < m0 ? T0[s0 + j] : T1[s1 + j - m0];
        const uint32_t low = valid ? (hash4(in, pos) & 0xFFu) : 0u;
        const uint64_t pm = peers8(low, valid);
        uint32_t pred = NONE16;
        if (valid) {
            const uint64_t below = pm & lt;
            if (below) pred = c0 + 63u - (uint32_t)__clzll((unsigned long long)below);
            else {
                const uint32_t hv = S.sm.tab[w][low];
                if ((hv >> 16) == tag) pred = hv & 0xFFFFu;
            }
            wpos[j] = (uint16_t)pos;
            wprev[j] = (uint16_t)pred;
            if ((pm >> lane) == 1ull) S.sm.tab[w][low] = (tag << 16) | j;
        }
        if (!LDSWS) __threadfence_block();
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            const uint32_t cap = n - pos < PRZM_MAX_MATCH ? n - pos : PRZM_MAX_MATCH;
            uint32_t best = 0, bd = 0, cand = pred;
            for (uint32_t k = 0; cand != NONE16 && k < depth; k++) {
                const uint32_t q = wpos[cand];
                if (in[pos + best] == in[q + best]) {
                    const uint32_t l = lcp(in, pos, q, cap);
                    if (l > best) { best = l; bd = pos - q; }
                    if (best >= cap) break;
                }
                cand = wprev[cand];
            }
            const uint32_t dist = best >= 4 ? bd : 0u;
            const uint32_t rv = pos | (dist << 16);
            if (j < m0) R0[s0 + j] = rv;
            else R1[s1 + j - m0] = rv;
            if (pos == PRZM_HALF) S.st[ST_DISTB] = dist;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

/* ---------------- parse of one half block ---------------- */

__device__ void parse_half(LzShared &S, uint32_t a, uint32_t b, uint32_t n, uint8_t *__restrict__ wsb) {
    const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    uint8_t *cmd = wsb + PRZM_WS_CMD, *off = wsb + PRZM_WS_OFF, *lit = wsb + PRZM_WS_LIT;
    uint8_t *op = wsb + PRZM_WS_OP, *meta = wsb + PRZM_WS_META, *var = wsb + PRZM_WS_VAR;
    const uint8_t *in = S.in;
    const uint32_t s = tid, ss = a + s * SEG;
    const bool segv = ss < b;
    const uint32_t se = segv ? (ss + SEG < b ? ss + SEG : b) : ss;

    /* 1. speculative walk of every segment from its start */
    if (segv) {
        uint64_t bm = 0;
        uint32_t p = ss, xp = 0;
        bool cross = false;
        while (p < se) {
            bm |= 1ull << (p - ss);
            const Tok t = eval_tok<false>(S, p, a, n, se);
            if (t.kind == T_LIT) { p++; continue; }
            if (t.len == LEN_CROSS) { cross = true; xp = p; break; }
            p += t.len;
        }
        S.sm.ps.bmp[s] = bm;
        S.sm.ps.xpos[s] = cross ? (xp | XCROSS) : p;
    }
    S.sm.ps.tstart[s] = NONE32;
    __syncthreads();

    /* 2. follow the true path (wave 0) */
    if (w == 0) {
        uint32_t cur = S.st[ST_ENTRY];
        while (cur < b) {
            const uint32_t sg = (cur - a) / SEG, sgs = a + sg * SEG, sge = sgs + SEG < b ? sgs + SEG : b;
            uint32_t xp = S.sm.ps.xpos[sg];
            if (!((S.sm.ps.bmp[sg] >> (cur - sgs)) & 1ull)) {
                /* the true entry is off the speculative path: walk the segment again */
                uint32_t p = cur;
                bool cross = false;
                while (p < sge) {
                    const Tok t = eval_tok<false>(S, p, a, n, sge);
                    if (t.kind == T_LIT) { p++; continue; }
                    if (t.len == LEN_CROSS) { cross = true; xp = p | XCROSS; break; }
                    p += t.len;
                }
                if (!cross) xp = p;
                if (lane == 0) S.sm.ps.xpos[sg] = xp;
            }
            if (lane == 0) S.sm.ps.tstart[sg] = cur;
            if (xp & XCROSS) {
                const uint32_t tp = xp & ~XCROSS;
                const Tok t = eval_tok<true>(S, tp, a, n, NONE32);
                if (lane == 0) S.sm.ps.xlen[sg] = t.len | (t.kind << 30);
                cur = tp + t.len;
            } else {
                cur = xp;
            }
        }
        if (lane == 0) S.st[ST_ENTRY] = cur;
    }
    __syncthreads();

    /* 3. count literals / runs / matches per segment on the true path */
    const uint32_t t0 = S.sm.ps.tstart[s];
    const bool live = segv && t0 != NONE32;
    uint32_t nl = 0, nr = 0, nm = 0, last_end = 0;
    if (live) {
        uint32_t p = t0;
        while (p < se) {
            Tok t = eval_tok<false>(S, p, a, n, se);
            if (t.kind == T_LIT) { nl++; p++; continue; }
            if (t.len == LEN_CROSS) t.len = S.sm.ps.xlen[s] & 0x3FFFFFFFu;
            if (t.kind == T_RLE) nr++;
            else nm++;
            p += t.len;
            last_end = p;
        }
    }
    uint32_t tot_l, tot_m;
    const uint32_t lit_base = wg_scan(S, nl + nr, &tot_l) + S.st[ST_LIT];
    const uint32_t m_base = wg_scan(S, nm, &tot_m);
    uint32_t ls_in = wg_scan_max(S, last_end);
    const uint32_t carried_ls = S.st[ST_LITSTART];
    if (ls_in < carried_ls) ls_in = carried_ls;

    /* 4. match distances in order */
    if (live) {
        uint32_t p = t0, mi = m_base;
        while (p < se) {
            Tok t = eval_tok<false>(S, p, a, n, se);
            if (t.kind == T_LIT) { p++; continue; }
            if (t.len == LEN_CROSS) t.len = S.sm.ps.xlen[s] & 0x3FFFFFFFu;
            if (t.kind == T_MATCH) S.sm.ps.tdist[mi++] = S.big.dist[p - a];
            p += t.len;
        }
    }
    __syncthreads();

    /* 5. byte counts per segment */
    uint32_t ccmd = 0, coff = 0, cop = 0, cmeta = 0, cvar = 0;
    if (live) {
        uint32_t p = t0, mi = m_base, ls = ls_in;
        while (p < se) {
            Tok t = eval_tok<false>(S, p, a, n, se);
            if (t.kind == T_LIT) { p++; continue; }
            if (t.len == LEN_CROSS) t.len = S.sm.ps.xlen[s] & 0x3FFFFFFFu;
            const uint32_t ll = p - ls;
            if (t.kind == T_RLE) {
                if (ll) { ccmd += 1; cop += 1; if (ll > 8) { const uint32_t q = leb_len(ll); ccmd += q; cvar += q; } }
                const uint32_t q = leb_len(t.len);
                ccmd += 1 + q; cop += 1; cvar += q;
            } else {
                const uint32_t d = S.big.dist[p - a];
                const int k = lru_index(S, mi, d);
                mi++;
                const uint32_t keep = k >= 0 ? (t.len <= 7 ? 7u : 15u) : 3u;
                if (ll > keep) {
                    const uint32_t f = ll - keep;
                    ccmd += 1; cop += 1;
                    if (f > 8) { const uint32_t q = leb_len(f); ccmd += q; cvar += q; }
                }
                const uint32_t xq = t.len - 4 > 14 ? leb_len(t.len - 19) : 0u;
                if (k >= 0 && t.len <= 7) { ccmd += 1; cop += 1; }
                else if (k >= 0) { ccmd += 2 + xq; cop += 1; cmeta += 1; cvar += xq; }
                else { ccmd += 1 + xq; cop += 1; cvar += xq; coff += leb_len(d); }
            }
            p += t.len;
            ls = p;
        }
    }
    uint32_t t_cmd, t_off, t_op, t_meta, t_var;
    uint32_t b_cmd = wg_scan(S, ccmd, &t_cmd) + S.st[ST_CMD];
    uint32_t b_off = wg_scan(S, coff, &t_off) + S.st[ST_OFF];
    uint32_t b_op = wg_scan(S, cop, &t_op) + S.st[ST_OP];
    uint32_t b_meta = wg_scan(S, cmeta, &t_meta) + S.st[ST_META];
    uint32_t b_var = wg_scan(S, cvar, &t_var) + S.st[ST_VAR];

    /* 6. write bytes */
    if (live) {
        uint32_t p = t0, mi = m_base, ls = ls_in, lb = lit_base;
        while (p < se) {
            Tok t = eval_tok<false>(S, p, a, n, se);
            if (t.kind == T_LIT) { lit[lb++] = in[p]; p++; continue; }
            if (t.len == LEN_CROSS) t.len = S.sm.ps.xlen[s] & 0x3FFFFFFFu;
            const uint32_t ll = p - ls;
            if (t.kind == T_RLE) {
                if (ll) {
                    if (ll <= 8) { cmd[b_cmd++] = (uint8_t)(0xF2 + ll); op[b_op++] = (uint8_t)(0xF2 + ll); }
                    else {
                        cmd[b_cmd++] = 0xF0; op[b_op++] = 0xF0;
                        b_cmd = put_leb(cmd, b_cmd, ll); b_var = put_leb(var, b_var, ll);
                    }
                }
                cmd[b_cmd++] = 0xF1; op[b_op++] = 0xF1;
                b_cmd = put_leb(cmd, b_cmd, t.len); b_var = put_leb(var, b_var, t.len);
                lit[lb++] = in[p];
            } else {
                const uint32_t d = S.big.dist[p - a];
                const int k = lru_index(S, mi, d);
                mi++;
                const uint32_t keep = k >= 0 ? (t.len <= 7 ? 7u : 15u) : 3u;
                uint32_t l2 = ll;
                if (ll > keep) {
                    const uint32_t f = ll - keep;
                    if (f <= 8) { cmd[b_cmd++] = (uint8_t)(0xF2 + f); op[b_op++] = (uint8_t)(0xF2 + f); }
                    else {
                        cmd[b_cmd++] = 0xF0; op[b_op++] = 0xF0;
                        b_cmd = put_leb(cmd, b_cmd, f); b_var = put_leb(var, b_var, f);
                    }
                    l2 = keep;
                }
                const uint32_t base = t.len - 4 > 14 ? 15u : t.len - 4;
                if (k >= 0 && t.len <= 7) {
                    const uint8_t o = (uint8_t)((l2 << 4) | ((t.len - 4) << 2) | (uint32_t)k);
                    cmd[b_cmd++] = o; op[b_op++] = o;
                } else if (k >= 0) {
                    const uint8_t tok = (uint8_t)((l2 << 4) | base);
                    cmd[b_cmd++] = (uint8_t)(0xC0 | k); cmd[b_cmd++] = tok;
                    op[b_op++] = (uint8_t)(0xC0 | k); meta[b_meta++] = tok;
                    if (base == 15) { b_cmd = put_leb(cmd, b_cmd, t.len - 19); b_var = put_leb(var, b_var, t.len - 19); }
                } else {
                    const uint8_t o = (uint8_t)(0x80 | (l2 << 4) | base);
                    cmd[b_cmd++] = o; op[b_op++] = o;
                    if (base == 15) { b_cmd = put_leb(cmd, b_cmd, t.len - 19); b_var = put_leb(var, b_var, t.len - 19); }
                    b_off = put_leb(off, b_off, d);
                }
            }
            p += t.len;
            ls = p;
        }
    }
    __syncthreads();

    /* 7. carry the uniform state into the next half */
    if (tid == 0) {
        uint32_t r0 = S.st[ST_R0], r1 = S.st[ST_R1], r2 = S.st[ST_R2];
        uint32_t out[3];
        int no = 0;
        for (uint32_t j = tot_m; j-- > 0 && no < 3;) {
            const uint32_t v = S.sm.ps.tdist[j];
            bool dup = false;
            for (int q = 0; q < no; q++) dup |= out[q] == v;
            if (!dup) out[no++] = v;
        }
        const uint32_t rr[3] = {r0, r1, r2};
        for (int j = 0; j < 3 && no < 3; j++) {
            bool dup = false;
            for (int q = 0; q < no; q++) dup |= out[q] == rr[j];
            if (!dup) out[no++] = rr[j];
        }
        S.st[ST_R0] = out[0]; S.st[ST_R1] = out[1]; S.st[ST_R2] = out[2];
        S.st[ST_CMD] += t_cmd; S.st[ST_OFF] += t_off; S.st[ST_OP] += t_op;
        S.st[ST_META] += t_meta; S.st[ST_VAR] += t_var; S.st[ST_LIT] += tot_l;
        /* last non-literal end of this half (or the carried one) */
        S.st[ST_NMATCH] = tot_m;
    }
    /* litStart for the next half: the exclusive max over all segments plus this one */
    {
        uint32_t v = live ? last_end : 0u;
        const uint32_t ex = wg_scan_max(S, v);
        const uint32_t incl = ex > v ? ex : v;
        /* thread NT-1 holds the maximum over all segments */
        if (tid == LZ_NT - 1) {
            const uint32_t c = S.st[ST_LITSTART];
            S.st[ST_LITSTART] = incl > c ? incl : c;
        }
    }
    uint32_t ntok_half;
    (void)wg_scan(S, nr + nm, &ntok_half);
    if (tid == 0) S.st[ST_TOK] += ntok_half;
    __syncthreads();
    (void)lane;
}

/* ---------------- kernel ---------------- */

extern "C" __global__ void __launch_bounds__(LZ_NT)
przm_lz_kernel(const uint8_t *__restrict__ src, const przm_block_t *__restrict__ blocks, uint32_t nblocks,
               uint32_t depth, uint8_t *__restrict__ ws, przm_meta_t *__restrict__ meta,
               uint8_t *__restrict__ scratch, uint64_t scratch_stride, uint16_t *__restrict__ dist_dbg) {
    __shared__ LzShared S;
    const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    uint8_t *my = scratch + (uint64_t)blockIdx.x * scratch_stride;
    uint16_t *T0 = (uint16_t *)my;
    uint16_t *T1 = T0 + PRZM_HALF;
    uint32_t *R0 = (uint32_t *)(T1 + PRZM_HALF);
    uint32_t *R1 = R0 + PRZM_HALF;
    uint16_t *gws = (uint16_t *)(R1 + PRZM_HALF) + (uint64_t)w * 2u * PRZM_BLOCK;

    for (uint32_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
        const przm_block_t bd = blocks[blk];
        const uint32_t n = bd.len;
        przm_meta_t *mt = meta + blk;
        uint8_t *wsb = ws + (uint64_t)blk * PRZM_WS_BYTES;
        if ((bd.flags & PRZM_BF_RAW) || n == 0) {
            if (tid < PRZM_S_N) mt->len[tid] = 0;
            continue;
        }
        /* load the block into LDS (16-byte loads when the block is aligned) */
        const uint8_t *g = src + bd.src_off;
        const bool al16 = (bd.src_off & 15u) == 0;
        for (uint32_t i = tid * 16; i < PRZM_BLOCK + 64; i += LZ_NT * 16) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (al16 && i + 16 <= n) v = *(const uint4 *)(g + i);
            else if (i < n) {
                uint8_t tmp[16];
                for (int k = 0; k < 16; k++) tmp[k] = i + k < n ? g[i + k] : 0;
                v = *(uint4 *)tmp;
            }
            if (i + 16 <= PRZM_BLOCK + 64) *(uint4 *)(S.in + i) = v;
        }
        if (tid < ST_N) S.st[tid] = 0;
        __syncthreads();
        if (tid == 0) { S.st[ST_R0] = 1; S.st[ST_R1] = 4; S.st[ST_R2] = 8; }

        const uint32_t nh = n >= 4 ? n - 3 : 0; /* positions p with p + 4 <= n */

        /* A. bucket positions by the hash's high byte, per half */
        for (uint32_t H = 0; H < 2; H++) {
            const uint32_t base = H * PRZM_HALF;
            const uint32_t end = nh < base + PRZM_HALF ? (nh > base ? nh : base) : base + PRZM_HALF;
            for (uint32_t i = tid; i < LZ_NW * 256; i += LZ_NT) (&S.sm.tab[0][0])[i] = 0;
            __syncthreads();
            const uint32_t wb = base + w * (PRZM_HALF / LZ_NW);
            for (uint32_t c = 0; c < PRZM_HALF / LZ_NW; c += 64) {
                const uint32_t p = wb + c + lane;
                if (p < end) atomicAdd(&S.sm.tab[w][hash4(S.in, p) >> 8], 1u);
            }
            __syncthreads();
            {
                const uint32_t dg = tid < 256 ? tid : 0;
                uint32_t tot = 0;
                if (tid < 256)
                    for (int k = 0; k < LZ_NW; k++) tot += S.sm.tab[k][dg];
                uint32_t all;
                const uint32_t st0 = wg_scan(S, tid < 256 ? tot : 0u, &all);
                if (tid < 256) {
                    uint32_t run = st0;
                    for (int k = 0; k < LZ_NW; k++) {
                        const uint32_t c = S.sm.tab[k][dg];
                        S.sm.tab[k][dg] = run;
                        run += c;
                    }
                    S.bstart[H][dg] = st0;
                }
                if (tid == 0) S.bstart[H][256] = all;
            }
            __syncthreads();
            const uint64_t lt = (1ull << lane) - 1ull;
            for (uint32_t c = 0; c < PRZM_HALF / LZ_NW; c += 64) {
                const uint32_t p = wb + c + lane;
                const bool valid = p < end;
                const uint32_t dg = valid ? hash4(S.in, p) >> 8 : 0u;
                const uint64_t pm = peers8(dg, valid);
                if (valid) {
                    const uint32_t cur = S.sm.tab[w][dg];
                    S.big.stage[cur + (uint32_t)__popcll(pm & lt)] = (uint16_t)p;
                    if ((pm & lt) == 0) S.sm.tab[w][dg] = cur + (uint32_t)__popcll(pm);
                }
                __builtin_amdgcn_wave_barrier();
            }
            __syncthreads();
            const uint32_t cnt = S.bstart[H][256];
            uint16_t *T = H ? T1 : T0;
            for (uint32_t i = tid; i < cnt; i += LZ_NT) T[i] = S.big.stage[i];
            __syncthreads();
        }

        /* C. per bucket: chains by the low byte, then the chain search */
        for (uint32_t i = tid; i < LZ_NW * 256; i += LZ_NT) (&S.sm.tab[0][0])[i] = 0;
        if (tid == 0) S.st[ST_BUCKET] = 0;
        __threadfence_block();
        __syncthreads();
        for (;;) {
            uint32_t dg = 0;
            if (lane == 0) dg = atomicAdd(&S.st[ST_BUCKET], 1u);
            dg = __shfl(dg, 0);
            if (dg >= 256) break;
            const uint32_t m = (S.bstart[0][dg + 1] - S.bstart[0][dg]) + (S.bstart[1][dg + 1] - S.bstart[1][dg]);
            if (m == 0) continue;
            if (m <= WS_CAP) bucket_search<true>(S, w, dg, n, depth, T0, T1, R0, R1, nullptr, nullptr);
            else bucket_search<false>(S, w, dg, n, depth, T0, T1, R0, R1, gws, gws + PRZM_BLOCK);
        }
        __threadfence_block();
        __syncthreads();

        /* E + parse, per half */
        const uint32_t nhalves = n > PRZM_HALF ? 2u : 1u;
        for (uint32_t H = 0; H < nhalves; H++) {
            const uint32_t a = H * PRZM_HALF, b = n < a + PRZM_HALF ? n : a + PRZM_HALF;
            for (uint32_t i = tid; i <= b - a; i += LZ_NT) S.big.dist[i] = 0;
            __syncthreads();
            const uint32_t cnt = S.bstart[H][256];
            const uint32_t *R = H ? R1 : R0;
            for (uint32_t i = tid; i < cnt; i += LZ_NT) {
                const uint32_t rv = R[i];
                S.big.dist[(rv & 0xFFFFu) - a] = (uint16_t)(rv >> 16);
            }
            if (H == 0 && b < n && tid == 0) S.big.dist[b - a] = (uint16_t)S.st[ST_DISTB];
            __syncthreads();
            if (dist_dbg)
                for (uint32_t i = tid; i < b - a; i += LZ_NT) dist_dbg[(uint64_t)blk * PRZM_BLOCK + a + i] = S.big.dist[i];
            if (S.st[ST_ENTRY] < b) parse_half(S, a, b, n, wsb);
            __syncthreads();
        }

        /* trailing literal run (przm_tokenizer.cuh:440-457) */
        if (tid == 0) {
            const uint32_t ls = S.st[ST_LITSTART];
            uint32_t ncmd = S.st[ST_CMD], nop = S.st[ST_OP], nvar = S.st[ST_VAR];
            if (ls < n) {
                const uint32_t f = n - ls;
                uint8_t *cmd = wsb + PRZM_WS_CMD, *op = wsb + PRZM_WS_OP, *var = wsb + PRZM_WS_VAR;
                if (f <= 8) { cmd[ncmd++] = (uint8_t)(0xF2 + f); op[nop++] = (uint8_t)(0xF2 + f); }
                else { cmd[ncmd++] = 0xF0; op[nop++] = 0xF0; ncmd = put_leb(cmd, ncmd, f); nvar = put_leb(var, nvar, f); }
            }
            mt->len[PRZM_S_CMD] = ncmd;
            mt->len[PRZM_S_OFF] = S.st[ST_OFF];
            mt->len[PRZM_S_LIT] = S.st[ST_LIT];
            mt->len[PRZM_S_OP] = nop;
            mt->len[PRZM_S_META] = S.st[ST_META];
            mt->len[PRZM_S_VAR] = nvar;
            mt->len[PRZM_S_TOK] = S.st[ST_TOK];
        }
        __syncthreads();
    }
}
