/*
 * przm_lz.hip -- per-block match finder + tokenizer for gfx950 (CDNA4).
 *
 * One workgroup (8 wave64s) owns one 64 KiB block at a time and loops over
 * blocks (persistent grid, blocks handed out by an atomic counter).  The
 * block's bytes live in LDS for the whole kernel.  Result: the block's
 * command / offset / literal streams and the Section-2 substreams (op, meta,
 * var), byte-identical to the oracle.
 *
 * Match finding (compress.go:681-707, 857-862, 374-387).  The chain of p is
 * "every earlier position with the same 16-bit hash4, newest first", at most
 * `depth` of them.  With 160 KiB of LDS there is no room for a 65536-entry
 * head table next to the block, so chains are built by grouping instead of
 * by sequential insertion:
 *   A. positions are bucketed by the hash's high byte with a stable LDS
 *      scatter (per-wave digit cursors; the lanes sharing a digit found
 *      through a per-wave bitmap word per digit), one half block at a time,
 *      and spilled to global scratch as 16-byte stores;
 *   C. each bucket is taken by one wave (large ones are spilled and searched
 *      by all eight): its positions are grouped stably by the hash's low
 *      byte, so each group is one full hash in position order (a chain read
 *      backward), and searched 64 elements at a time from a per-wave LDS ring
 *      of 12-byte keys; a candidate's score is length << 16 | position, so
 *      the best of a chain is a plain maximum.  Results go to a per-position
 *      array in global scratch;
 *   E. per quarter block, the results are read back into LDS (dist, length,
 *      and a "has a match" mask per 32 positions).
 *
 * Tokenizing (przm_tokenizer.cuh:115-466, one tile == a quarter block).  The
 * reference walks tokens with one lane.  Here every 32-byte segment of the
 * quarter is classified and walked speculatively by its own thread from the
 * segment start; the workgroup then follows the true path across segments by
 * pointer doubling over segment -> segment-of-exit, re-walking segments whose
 * entry is off their speculative path and resolving tokens that cross
 * segment ends with wave-wide compares.  Emission is two scans over segments
 * (literal and token counts, then the five per-stream byte counts) and a
 * write pass that keeps each segment's tokens in registers; the literal, cmd
 * and op bytes are staged in LDS and stored as aligned 16-byte words.  The
 * ring-register index of each match comes from a backward scan for the last
 * three distinct distances.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "przm_common.h"

/* a byte in global memory (address space 1), for pointers the compiler cannot trace */
typedef __attribute__((address_space(1))) uint8_t gu8;

#define LZ_NT 512
#define LZ_NW 8
#define QP 16384                  /* parse tile: a quarter block */
#define SEG 32                    /* speculative-walk segment */
#define NSEG (QP / SEG)           /* 512: one per thread */
#define WS_CAP 1536               /* bucket elements kept in LDS per wave */
#define RING 128                  /* staged search keys per wave: previous + current chunk */
/* chain candidates read per step of the search: a template parameter B of the kernels (2, 4 or
 * 8, picked per workload by the autotuner; 4 by default) */
#define WS_MAXV 5                 /* values per workgroup scan */
#define NONE16 0xFFFFu
#define NONE32 0xFFFFFFFFu
#define XCROSS 0x80000000u
#define LEN_CROSS 0xFFFFFFFFu
#define XROOM 128u                /* bytes past a segment end a thread may look to settle a token */
#define NX_STOP 0xFFFEu
#define NX_END 0xFFFFu
/* padded LDS indices: one spare u16 / four spare bytes per 32 entries, so the
 * lanes of a wave walking consecutive segments hit different banks */
#define DIX(i) ((i) + ((i) >> 5))
#define LIX(i) ((i) + (((i) >> 5) << 2))
#define DSZ ((DIX(QP + 1) + 3) & ~1) /* even, so that lc below starts 4-byte aligned */
#define LSZ (LIX(QP + 1) + 8)
#define GWS_SLOT (2u * PRZM_BLOCK * 2u + 256u * 4u) /* one large bucket: positions, order, group starts */
#define GWS_STRIDE (2u * GWS_SLOT)                   /* per wave: the queued bucket + one worked in place */
/* global scratch of one LZ workgroup: the half blocks' bucketed positions (T0, T1: u16), the
 * per-position results (D: u32), 2 * PRZM_HALF * 2 bytes unused, the waves' workspaces */
constexpr uint64_t LZ_SCRATCH_STRIDE = 2ull * PRZM_HALF * 2 + 2ull * PRZM_HALF * 4 + 2ull * PRZM_HALF * 2 +
                                       (uint64_t)LZ_NW * GWS_STRIDE;

enum { T_LIT = 0, T_RLE = 1, T_MATCH = 2 };

/* uniform per-block state words */
enum {
    ST_ENTRY = 0, ST_LITSTART, ST_R0, ST_R1, ST_R2,
    ST_CMD, ST_OFF, ST_LIT, ST_OP, ST_META, ST_VAR, ST_TOK,
    ST_BUCKET, ST_DISTB, ST_LB, ST_NMATCH, ST_E0, ST_BAD, ST_N
};

struct LzShared {
    uint8_t in[PRZM_BLOCK + 64];
    union {
        uint16_t stage[PRZM_HALF];
        struct {
            uint16_t pos[LZ_NW][WS_CAP];
            uint16_t srt[LZ_NW][WS_CAP];
            uint4 ring[LZ_NW][RING]; /* bytes p..p+11 of a staged bucket element, and p */
        } ws;
        struct {
            uint16_t dist[DSZ];
            uint8_t lc[LSZ];
            __align__(16) uint16_t tlen[NSEG][8]; /* cached non-literal tokens of each segment (swizzled) */
            uint8_t toff[NSEG][8];  /* start offset in segment | kind << 5 */
        } q;
    } big;
    union {
        struct {
            uint32_t tab[LZ_NW][256];
            uint64_t pbits[LZ_NW][256]; /* per wave and digit: the lanes holding it (zero between uses) */
            uint16_t gstart[LZ_NW][256]; /* per wave and low digit: where its group starts in the bucket */
        };
        struct {
            uint32_t bmp[NSEG];
            uint32_t xpos[NSEG];
            uint32_t xlen[NSEG];
            uint32_t ent[NSEG];
            uint16_t nxt[NSEG];
            uint16_t jmp[NSEG];
            uint16_t prv[NSEG];
            uint8_t vis[NSEG];
            uint8_t ntk[NSEG];
            uint16_t tdist[QP / 4];
            uint32_t hm[NSEG + 1]; /* per segment: positions with a match (bit k: position 32 s + k) */
        } ps;
    } sm;
    uint32_t bstart[2][257];
    uint32_t wsum[LZ_NW][WS_MAXV];
    uint32_t st[ST_N + 3];
    uint32_t bigq[LZ_NW];
    uint32_t nextb;
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

/* DPP move of x with control CTRL on rows ROWMASK; lanes whose source is out of range, and
 * rows outside the mask, read 0 */
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dppz(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWMASK, 0xF, false);
}

/* inclusive wave scans: row_shr 1, 2, 4, 8 inside each row of 16 lanes, then row 0's last lane
 * into row 1 and row 2's into row 3 (row_bcast:15), then lane 31 into rows 2 and 3
 * (row_bcast:31) */
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x) {
    x += dppz<0x111, 0xF>(x);
    x += dppz<0x112, 0xF>(x);
    x += dppz<0x114, 0xF>(x);
    x += dppz<0x118, 0xF>(x);
    x += dppz<0x142, 0xA>(x);
    x += dppz<0x143, 0xC>(x);
    return x;
}
__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t x) {
    x = umax(x, dppz<0x111, 0xF>(x));
    x = umax(x, dppz<0x112, 0xF>(x));
    x = umax(x, dppz<0x114, 0xF>(x));
    x = umax(x, dppz<0x118, 0xF>(x));
    x = umax(x, dppz<0x142, 0xA>(x));
    x = umax(x, dppz<0x143, 0xC>(x));
    return x;
}
/* the value of the lane below (0 in lane 0): wave_shr:1 */
__device__ __forceinline__ uint32_t wave_prev(uint32_t x) { return dppz<0x138, 0xF>(x); }

/*
 * Exclusive workgroup scans of NS sums (sv: value in, exclusive prefix out; stot: totals)
 * and NM running maxima (mv: value in, exclusive running maximum out) in one pass: one
 * barrier to publish the per-wave totals, one to release them.  Every thread calls it.
 */
template <int NS, int NM>
__device__ __forceinline__ void wg_scan_multi(LzShared &S, uint32_t *sv, uint32_t *stot, uint32_t *mv) {
    static_assert(NS + NM <= WS_MAXV, "wsum size");
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t xs[NS > 0 ? NS : 1], xm[NM > 0 ? NM : 1];
#pragma unroll
    for (int k = 0; k < NS; k++) xs[k] = wave_incl_sum(sv[k]);
#pragma unroll
    for (int k = 0; k < NM; k++) xm[k] = wave_incl_max(mv[k]);
    if (lane == 63) {
#pragma unroll
        for (int k = 0; k < NS; k++) S.wsum[w][k] = xs[k];
#pragma unroll
        for (int k = 0; k < NM; k++) S.wsum[w][NS + k] = xm[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NS; k++) {
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < LZ_NW; i++) {
            const uint32_t t = S.wsum[i][k];
            pre += (uint32_t)i < w ? t : 0u;
            tot += t;
        }
        sv[k] = pre + xs[k] - sv[k];
        stot[k] = tot;
    }
#pragma unroll
    for (int k = 0; k < NM; k++) {
        uint32_t pre = 0;
#pragma unroll
        for (int i = 0; i < LZ_NW; i++) pre = (uint32_t)i < w ? umax(pre, S.wsum[i][NS + k]) : pre;
        mv[k] = umax(pre, wave_prev(xm[k]));
    }
    __syncthreads();
}

/* exclusive prefix sum over the workgroup; every thread calls it */
__device__ __forceinline__ uint32_t wg_scan(LzShared &S, uint32_t v, uint32_t *total) {
    uint32_t sv[1] = {v}, st[1];
    wg_scan_multi<1, 0>(S, sv, st, nullptr);
    *total = st[0];
    return sv[0];
}

__device__ __forceinline__ uint32_t leb_len(uint32_t v) {
    return v < 0x80u ? 1u : v < 0x4000u ? 2u : v < 0x200000u ? 3u : v < 0x10000000u ? 4u : 5u;
}
template <typename P>
__device__ __forceinline__ uint32_t put_leb(P b, uint32_t n, uint32_t v) {
    while (v >= 0x80u) { b[n++] = (uint8_t)(v | 0x80u); v >>= 7; }
    b[n++] = (uint8_t)v;
    return n;
}

/* ---------------- token evaluation ---------------- */

struct Tok { uint32_t kind, len; };

__device__ __forceinline__ uint32_t dget(const LzShared &S, uint32_t i) { return S.big.q.dist[DIX(i)]; }
__device__ __forceinline__ uint32_t lget(const LzShared &S, uint32_t i) { return S.big.q.lc[LIX(i)]; }

/* exact descriptor length at p (tile index i) with distance d; lc keeps min(L, 255) */
template <bool COOP>
__device__ __forceinline__ uint32_t desc_len(const LzShared &S, uint32_t p, uint32_t i, uint32_t d, uint32_t n) {
    const uint32_t l = lget(S, i);
    if (l < 255) return l;
    const uint32_t cap = n - p < PRZM_MAX_MATCH ? n - p : PRZM_MAX_MATCH;
    return 255 + LCP<COOP>(S.in, p + 255, p - d + 255, cap - 255);
}

/*
 * The token the reference tokenizer emits at position p of the tile [a, b)
 * (dist / lc hold [a, b]), restated from przm_tokenizer.cuh:171-320.  The
 * descriptor length L of every position comes from the search, so a match
 * shorter than the 1024 cap has its exact length without any compare; only a
 * capped match (extension) or a byte run may need compares.  Those whose end
 * is at or beyond lim come back with len == LEN_CROSS.
 */
template <bool COOP>
__device__ Tok eval_tok(const LzShared &S, uint32_t p, uint32_t a, uint32_t n, uint32_t lim) {
    const uint8_t *in = S.in;
    const uint32_t i = p - a;
    /* everything a position's decision reads, loaded together (one LDS round trip
     * instead of a chain); entries i + 1 <= the tile end and bytes p..p+3 are in range */
    const uint32_t d = dget(S, i), lc = lget(S, i), d1r = dget(S, i + 1), l1c = lget(S, i + 1);
    const uint32_t w4 = ld32(in, p);
    if (!d) return {T_LIT, 1};
    const uint32_t L = lc < 255 ? lc : desc_len<COOP>(S, p, i, d, n);
    if (p + 1 < n) { /* lookahead, :196-216 */
        const uint32_t d1 = d1r;
        if (d1) {
            uint32_t l1 = l1c;
            if (l1c == 255 && L + 1 >= 255) l1 = desc_len<COOP>(S, p + 1, i + 1, d1, n);
            if (l1 > L + 1) return {T_LIT, 1};
        }
    }
    const uint32_t room = lim - p;
    if (p + 12 <= n) { /* run-length, :218-286 */
        const uint32_t b0 = w4 & 0xFFu;
        if (w4 == b0 * 0x01010101u) {
            const uint32_t need = L > 12 ? L : 12;
            const uint32_t maxr = n - p < PRZM_RLE_CAP ? n - p : PRZM_RLE_CAP;
            if (need <= maxr && RUN<COOP>(in, p, b0, need) >= need) {
                const uint32_t c = maxr < room ? maxr : room;
                const uint32_t r = c <= need ? c : need + RUN<COOP>(in, p + need, b0, c - need);
                if (r < c || maxr <= room) return {T_RLE, r};
                return {T_RLE, LEN_CROSS};
            }
        }
    }
    /* match (valid by construction); the extension only runs past a capped length, :288-320 */
    if (L < PRZM_MAX_MATCH || L >= n - p) return {T_MATCH, L};
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


/* Orders LDS accesses of one wave for the compiler only: the LDS executes a
 * wave's DS instructions in issue order, so no hardware wait is needed.  It
 * also marks the points where the code relies on the 64 lanes moving in
 * lockstep (all lanes' reads of an instruction before any lane's later write). */
__device__ __forceinline__ void lds_fence() {
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("" ::: "memory");
}

/* Lanes of the wave whose digit d (< 256) equals this lane's, valid lanes only, and the
 * counter cnt[d] as it was before this step: each lane sets its bit in the wave's word pb[d]
 * (one LDS atomic OR), then reads the word and the counter back and clears the word, so the
 * words are zero between calls.  The LDS executes a wave's operations in order, so all ORs
 * land before any read and all reads before any clear. */
__device__ __forceinline__ uint64_t peers_lds(uint64_t *pb, const uint32_t *cnt, uint32_t d, bool valid,
                                              uint32_t *cur) {
    const uint64_t me = 1ull << __lane_id();
    if (valid) atomicOr((unsigned long long *)&pb[d], (unsigned long long)me);
    lds_fence();
    uint64_t pm = 0;
    uint32_t c = 0;
    if (valid) {
        pm = pb[d];
        c = cnt[d];
    }
    lds_fence();
    if (valid) pb[d] = 0ull;
    *cur = c;
    return pm;
}

/* ---------------- phase C: one bucket ---------------- */

/* a match result whose store is held back (one per lane): stores count in the same counter as
 * loads, so the positions loaded at the next bucket's start would otherwise wait for the last
 * chunk's stores; they are written once those loads are used.  a == NONE32: none */
struct Pend {
    uint32_t v, a;
};

struct Scr {
    const uint16_t *T0, *T1;
    uint32_t *D;      /* per position: match distance | min(length, 255) << 16, 0 for no match */
    uint8_t *gws; /* per-wave workspaces, GWS_STRIDE bytes each */
};

/* index of the lowest set bit, 0xFFFFFFFF when x == 0: v_ffbl_b32 itself, so that the compiler
 * does not add a zero test around it (the C form is for host-side compilation of this file) */
__device__ __forceinline__ uint32_t ffbl(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    __asm__("v_ffbl_b32 %0, %1" : "=v"(r) : "v"(x));
    return r;
#else
    return x ? (uint32_t)__builtin_ctz(x) : 0xFFFFFFFFu;
#endif
}
__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

/* Bit index of the first difference of two 12-byte keys given as the xor of their words,
 * 0xFFFFFFFF when they are equal: the length of their common prefix is that index / 8.
 * ffbl(x) < 32 for x != 0, so "| 32" adds 32 and keeps 0xFFFFFFFF for equal words. */
__device__ __forceinline__ uint32_t diff12(uint32_t x0, uint32_t x1, uint32_t x2) {
    return umin(umin(ffbl(x0), ffbl(x1) | 32u), ffbl(x2) | 64u);
}

/* bytes p..p+15 as four little-endian words */
__device__ __forceinline__ uint4 ldw16(const uint8_t *in, uint32_t p) {
    const uint32_t *wp = (const uint32_t *)(in + (p & ~3u));
    const uint32_t sh = p & 3u, w0 = wp[0], w1 = wp[1], w2 = wp[2], w3 = wp[3], w4 = wp[4];
    return make_uint4(__builtin_amdgcn_alignbyte(w1, w0, sh), __builtin_amdgcn_alignbyte(w2, w1, sh),
                      __builtin_amdgcn_alignbyte(w3, w2, sh), __builtin_amdgcn_alignbyte(w4, w3, sh));
}

/* number of equal leading bytes (0..16) of two 16-byte strings, branch-free */
__device__ __forceinline__ uint32_t eq16(uint4 a, uint4 b) {
    const uint32_t t = umin(diff12(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z), ffbl(a.w ^ b.w) | 96u);
    return umin(t >> 3, 16u);
}

/* Common prefix of in[p..] and in[q..] (q < p) known to agree on 12 bytes, at most cap (> 12).
 * One LDS round trip per 16 bytes: the ten words of both sides are loaded before any is used. */
__device__ __forceinline__ uint32_t lcp_ext(const uint8_t *in, uint32_t p, uint32_t q, uint32_t cap) {
    uint32_t l = 12u, e;
    do {
        const uint32_t pa = p + l, pb = q + l;
        const uint32_t *wa = (const uint32_t *)(in + (pa & ~3u)), *wb = (const uint32_t *)(in + (pb & ~3u));
        const uint32_t a0 = wa[0], a1 = wa[1], a2 = wa[2], a3 = wa[3], a4 = wa[4];
        const uint32_t b0 = wb[0], b1 = wb[1], b2 = wb[2], b3 = wb[3], b4 = wb[4];
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t sa = pa & 3u, sb = pb & 3u;
        const uint4 a = make_uint4(__builtin_amdgcn_alignbyte(a1, a0, sa), __builtin_amdgcn_alignbyte(a2, a1, sa),
                                   __builtin_amdgcn_alignbyte(a3, a2, sa), __builtin_amdgcn_alignbyte(a4, a3, sa));
        const uint4 b = make_uint4(__builtin_amdgcn_alignbyte(b1, b0, sb), __builtin_amdgcn_alignbyte(b2, b1, sb),
                                   __builtin_amdgcn_alignbyte(b3, b2, sb), __builtin_amdgcn_alignbyte(b4, b3, sb));
        e = eq16(a, b);
        l += e;
    } while (e == 16u && l < cap);
    return umin(l, cap);
}

/*
 * Search elements [i0, i1) (sorted order) of a bucket whose positions in group
 * order are SRT and per-digit group starts GE.  Element i's chain is
 * the up to `depth` elements before it in its group (same 16-bit hash,
 * newest first, compress.go:681-707); the longest common prefix wins, ties
 * keep the nearest, and only lengths >= 4 count.
 *
 * Each wave stages the first 12 bytes and the position of the current and the
 * previous chunk of 64 elements in an LDS ring, so a candidate costs one
 * conflict-free 16-byte read and a branch-free compare.  A candidate's score is
 * length << 16 | its position, so the best of a chain (longest, then nearest) is
 * a plain maximum of scores.
 *
 * No per-candidate range test: a ring entry before the element's group start
 * is another element of the same bucket with a different 16-bit hash, whose
 * first 4 bytes therefore differ (length <= 3, never taken), and the ring slots
 * of the previous chunk hold, on the first chunk, a key whose hash lies in
 * another bucket.  Chains longer than the ring (depth > 64) read the rest from
 * the block, with the same filler key past the group start.
 */
template <bool DEEP, uint32_t B, typename PT, typename ET>
__device__ __forceinline__ void bucket_search_range(LzShared &S, const Scr &sc, uint32_t dgt, uint32_t n,
                                                    uint32_t depth, const PT *SRT, const ET *GE,
                                                    uint32_t i0, uint32_t i1, uint32_t step, Pend &pd,
                                                    uint64_t *pt = nullptr) {
    const bool prof = pt && threadIdx.x == 0;
    const uint8_t *in = S.in;
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint4 *ring = S.big.ws.ring[w];
    /* a 4-byte word whose hash4 high byte is not dgt: hash4(0) = 0, hash4(1) = 0x1E35 */
    const uint32_t fz = dgt == 0 ? 1u : 0u;
    const uint4 filler = make_uint4(fz, fz, fz, 0u);
    /* this lane's position in the current chunk (pc) and the next (pn), and the four words that
     * hold the current key's bytes: loaded one chunk ahead (indexes past the range repeat its
     * last element), so that staging a chunk waits only for its group starts */
    const uint32_t last = i1 - 1u;
    uint32_t pc = 0, pn = 0, kw0 = 0, kw1 = 0, kw2 = 0, kw3 = 0;
    if (i0 < i1) {
        pc = SRT[umin(i0 + lane, last)];
        pn = SRT[umin(i0 + step + lane, last)];
        const uint32_t *wp = (const uint32_t *)(in + (pc & ~3u));
        kw0 = wp[0]; kw1 = wp[1]; kw2 = wp[2]; kw3 = wp[3];
    }
    uint4 key = make_uint4(0, 0, 0, 0);
    for (uint32_t c0 = i0; c0 < i1; c0 += step) {
        const uint64_t t_a = prof ? __builtin_amdgcn_s_memrealtime() : 0;
        const uint32_t i = c0 + lane;
        const bool valid = i < i1;
        /* ring[0, 64): the previous chunk, ring[64, 128): this chunk, so candidate k of
         * lane l is entry 64 + l - k (k <= 64 + l) */
        if (c0 < 64) ring[lane] = filler;
        else if (step == 64 && c0 != i0) ring[lane] = key; /* this lane's key of the last chunk, which was full */
        else {
            const uint32_t ip = c0 - 64 + lane, pp = SRT[ip];
            const uint4 k = ldw16(in, pp);
            ring[lane] = make_uint4(k.x, k.y, k.z, pp);
        }
        const uint32_t sh = pc & 3u;
        key = make_uint4(__builtin_amdgcn_alignbyte(kw1, kw0, sh), __builtin_amdgcn_alignbyte(kw2, kw1, sh),
                         __builtin_amdgcn_alignbyte(kw3, kw2, sh), pc);
        /* lanes past the range stage a copy of its last element, which no chain of the range reads */
        ring[64 + lane] = key;
        uint32_t p = 0, kmax = 0, cap = 0;
        if (valid) {
            p = pc;
            const uint32_t low = ((key.x * 0x1E35A7BDu) >> 16) & 0xFFu; /* hash4 of p from its key */
            const uint32_t gs = GE[low]; /* where the element's group starts */
            cap = n - p < PRZM_MAX_MATCH ? n - p : PRZM_MAX_MATCH;
            kmax = i - gs < depth ? i - gs : depth;
        } else {
            key = make_uint4(0, 0, 0, 0);
        }
        if (c0 + step < i1) { /* the next chunk's key words and the position after it */
            const uint32_t *wp = (const uint32_t *)(in + (pn & ~3u));
            kw0 = wp[0]; kw1 = wp[1]; kw2 = wp[2]; kw3 = wp[3];
            pc = pn;
            pn = SRT[umin(c0 + 2u * step + lane, last)];
        }
        lds_fence();
        const uint64_t t_b = prof ? __builtin_amdgcn_s_memrealtime() : 0;
        /* best = L << 16 | q of the best candidate so far.  A group lists its positions in
         * ascending order, so a chain's candidates come nearest first with descending q and a
         * plain maximum keeps the nearest of equal lengths.  L = 3 to start: lengths below 4 do
         * not count.  capb caps a key compare at min(cap, 12) bytes; lim12 is the score from
         * which a candidate agrees on all 12 key bytes with more to compare (never when cap <= 12). */
        uint32_t best = 3u << 16;
        const uint32_t capb = 8u * umin(cap, 12u) + 7u, bcap = cap << 16;
        const uint32_t lim12 = cap > 12u ? (12u << 16) : 0xFFFFFFFFu;
        for (uint32_t k0 = 1; k0 <= depth; k0 += B) { /* depth is a multiple of B */
            if (!__ballot(k0 <= kmax && best < bcap)) break;
            uint32_t sco[B];
            /* candidates k0 .. k0 + B - 1 sit below one another in the ring */
            const uint4 *rq = ring + (64 + (int32_t)lane - (int32_t)k0 - (int32_t)(B - 1));
            uint32_t mc = 0;
#pragma unroll
            for (uint32_t u = 0; u < B; u++) { /* in flight together */
                const uint32_t k = k0 + u;
                uint4 qk;
                if (DEEP && k > lane + 64u) {
                    if (k <= kmax) {
                        const uint32_t q = (uint32_t)SRT[i - k];
                        const uint4 b = ldw16(in, q);
                        qk = make_uint4(b.x, b.y, b.z, q);
                    } else {
                        qk = filler;
                    }
                } else {
                    qk = rq[B - 1 - u];
                }
                const uint32_t lb = umin(diff12(key.x ^ qk.x, key.y ^ qk.y, key.z ^ qk.z), capb);
                sco[u] = ((lb >> 3) << 16) | qk.w;
                mc = mc > sco[u] ? mc : sco[u];
            }
            if (__ballot(mc >= lim12)) {
                /* Some candidate agrees on all 12 key bytes: its exact length decides.  While
                 * the best is shorter than 12 bytes, the nearest such candidate of the step is
                 * extended first (one lcp_ext per lane).  Then a candidate can only beat the
                 * best if it agrees at byte bl, the best's length (one that differs there is at
                 * most as long and farther, so it loses a tie): those bytes of the step's other
                 * 12-byte candidates are read together, and only the ones that agree are
                 * extended. */
                uint32_t qf = 0, uf = B;
#pragma unroll
                for (uint32_t u = B; u-- > 0;)
                    if (sco[u] >= lim12) { qf = sco[u] & 0xFFFFu; uf = u; }
                /* the step's key scores are exact below 12 bytes and lower bounds at 12, which
                 * the extended nearest candidate, or a best of 12 bytes or more from a nearer
                 * step, dominates */
                if (best < lim12 && uf < B) {
                    const uint32_t sf = (lcp_ext(in, p, qf, cap) << 16) | qf;
                    best = best > sf ? best : sf;
                } else {
                    uf = B; /* no candidate extended in this step */
                }
                best = best > mc ? best : mc;
                const uint32_t bl = best >> 16;
                if (bl < cap) {
                    const uint32_t pb = in[p + bl];
                    uint32_t pass = 0;
#pragma unroll
                    for (uint32_t u = 0; u < B; u++) {
                        const uint32_t q = sco[u] & 0xFFFFu;
                        const bool cand = sco[u] >= lim12 && u != uf;
                        pass |= (cand && in[q + bl] == pb) ? 1u << u : 0u;
                    }
#pragma unroll
                    for (uint32_t u = 0; u < B; u++) {
                        if ((pass >> u) & 1u) {
                            const uint32_t q = sco[u] & 0xFFFFu, su = (lcp_ext(in, p, q, cap) << 16) | q;
                            best = best > su ? best : su;
                        }
                    }
                }
            } else {
                best = best > mc ? best : mc;
            }
        }
        lds_fence();
        const uint64_t t_c = prof ? __builtin_amdgcn_s_memrealtime() : 0;
        if (pd.a != NONE32) sc.D[pd.a] = pd.v; /* the previous chunk's result */
        { /* by position: dist | min(len, 255) << 16; "no match" (0) is already in place */
            const uint32_t L = best >> 16;
            pd.v = (p - (best & 0xFFFFu)) | (umin(L, 255u) << 16);
            pd.a = valid && L >= 4u ? p : NONE32;
        }
        if (prof) {
            const uint64_t t_d = __builtin_amdgcn_s_memrealtime();
            pt[1] += t_b - t_a; pt[2] += t_c - t_b; pt[3] += t_d - t_c; pt[4] += 1;
        }
    }
}

/* group the bucket by the hash's low byte (stable), then search it unless it
 * is large; returns true when the bucket was queued for the all-wave search */
/* the first 256 positions of bucket dgt, element u * 64 + lane in pre[u] (undefined past the
 * bucket's end).  The loads are unconditional (an index past the end reads T0[0]) so that no
 * branch makes their number unknown to the wait-count bookkeeping. */
__device__ __forceinline__ void bucket_fetch(const LzShared &S, const Scr &sc, uint32_t dgt, uint32_t pre[4]) {
    const uint32_t lane = __lane_id();
    uint32_t s0 = 0, m0 = 0, s1 = 0, m1 = 0;
    if (dgt < 256) {
        s0 = S.bstart[0][dgt]; m0 = S.bstart[0][dgt + 1] - s0;
        s1 = S.bstart[1][dgt]; m1 = S.bstart[1][dgt + 1] - s1;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const uint32_t j = u * 64 + lane;
        const uint16_t *a = j < m0 ? sc.T0 + s0 + j : j < m0 + m1 ? sc.T1 + s1 + (j - m0) : sc.T0;
        pre[u] = *a;
    }
}

template <bool LDSWS, uint32_t B>
__device__ bool bucket_sort(LzShared &S, const Scr &sc, uint32_t w, uint32_t dgt, uint32_t n, uint32_t depth,
                            bool may_queue, const uint32_t pre[4], Pend &pd, uint64_t *pt) {
    /* pt (diagnostics, thread 0 only): accumulated sort, chunk staging, candidate scan and result
     * write times, and chunks */
    const bool prof = pt && threadIdx.x == 0;
    const uint64_t t_a = prof ? __builtin_amdgcn_s_memrealtime() : 0;
    uint64_t t_c = 0;
    const uint8_t *in = S.in;
    const uint32_t lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t s0 = S.bstart[0][dgt], m0 = S.bstart[0][dgt + 1] - s0;
    const uint32_t s1 = S.bstart[1][dgt], m1 = S.bstart[1][dgt + 1] - s1;
    const uint32_t m = m0 + m1;
    uint8_t *gw = sc.gws + (uint64_t)w * GWS_STRIDE + (may_queue ? 0u : GWS_SLOT);
    uint16_t *P = LDSWS ? S.big.ws.pos[w] : (uint16_t *)gw;
    uint16_t *SRT = LDSWS ? S.big.ws.srt[w] : (uint16_t *)gw + PRZM_BLOCK;
    uint32_t *cnt = S.sm.tab[w];
    for (int k = 0; k < 4; k++) cnt[lane * 4 + k] = 0;
    lds_fence();
    /* the first 256 elements' positions and low digits stay in registers for the scatter, */
    uint32_t lw0[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const uint32_t j = u * 64 + lane;
        lw0[u] = hash4(in, pre[u]) & 0xFFu;
        if (j < m) {
            P[j] = (uint16_t)pre[u];
            atomicAdd(&cnt[lw0[u]], 1u);
        }
    }
    /* and so do the next 256 (most buckets hold fewer than 512) */
    uint32_t pv1[4] = {0u, 0u, 0u, 0u}, lw1[4] = {0u, 0u, 0u, 0u};
    if (m > 256) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t j = 256 + u * 64 + lane;
            pv1[u] = j < m ? (j < m0 ? sc.T0[s0 + j] : sc.T1[s1 + j - m0]) : 0u;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t j = 256 + u * 64 + lane;
            lw1[u] = hash4(in, pv1[u]) & 0xFFu;
            if (j < m) {
                P[j] = (uint16_t)pv1[u];
                atomicAdd(&cnt[lw1[u]], 1u);
            }
        }
    }
    for (uint32_t c0 = 512; c0 < m; c0 += 256) {
        uint32_t pv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t j = c0 + u * 64 + lane;
            pv[u] = j < m ? (j < m0 ? sc.T0[s0 + j] : sc.T1[s1 + j - m0]) : 0u;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t j = c0 + u * 64 + lane;
            if (j < m) {
                P[j] = (uint16_t)pv[u];
                atomicAdd(&cnt[hash4(in, pv[u]) & 0xFFu], 1u);
            }
        }
    }
    if (!LDSWS) __threadfence_block();
    lds_fence();
    if (prof) pt[-2] += __builtin_amdgcn_s_memrealtime() - t_a;
    if (pd.a != NONE32) { /* this bucket's positions are in: the held-back result can go now */
        sc.D[pd.a] = pd.v;
        pd.a = NONE32;
    }
    /* Group starts.  A chunk of 64 elements takes as many search steps as its longest chain, so
     * the groups are laid out by class, each class in digit order: chunks then hold chains of
     * like length.  A chain only needs its group to be contiguous and in position order, so
     * the layout changes no result.  One wave scan serves all classes, each counted in its own
     * bit field of the scanned value: for a bucket in LDS three 11-bit fields (chains longer
     * than depth, shorter chains, single positions), otherwise two (chains in the low 20 bits,
     * at most 65536 elements; single positions from bit 20, at most 256). */
    uint32_t msing; /* where the single positions start: they have no chain, hence no match */
    {
        uint16_t *GS = S.sm.gstart[w];
        uint32_t c4[4];
#pragma unroll
        for (int k = 0; k < 4; k++) c4[k] = cnt[lane * 4 + k];
        if (m <= WS_CAP) {
            /* a bucket kept in LDS (at most 1536 elements) has room for three 11-bit fields:
             * chains longer than depth first (their chunks all take every step), then the
             * shorter chains, then the single positions */
            uint32_t tp = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) tp += c4[k] == 1u ? (1u << 22) : c4[k] > depth ? c4[k] : (c4[k] << 11);
            const uint32_t incl = wave_incl_sum(tp), excl = incl - tp;
            const uint32_t all = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            const uint32_t nf = all & 0x7FFu, nsh = (all >> 11) & 0x7FFu;
            uint32_t rf = excl & 0x7FFu, rh = nf + ((excl >> 11) & 0x7FFu), rs = nf + nsh + (excl >> 22);
            msing = nf + nsh;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const bool single = c4[k] == 1u, full = c4[k] > depth;
                const uint32_t st = single ? rs : full ? rf : rh;
                rs += single ? 1u : 0u;
                rf += full ? c4[k] : 0u;
                rh += (!single && !full) ? c4[k] : 0u;
                cnt[lane * 4 + k] = st;
                GS[lane * 4 + k] = (uint16_t)st;
            }
        } else {
            uint32_t tp = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) tp += c4[k] == 1u ? (1u << 20) : c4[k];
            const uint32_t incl = wave_incl_sum(tp), excl = incl - tp;
            const uint32_t all = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            uint32_t rc = excl & 0xFFFFFu, rs = (all & 0xFFFFFu) + (excl >> 20);
            msing = all & 0xFFFFFu;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const bool single = c4[k] == 1u;
                const uint32_t st = single ? rs : rc;
                rs += single ? 1u : 0u;
                rc += single ? 0u : c4[k];
                cnt[lane * 4 + k] = st;
                GS[lane * 4 + k] = (uint16_t)st;
            }
        }
    }
    lds_fence();
    auto scatter = [&](uint32_t j, uint32_t pj, uint32_t low) {
        const bool valid = j < m;
        low = valid ? low : 0u;
        uint32_t base;
        const uint64_t pm = peers_lds(S.sm.pbits[w], cnt, low, valid, &base);
        if (valid) {
            SRT[base + (uint32_t)__popcll(pm & lt)] = (uint16_t)pj; /* the position itself, in group order */
            if ((pm & lt) == 0) cnt[low] = base + (uint32_t)__popcll(pm);
        }
        lds_fence();
    };
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) {
        if (u * 64 >= m) break;
        scatter(u * 64 + lane, pre[u], lw0[u]);
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) {
        if (256 + u * 64 >= m) break;
        scatter(256 + u * 64 + lane, pv1[u], lw1[u]);
    }
    for (uint32_t c0 = 512; c0 < m; c0 += 64) {
        const uint32_t j = c0 + lane;
        const uint32_t pj = j < m ? P[j] : 0u;
        scatter(j, pj, hash4(in, pj) & 0xFFu);
    }
    if (!LDSWS) __threadfence_block();
    lds_fence();
    if (prof) t_c = __builtin_amdgcn_s_memrealtime();
    if (!LDSWS && may_queue) {
        uint32_t *ge = (uint32_t *)(gw + 2u * PRZM_BLOCK * 2u);
        for (int k = 0; k < 4; k++) ge[lane * 4 + k] = S.sm.gstart[w][lane * 4 + k];
        __threadfence_block();
        return true;
    }
    if (prof) pt[0] += t_c - t_a;
    const uint16_t *GS = S.sm.gstart[w];
    /* the groups with a chain are searched; the single positions come last and keep "no match" */
    if (depth > 64u) bucket_search_range<true, B>(S, sc, dgt, n, depth, SRT, GS, 0, msing, 64, pd, pt);
    else if (depth < B) bucket_search_range<false, 2>(S, sc, dgt, n, depth, SRT, GS, 0, msing, 64, pd, pt);
    else bucket_search_range<false, B>(S, sc, dgt, n, depth, SRT, GS, 0, msing, 64, pd, pt);
    return false;
}

/*
 * A segment's cached tokens (at most 8) in registers: toff and tlen rows as they lie in LDS
 * (entry k of the segment at index k ^ (s & 7)), and the distances of its matches in order
 * (tdist[m_base ..]).
 */
struct SegTok {
    uint64_t off;      /* toff row: byte i */
    uint64_t len[2];   /* tlen row: u16 i in len[i >> 2] */
    uint64_t dist[2];  /* u16 r: distance of the segment's match number r */
};
__device__ __forceinline__ uint32_t u16_of(const uint64_t (&v)[2], uint32_t i) {
    return (uint32_t)(((i < 4u ? v[0] : v[1]) >> (16u * (i & 3u))) & 0xFFFFu);
}
__device__ __forceinline__ SegTok seg_tok_load(const LzShared &S, uint32_t s, uint32_t m_base) {
    SegTok t;
    t.off = *(const uint64_t *)&S.big.q.toff[s][0];
    const uint4 l = *(const uint4 *)&S.big.q.tlen[s][0];
    t.len[0] = ((uint64_t)l.y << 32) | l.x;
    t.len[1] = ((uint64_t)l.w << 32) | l.z;
    uint32_t d[8];
#pragma unroll
    for (uint32_t r = 0; r < 8; r++) d[r] = m_base + r < QP / 4 ? S.sm.ps.tdist[m_base + r] : 0u;
    t.dist[0] = ((uint64_t)(d[3] << 16 | d[2]) << 32) | (d[1] << 16 | d[0]);
    t.dist[1] = ((uint64_t)(d[7] << 16 | d[6]) << 32) | (d[5] << 16 | d[4]);
    return t;
}

/* The three most recent distinct match distances before match number i (most recent first),
 * from tdist[i - 1], tdist[i - 2], ... and then the ring carried into the tile
 * (przm_tokenizer.cuh ring registers). */
__device__ __forceinline__ void ring_before(const LzShared &S, uint32_t i, uint32_t &r0, uint32_t &r1,
                                            uint32_t &r2) {
    uint32_t g0 = 0, g1 = 0, g2 = 0, ns = 0;
    for (uint32_t j = i; j-- > 0 && ns < 3;) {
        const uint32_t v = S.sm.ps.tdist[j];
        if (ns == 0) { g0 = v; ns = 1; }
        else if (v != g0 && (ns == 1 || v != g1)) {
            if (ns == 1) { g1 = v; ns = 2; }
            else { g2 = v; ns = 3; }
        }
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const uint32_t v = j == 0 ? S.st[ST_R0] : j == 1 ? S.st[ST_R1] : S.st[ST_R2];
        if (ns == 0) { g0 = v; ns = 1; }
        else if (ns < 3 && v != g0 && (ns == 1 || v != g1)) {
            if (ns == 1) { g1 = v; ns = 2; }
            else { g2 = v; ns = 3; }
        }
    }
    r0 = g0; r1 = g1; r2 = g2;
}

/* ring index of distance d (0..2, or -1 for an explicit offset), and the ring after it */
__device__ __forceinline__ int ring_use(uint32_t d, uint32_t &r0, uint32_t &r1, uint32_t &r2) {
    const int kk = d == r0 ? 0 : d == r1 ? 1 : d == r2 ? 2 : -1;
    r2 = kk == 0 || kk == 1 ? r2 : r1;
    r1 = kk == 0 ? r1 : r0;
    r0 = d;
    return kk;
}

/* the byte staging area of the write pass: the dist and lc arrays, dead once the tokens
 * are listed; each stream takes its quarter's bytes plus 32 of slack for the 16-byte reads
 * of the copy-out, rounded up to 16 */
#define STG_BYTES ((DSZ * 2u + LSZ) & ~15u)
#ifndef STG_LIMIT
#define STG_LIMIT STG_BYTES /* a build with 0 here writes every quarter directly */
#endif
__device__ __forceinline__ uint32_t stg_room(uint32_t n) { return (n + 32u + 15u) & ~15u; }

/* the literal bytes in[cur .. p) (at most 32, inside one segment; none when the previous token
 * ends at or past p) to lit[lb ..]: the 36 bytes around them are loaded at once, then stored
 * one by one */
template <typename P>
__device__ __forceinline__ void copy_run(P lit, uint32_t &lb, const uint8_t *in, uint32_t cur, uint32_t p) {
    if (p <= cur) return;
    const uint32_t L = p - cur;
    const uint32_t *w = (const uint32_t *)(in + (cur & ~3u));
    const uint32_t sh = cur & 3u;
    if (L <= 4) { /* the common short gap between tokens: two words */
        const uint32_t a4 = __builtin_amdgcn_alignbyte(w[1], w[0], sh);
        lit[lb] = (uint8_t)a4;
        if (L > 1) lit[lb + 1] = (uint8_t)(a4 >> 8);
        if (L > 2) lit[lb + 2] = (uint8_t)(a4 >> 16);
        if (L > 3) lit[lb + 3] = (uint8_t)(a4 >> 24);
        lb += L;
        return;
    }
    uint32_t wd[9];
#pragma unroll
    for (int k = 0; k < 9; k++) wd[k] = w[k];
#pragma unroll
    for (uint32_t g = 0; g < 8; g++) { /* four bytes per step */
        if (4u * g >= L) break;
        const uint32_t v = __builtin_amdgcn_alignbyte(wd[g + 1], wd[g], sh);
        const uint32_t o = lb + 4u * g, r = L - 4u * g;
        lit[o] = (uint8_t)v;
        if (r > 1) lit[o + 1] = (uint8_t)(v >> 8);
        if (r > 2) lit[o + 2] = (uint8_t)(v >> 16);
        if (r > 3) lit[o + 3] = (uint8_t)(v >> 24);
    }
    lb += L;
}

/*
 * The byte writes of one segment's tokens (przm_tokenizer.cuh:330-437 emission): literal gaps,
 * then each token.  Each stream is a base pointer and this segment's first index in it: the
 * LDS staging areas with offsets inside the quarter, or the streams themselves with absolute
 * offsets (the function is inlined into both, so each keeps its address space).
 */
template <typename LP>
__device__ __forceinline__ void write_tokens(LzShared &S, uint32_t s, uint32_t ss, uint32_t se, uint32_t t0,
                                             uint32_t nt, const SegTok &tk, uint32_t kkp, uint32_t ls_in,
                                             LP lit, uint32_t lb, LP cmd, uint32_t b_cmd, LP op, uint32_t b_op,
                                             gu8 *meta, uint32_t b_meta, gu8 *var, uint32_t b_var, gu8 *off,
                                             uint32_t b_off) {
    const uint8_t *in = S.in;
    uint32_t mr = 0, ls = ls_in, cur = t0;
    for (uint32_t k = 0; k <= nt; k++) {
        const uint32_t sk = k ^ (s & 7u);
        const bool tok = k < nt;
        const uint32_t o = tok ? (uint32_t)(tk.off >> (8u * (sk & 7u))) & 0xFFu : 0u;
        const uint32_t len = tok ? u16_of(tk.len, sk & 7u) : 0u;
        const uint32_t p = tok ? ss + (o & 31u) : se;
        copy_run(lit, lb, in, cur, p);
        if (!tok) break;
        const uint32_t ll = p - ls;
        if (o & 0x20u) {
            if (ll) {
                if (ll <= 8) { cmd[b_cmd++] = (uint8_t)(0xF2 + ll); op[b_op++] = (uint8_t)(0xF2 + ll); }
                else {
                    cmd[b_cmd++] = 0xF0; op[b_op++] = 0xF0;
                    b_cmd = put_leb(cmd, b_cmd, ll); b_var = put_leb(var, b_var, ll);
                }
            }
            cmd[b_cmd++] = 0xF1; op[b_op++] = 0xF1;
            b_cmd = put_leb(cmd, b_cmd, len); b_var = put_leb(var, b_var, len);
            lit[lb++] = in[p];
        } else {
            const uint32_t d = u16_of(tk.dist, mr++);         /* this match's distance */
            const int kk = (int)((kkp >> (2u * k)) & 3u) - 1; /* the ring index the size pass found */
            const uint32_t keep = kk >= 0 ? (len <= 7 ? 7u : 15u) : 3u;
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
            const uint32_t base = len - 4 > 14 ? 15u : len - 4;
            if (kk >= 0 && len <= 7) {
                const uint8_t oc = (uint8_t)((l2 << 4) | ((len - 4) << 2) | (uint32_t)kk);
                cmd[b_cmd++] = oc; op[b_op++] = oc;
            } else if (kk >= 0) {
                const uint8_t tok = (uint8_t)((l2 << 4) | base);
                cmd[b_cmd++] = (uint8_t)(0xC0 | kk); cmd[b_cmd++] = tok;
                op[b_op++] = (uint8_t)(0xC0 | kk); meta[b_meta++] = tok;
                if (base == 15) { b_cmd = put_leb(cmd, b_cmd, len - 19); b_var = put_leb(var, b_var, len - 19); }
            } else {
                const uint8_t oc = (uint8_t)(0x80 | (l2 << 4) | base);
                cmd[b_cmd++] = oc; op[b_op++] = oc;
                if (base == 15) { b_cmd = put_leb(cmd, b_cmd, len - 19); b_var = put_leb(var, b_var, len - 19); }
                b_off = put_leb(off, b_off, d);
            }
        }
        ls = p + len;
        cur = p + len;
    }
}

/* n bytes from the LDS area src (4-byte aligned, readable 20 bytes past n) to dst: the
 * workgroup stores the 16-byte aligned body of dst as uint4s, the ends as bytes */
__device__ __forceinline__ void lds_copy_out(gu8 *__restrict__ dst, const uint8_t *src, uint32_t n) {
    const uint32_t tid = threadIdx.x;
    const uint32_t mis = (uint32_t)((uintptr_t)(const void *)(const uint8_t *)dst & 15u);
    const uint32_t head = umin(n, (16u - mis) & 15u);
    if (tid < head) dst[tid] = src[tid];
    const uint32_t body = (n - head) & ~15u;
    typedef __attribute__((address_space(1))) uint32_t gu32;
    for (uint32_t o = head + tid * 16u; o < head + body; o += LZ_NT * 16u) {
        const uint4 v = ldw16(src, o);
        gu32 *d4 = (gu32 *)(dst + o);
        d4[0] = v.x; d4[1] = v.y; d4[2] = v.z; d4[3] = v.w;
    }
    const uint32_t t0 = head + body;
    if (tid < n - t0) dst[t0 + tid] = src[t0 + tid];
}

/* ---------------- parse of one tile ---------------- */

/* bits [x, y) of a 32-bit mask, 0 <= x <= y <= 32 */
__device__ __forceinline__ uint32_t bits_range(uint32_t x, uint32_t y) {
    return (uint32_t)((1ull << y) - (1ull << x));
}

/* the first position at or after p (ss <= p < se) that is not a certain literal, or se */
__device__ __forceinline__ uint32_t lit_run_end(uint32_t litm, uint32_t ss, uint32_t p, uint32_t se) {
    const uint32_t nl = ~litm & (0xFFFFFFFFu << (p - ss));
    const uint32_t q = nl ? ss + (uint32_t)__builtin_ctz(nl) : ss + 32u;
    return q < se ? q : se;
}

/* the token at p, a position of segment ss that is not a certain literal: eval_tok where the
 * class mask says it must decide, else a match of the stored length */
__device__ __forceinline__ Tok seg_tok(const LzShared &S, uint32_t cxm, uint32_t ss, uint32_t p, uint32_t a,
                                       uint32_t n, uint32_t lim) {
    if ((cxm >> (p - ss)) & 1u) return eval_tok<false>(S, p, a, n, lim);
    return {T_MATCH, lget(S, p - a)};
}


__device__ void parse_tile(LzShared &S, uint32_t a, uint32_t b, uint32_t n, uint8_t *__restrict__ wsb,
                           uint64_t *__restrict__ pts) {
#ifdef PRZM_DIAG
#define QTS(i) do { if (pts && threadIdx.x == 0) ((__attribute__((address_space(1))) uint64_t *)pts)[i] = __builtin_amdgcn_s_memrealtime(); } while (0)
#else
#define QTS(i) ((void)0)
#endif
    const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    /* the streams are in global memory: said in their type, since this function is called
     * (not inlined) and would otherwise store through flat addresses, whose completion every
     * later LDS wait would also wait for */
    gu8 *const wsg = (gu8 *)wsb;
    gu8 *cmd = wsg + PRZM_WS_CMD, *off = wsg + PRZM_WS_OFF, *lit = wsg + PRZM_WS_LIT;
    gu8 *op = wsg + PRZM_WS_OP, *meta = wsg + PRZM_WS_META, *var = wsg + PRZM_WS_VAR;
    const uint8_t *in = S.in;
    const uint32_t s = tid, ss = a + s * SEG;
    const bool segv = ss < b;
    const uint32_t se = segv ? (ss + SEG < b ? ss + SEG : b) : ss;
    const uint32_t nseg = (b - a + SEG - 1) / SEG;

    /* 0. classes of the segment's positions, from the decision inputs alone
     *    (przm_tokenizer.cuh:171-320 as in eval_tok): litm = certainly a literal (no match,
     *    or the lookahead prefers the next position); cxm = only eval_tok can decide (a
     *    length of 255 or more, a lookahead that needs one, a possible byte run); any other
     *    position is a match of its stored length.  The walks below then take a literal run
     *    in one step. */
    uint32_t litm = 0, cxm = 0;
    if (segv) {
        const uint32_t h0 = S.sm.ps.hm[s], h1 = S.sm.ps.hm[s + 1];
        const uint32_t *lwp = (const uint32_t *)(S.big.q.lc + LIX(s * SEG));
        const uint32_t *iwp = (const uint32_t *)(in + ss);
        uint32_t lw[10], iw[9];
#pragma unroll
        for (int j = 0; j < 10; j++) lw[j] = lwp[j];
#pragma unroll
        for (int j = 0; j < 9; j++) iw[j] = iwp[j];
        /* per position k the byte predicates as bits of 32-bit masks; the position
         * conditions (p + 1 < n, p + 12 <= n, p < se) and the logic then act on whole masks */
        uint32_t E = 0, G = 0, E1 = 0, GT = 0, RW = 0;
#pragma unroll
        for (uint32_t k = 0; k < SEG; k++) {
            const uint32_t lc = (lw[k >> 2] >> (8u * (k & 3u))) & 0xFFu;
            const uint32_t l1c = k + 1 < SEG ? (lw[(k + 1) >> 2] >> (8u * ((k + 1) & 3u))) & 0xFFu : lw[9] & 0xFFu;
            const uint32_t w4 = (k & 3u) ? __builtin_amdgcn_alignbyte(iw[(k >> 2) + 1], iw[k >> 2], k & 3u) : iw[k >> 2];
            E |= (uint32_t)(lc == 255u) << k;
            G |= (uint32_t)(lc >= 254u) << k;
            E1 |= (uint32_t)(l1c == 255u) << k;
            GT |= (uint32_t)(l1c > lc + 1u) << k;
            RW |= (uint32_t)(w4 == __builtin_amdgcn_alignbit(w4, w4, 8u)) << k; /* four equal bytes: w4 is its own byte rotation */
        }
        /* positions k < lim of the segment, lim clamped to [0, 32] */
        auto below = [](int32_t lim) -> uint32_t {
            return lim >= 32 ? 0xFFFFFFFFu : lim <= 0 ? 0u : (1u << lim) - 1u;
        };
        const uint32_t PM1 = below((int32_t)n - 1 - (int32_t)ss);   /* p + 1 < n */
        const uint32_t PM12 = below((int32_t)n - 11 - (int32_t)ss); /* p + 12 <= n */
        const uint32_t SE = below((int32_t)(se - ss));               /* p < se */
        const uint32_t D = h0, D1 = (h0 >> 1) | (h1 << 31);
        const uint32_t LA = D1 & PM1;
        const uint32_t LACX = LA & E1 & G;
        const uint32_t LALIT = LA & ~LACX & GT;
        const uint32_t RUN = RW & PM12;
        litm = (~D | (~E & LALIT)) & SE;
        cxm = (D & (E | (~LALIT & (LACX | RUN)))) & SE;
    }

    QTS(8);
    /* 1. speculative walk of every segment from its start */
    if (segv) {
        uint32_t bm = 0, p = ss, xp = 0;
        bool cross = false;
        while (p < se) {
            const uint32_t q = lit_run_end(litm, ss, p, se);
            bm |= bits_range(p - ss, q - ss);
            p = q;
            if (p >= se) break;
            bm |= 1u << (p - ss);
            const Tok t = seg_tok(S, cxm, ss, p, a, n, se);
            if (t.kind == T_LIT) { p++; continue; }
            if (t.len == LEN_CROSS) {
                const Tok t2 = eval_tok<false>(S, p, a, n, se + XROOM);
                if (t2.len == LEN_CROSS) { cross = true; xp = p; break; }
                S.sm.ps.xlen[s] = t2.len | (t2.kind << 30);
                p += t2.len;
                break;
            }
            p += t.len;
        }
        S.sm.ps.bmp[s] = bm;
        S.sm.ps.xpos[s] = cross ? (xp | XCROSS) : p;
    }
    S.sm.ps.ent[s] = NONE32;
    __syncthreads();
    /* 1b. entries from the previous segment's exit, in parallel: when that exit
     * lands in this segment off its path, walk from it until the old path is met */
    for (int round = 0; round < 2; round++) {
        uint32_t xp = NONE32;
        if (segv && s > 0) xp = S.sm.ps.xpos[s - 1];
        __syncthreads();
        if (segv && s > 0 && !(xp & XCROSS) && xp >= ss && xp < se) {
            const uint32_t old = S.sm.ps.bmp[s];
            if (!((old >> (xp - ss)) & 1u)) {
                uint32_t p = xp, nb = 0;
                bool done = false;
                while (p < se) {
                    const uint32_t q = lit_run_end(litm, ss, p, se);
                    const uint32_t run = bits_range(p - ss, q - ss), hit = old & run;
                    if (hit) {
                        nb |= bits_range(p - ss, (uint32_t)__builtin_ctz(hit));
                        S.sm.ps.bmp[s] = old | nb; done = true; break;
                    }
                    nb |= run;
                    p = q;
                    if (p >= se) break;
                    if ((old >> (p - ss)) & 1u) { S.sm.ps.bmp[s] = old | nb; done = true; break; }
                    nb |= 1u << (p - ss);
                    const Tok t = seg_tok(S, cxm, ss, p, a, n, se);
                    if (t.kind == T_LIT) { p++; continue; }
                    if (t.len == LEN_CROSS) {
                        const Tok t2 = eval_tok<false>(S, p, a, n, se + XROOM);
                        if (t2.len == LEN_CROSS) { S.sm.ps.xpos[s] = p | XCROSS; S.sm.ps.bmp[s] = nb; done = true; break; }
                        S.sm.ps.xlen[s] = t2.len | (t2.kind << 30);
                        p += t2.len;
                        break;
                    }
                    p += t.len;
                }
                if (!done) { S.sm.ps.xpos[s] = p; S.sm.ps.bmp[s] = nb; }
            }
        }
        __syncthreads();
    }
    QTS(0);

    /* 2. the true path: orbit of the entry segment under segment ->
     *    segment-of-exit, by pointer doubling with one thread per segment;
     *    entries that are off their segment's speculative path are re-walked
     *    (merging into the old path where they meet) and the tokens crossing
     *    segment ends with an unknown length are resolved (wave 0, wave-wide
     *    compares), until the path is consistent. */
    if (tid == 0) S.st[ST_E0] = S.st[ST_ENTRY];
    __syncthreads();
    {
        const uint32_t sg = s;
        uint32_t iters = 0;
        for (;;) {
            iters++;
            const uint32_t e0 = S.st[ST_E0];
            const uint32_t s0 = (e0 - a) / SEG;
            const bool in = sg >= s0 && sg < nseg;
            if (in) {
                const uint32_t x = S.sm.ps.xpos[sg];
                const uint32_t nx = (x & XCROSS) ? NX_STOP : (x >= b ? NX_END : (x - a) / SEG);
                S.sm.ps.nxt[sg] = (uint16_t)nx;
                S.sm.ps.jmp[sg] = (uint16_t)nx;
                S.sm.ps.vis[sg] = sg == s0;
            }
            if (tid == 0) S.st[ST_BAD] = NSEG;
            __syncthreads();
            for (int r = 0; r < 9; r++) { /* after round r every segment within 2^(r+1) - 1 steps is marked */
                uint32_t jn = NX_END;
                if (in) {
                    const uint32_t jm = S.sm.ps.jmp[sg];
                    if (jm < NSEG) {
                        if (S.sm.ps.vis[sg]) S.sm.ps.vis[jm] = 1;
                        jn = S.sm.ps.jmp[jm];
                    } else jn = jm;
                }
                __syncthreads();
                if (in) S.sm.ps.jmp[sg] = (uint16_t)jn;
                __syncthreads();
            }
            const bool vis = in && S.sm.ps.vis[sg];
            const uint32_t nx = in ? S.sm.ps.nxt[sg] : NX_END;
            if (vis && nx < NSEG) S.sm.ps.prv[nx] = (uint16_t)sg;
            __syncthreads();
            bool ok = false;
            if (in) {
                if (!vis) S.sm.ps.ent[sg] = NONE32;
                else {
                    const uint32_t e = sg == s0 ? e0 : S.sm.ps.xpos[S.sm.ps.prv[sg]];
                    S.sm.ps.ent[sg] = e;
                    ok = (S.sm.ps.bmp[sg] >> (e - (a + sg * SEG))) & 1u;
                    if (!ok || nx == NX_STOP) atomicMin(&S.st[ST_BAD], sg);
                }
            }
            __syncthreads();
            const uint32_t bad = S.st[ST_BAD];
            if (bad == NSEG) {
                /* consistent: the quarter's exit is the exit of the visited segment leaving it */
                if (vis && nx == NX_END) S.st[ST_ENTRY] = S.sm.ps.xpos[sg];
                if (pts && tid == 0) pts[10] = iters;
                break;
            }
            /* re-walk every visited segment from the first problem on whose entry is off its path */
            if (vis && !ok && sg >= bad) {
                const uint32_t sgs = a + sg * SEG, sge = sgs + SEG < b ? sgs + SEG : b;
                uint32_t p = S.sm.ps.ent[sg], nb = 0;
                const uint32_t old = S.sm.ps.bmp[sg];
                bool done = false;
                while (p < sge) {
                    const uint32_t q = lit_run_end(litm, sgs, p, sge);
                    const uint32_t run = bits_range(p - sgs, q - sgs), hit = old & run;
                    if (hit) {
                        nb |= bits_range(p - sgs, (uint32_t)__builtin_ctz(hit));
                        S.sm.ps.bmp[sg] = old | nb; done = true; break;
                    }
                    nb |= run;
                    p = q;
                    if (p >= sge) break;
                    if ((old >> (p - sgs)) & 1u) { S.sm.ps.bmp[sg] = old | nb; done = true; break; }
                    nb |= 1u << (p - sgs);
                    const Tok t = seg_tok(S, cxm, sgs, p, a, n, sge);
                    if (t.kind == T_LIT) { p++; continue; }
                    if (t.len == LEN_CROSS) {
                        const Tok t2 = eval_tok<false>(S, p, a, n, sge + XROOM);
                        if (t2.len == LEN_CROSS) { S.sm.ps.xpos[sg] = p | XCROSS; S.sm.ps.bmp[sg] = nb; done = true; break; }
                        S.sm.ps.xlen[sg] = t2.len | (t2.kind << 30);
                        p += t2.len;
                        break;
                    }
                    p += t.len;
                }
                if (!done) { S.sm.ps.xpos[sg] = p; S.sm.ps.bmp[sg] = nb; }
            }
            __syncthreads();
            /* the first problem segment, if consistent now, may still end in an unresolved token */
            if (w == 0) {
                const uint32_t eb = S.sm.ps.ent[bad];
                const uint32_t xb = S.sm.ps.xpos[bad];
                const bool okb = (S.sm.ps.bmp[bad] >> (eb - (a + bad * SEG))) & 1u;
                if (okb && (xb & XCROSS)) {
                    const uint32_t tp = xb & ~XCROSS;
                    const Tok t = eval_tok<true>(S, tp, a, n, NONE32);
                    if (lane == 0) {
                        S.sm.ps.xlen[bad] = t.len | (t.kind << 30);
                        S.sm.ps.xpos[bad] = tp + t.len;
                    }
                }
                if (lane == 0) S.st[ST_E0] = eb;
            }
            __syncthreads();
        }
    }
    __syncthreads();

    QTS(1);
    /* 3. walk the true path of every segment once, caching its tokens */
    const uint32_t t0 = S.sm.ps.ent[s];
    const bool live = segv && t0 != NONE32;
    uint32_t nl = 0, nr = 0, nm = 0, last_end = 0, nt = 0;
    if (live) {
        uint32_t p = t0;
        while (p < se) {
            const uint32_t q = lit_run_end(litm, ss, p, se);
            nl += q - p;
            p = q;
            if (p >= se) break;
            Tok t = seg_tok(S, cxm, ss, p, a, n, se);
            if (t.kind == T_LIT) { nl++; p++; continue; }
            if (t.len == LEN_CROSS) t.len = S.sm.ps.xlen[s] & 0x3FFFFFFFu;
            if (t.kind == T_RLE) nr++;
            else nm++;
            const uint32_t k = nt ^ (s & 7u);
            S.big.q.tlen[s][k] = (uint16_t)t.len;
            S.big.q.toff[s][k] = (uint8_t)((p - ss) | (t.kind == T_RLE ? 0x20u : 0u));
            nt++;
            p += t.len;
            last_end = p;
        }
    }
    if (segv) S.sm.ps.ntk[s] = (uint8_t)nt;
    uint32_t sv3[2] = {nl + nr, nm}, st3[2], mv3[1] = {last_end};
    wg_scan_multi<2, 1>(S, sv3, st3, mv3);
    const uint32_t tot_l = st3[0], tot_m = st3[1];
    const uint32_t lit_base = sv3[0] + S.st[ST_LIT];
    const uint32_t m_base = sv3[1];
    uint32_t ls_in = mv3[0];
    const uint32_t carried_ls = S.st[ST_LITSTART];
    if (ls_in < carried_ls) ls_in = carried_ls;
    QTS(2);

    /* 4. match distances in order */
    if (live) {
        uint32_t mi = m_base;
        for (uint32_t k = 0; k < nt; k++) {
            const uint32_t sk = k ^ (s & 7u);
            const uint32_t o = S.big.q.toff[s][sk];
            if (!(o & 0x20u)) S.sm.ps.tdist[mi++] = (uint16_t)dget(S, ss + (o & 31u) - a);
        }
    }
    __syncthreads();
    QTS(3);

    /* 5. byte counts per segment, with the segment's tokens and match distances in registers
     *    and the ring of the last three distances kept up to date from the segment's start */
    uint32_t ccmd = 0, coff = 0, cop = 0, cmeta = 0, cvar = 0, kkp = 0;
    SegTok tk = {};
    if (live) {
        tk = seg_tok_load(S, s, m_base);
        uint32_t r0, r1, r2;
        ring_before(S, m_base, r0, r1, r2);
        uint32_t mr = 0, ls = ls_in;
        for (uint32_t k = 0; k < nt; k++) {
            const uint32_t sk = k ^ (s & 7u);
            const uint32_t o = (uint32_t)(tk.off >> (8u * sk)) & 0xFFu, len = u16_of(tk.len, sk);
            const uint32_t p = ss + (o & 31u);
            const uint32_t ll = p - ls;
            if (o & 0x20u) {
                if (ll) { ccmd += 1; cop += 1; if (ll > 8) { const uint32_t q = leb_len(ll); ccmd += q; cvar += q; } }
                const uint32_t q = leb_len(len);
                ccmd += 1 + q; cop += 1; cvar += q;
            } else {
                const uint32_t d = u16_of(tk.dist, mr++);
                const int kk = ring_use(d, r0, r1, r2);
                kkp |= (uint32_t)(kk + 1) << (2u * k); /* kept for the write pass */
                const uint32_t keep = kk >= 0 ? (len <= 7 ? 7u : 15u) : 3u;
                if (ll > keep) {
                    const uint32_t f = ll - keep;
                    ccmd += 1; cop += 1;
                    if (f > 8) { const uint32_t q = leb_len(f); ccmd += q; cvar += q; }
                }
                const uint32_t xq = len - 4 > 14 ? leb_len(len - 19) : 0u;
                if (kk >= 0 && len <= 7) { ccmd += 1; cop += 1; }
                else if (kk >= 0) { ccmd += 2 + xq; cop += 1; cmeta += 1; cvar += xq; }
                else { ccmd += 1 + xq; cop += 1; cvar += xq; coff += leb_len(d); }
            }
            ls = p + len;
        }
    }
    uint32_t sv5[5] = {ccmd, coff, cop, cmeta, cvar}, st5[5];
    wg_scan_multi<5, 0>(S, sv5, st5, nullptr);
    const uint32_t t_cmd = st5[0], t_off = st5[1], t_op = st5[2], t_meta = st5[3], t_var = st5[4];
    uint32_t b_cmd = sv5[0] + S.st[ST_CMD];
    uint32_t b_off = sv5[1] + S.st[ST_OFF];
    uint32_t b_op = sv5[2] + S.st[ST_OP];
    uint32_t b_meta = sv5[3] + S.st[ST_META];
    uint32_t b_var = sv5[4] + S.st[ST_VAR];
    QTS(4);

    /* 6. write bytes: literal gaps, then each token.  dist and lc are dead after step 4, so
     *    when the quarter's literal, cmd and op bytes fit there they are staged in LDS and
     *    copied out with aligned 16-byte stores (scattered byte stores otherwise); the small
     *    meta, var and off streams are written directly */
    const uint32_t lit0 = S.st[ST_LIT];
    const uint32_t o_cmd = stg_room(tot_l), o_op = o_cmd + stg_room(t_cmd);
    const bool stage = o_op + stg_room(t_op) <= STG_LIMIT; /* uniform */
    uint8_t *const LB = (uint8_t *)S.big.q.dist;
    if (live) {
        if (stage)
            write_tokens(S, s, ss, se, t0, nt, tk, kkp, ls_in, LB, lit_base - lit0, LB + o_cmd, sv5[0], LB + o_op,
                         sv5[2], meta, b_meta, var, b_var, off, b_off);
        else
            write_tokens(S, s, ss, se, t0, nt, tk, kkp, ls_in, lit, lit_base, cmd, b_cmd, op, b_op, meta, b_meta,
                         var, b_var, off, b_off);
    }
    __syncthreads();
    QTS(7);
    if (stage) {
        lds_copy_out(lit + lit0, LB, tot_l);
        lds_copy_out(cmd + (b_cmd - sv5[0]), LB + o_cmd, t_cmd);
        lds_copy_out(op + (b_op - sv5[2]), LB + o_op, t_op);
    }
    __syncthreads();
    QTS(5);

    /* 7. carry the uniform state into the next tile */
    if (tid == 0) {
        uint32_t out[3];
        int no = 0;
        for (uint32_t j = tot_m; j-- > 0 && no < 3;) {
            const uint32_t v = S.sm.ps.tdist[j];
            bool dup = false;
            for (int q = 0; q < no; q++) dup |= out[q] == v;
            if (!dup) out[no++] = v;
        }
        const uint32_t rr[3] = {S.st[ST_R0], S.st[ST_R1], S.st[ST_R2]};
        for (int j = 0; j < 3 && no < 3; j++) {
            bool dup = false;
            for (int q = 0; q < no; q++) dup |= out[q] == rr[j];
            if (!dup) out[no++] = rr[j];
        }
        S.st[ST_R0] = out[0]; S.st[ST_R1] = out[1]; S.st[ST_R2] = out[2];
        S.st[ST_CMD] += t_cmd; S.st[ST_OFF] += t_off; S.st[ST_OP] += t_op;
        S.st[ST_META] += t_meta; S.st[ST_VAR] += t_var; S.st[ST_LIT] += tot_l;
        S.st[ST_NMATCH] = tot_m;
    }
    {
        const uint32_t v = live ? last_end : 0u;
        uint32_t sv7[1] = {nr + nm}, st7[1], mv7[1] = {v};
        wg_scan_multi<1, 1>(S, sv7, st7, mv7);
        const uint32_t incl = mv7[0] > v ? mv7[0] : v;
        if (tid == LZ_NT - 1) {
            const uint32_t c = S.st[ST_LITSTART];
            S.st[ST_LITSTART] = incl > c ? incl : c;
        }
        if (tid == 0) S.st[ST_TOK] += st7[0];
    }
    __syncthreads();
    QTS(6);
    (void)lane;
}

/* ---------------- kernel ---------------- */

/*
 * The kernel body.  MATCH_OUT (the match-stage entry point only) also writes every position's
 * match distance to dist_out; phase_ts (builds with PRZM_DIAG only) receives s_memrealtime
 * stamps at the phase edges of every block.
 */
template <bool MATCH_OUT, uint32_t B>
__device__ __forceinline__ void lz_body(LzShared &S, const uint8_t *__restrict__ src,
                                        const przm_block_t *__restrict__ blocks, uint32_t nblocks, uint32_t depth,
                                        uint8_t *__restrict__ ws, przm_meta_t *__restrict__ meta,
                                        uint8_t *__restrict__ scratch, uint64_t scratch_stride,
                                        uint16_t *__restrict__ dist_out, uint64_t *__restrict__ phase_ts,
                                        uint32_t *__restrict__ next_block) {
    const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    uint8_t *my = scratch + (uint64_t)blockIdx.x * scratch_stride;
    Scr sc;
    sc.T0 = (const uint16_t *)my;
    sc.T1 = sc.T0 + PRZM_HALF;
    sc.D = (uint32_t *)(my + 2u * PRZM_HALF * 2u);
    sc.gws = (uint8_t *)(my + 2u * PRZM_HALF * 2u + 2u * PRZM_HALF * 4u + 2u * PRZM_HALF * 2u);
    uint16_t *T0w = (uint16_t *)my, *T1w = T0w + PRZM_HALF;

    /* blocks are handed out dynamically (next_block starts at 0 for each launch),
     * so a workgroup that drew cheap blocks takes more of them */
    for (;;) {
        if (tid == 0) S.nextb = atomicAdd(next_block, 1u);
        __syncthreads();
        const uint32_t blk = S.nextb;
        __syncthreads();
        if (blk >= nblocks) break;
        const przm_block_t bd = blocks[blk];
        const uint32_t n = bd.len;
        przm_meta_t *mt = meta + blk;
        uint8_t *wsb = ws + (uint64_t)blk * PRZM_WS_BYTES;
        uint64_t *ts = phase_ts ? phase_ts + (uint64_t)blk * 32 : nullptr;
#ifdef PRZM_DIAG
#define PTS(i) do { if (ts && tid == 0) ts[i] = __builtin_amdgcn_s_memrealtime(); } while (0)
#else
#define PTS(i) ((void)0)
#endif
        PTS(0);
        if ((bd.flags & PRZM_BF_RAW) || n == 0) {
            if (tid < PRZM_S_N) mt->len[tid] = 0;
            continue;
        }
        /* load the block into LDS.  A full aligned block (the common case): the eight rounds'
         * 16-byte loads are all issued before the first store, so the load waits for one memory
         * latency; the 64 bytes past the block are zeroed.  Any other block: 16 bytes a round,
         * byte by byte past its end. */
        const uint8_t *g = src + bd.src_off;
        if ((bd.src_off & 15u) == 0 && n == PRZM_BLOCK) {
            constexpr uint32_t LR = PRZM_BLOCK / (LZ_NT * 16);
            const uint4 *g4 = (const uint4 *)g + tid;
            uint4 lv[LR];
#pragma unroll
            for (uint32_t r = 0; r < LR; r++) lv[r] = g4[r * LZ_NT];
#pragma unroll
            for (uint32_t r = 0; r < LR; r++) *(uint4 *)(S.in + 16u * (tid + r * LZ_NT)) = lv[r];
            if (tid < 4) *(uint4 *)(S.in + PRZM_BLOCK + 16u * tid) = make_uint4(0, 0, 0, 0);
        } else {
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
        }
        if (tid < ST_N) S.st[tid] = 0;
        if (tid < LZ_NW) S.bigq[tid] = 0;
        for (uint32_t i = tid; i < LZ_NW * 256; i += LZ_NT) (&S.sm.pbits[0][0])[i] = 0ull;
        __syncthreads();
        if (tid == 0) { S.st[ST_R0] = 1; S.st[ST_R1] = 4; S.st[ST_R2] = 8; }

        const uint32_t nh = n >= 4 ? n - 3 : 0; /* positions p with p + 4 <= n */
        PTS(1);

        /* A. bucket positions by the hash's high byte, per half */
        for (uint32_t H = 0; H < 2; H++) {
            const uint32_t base = H * PRZM_HALF;
            const uint32_t end = nh < base + PRZM_HALF ? (nh > base ? nh : base) : base + PRZM_HALF;
            for (uint32_t i = tid; i < LZ_NW * 256; i += LZ_NT) (&S.sm.tab[0][0])[i] = 0;
            __syncthreads();
            { /* the half's match results start as "no match" (coalesced): the search writes matches only */
                uint4 *d4 = (uint4 *)(sc.D + base);
                for (uint32_t i = tid; i < PRZM_HALF / 4u; i += LZ_NT) d4[i] = make_uint4(0u, 0u, 0u, 0u);
            }
            const uint32_t wb = base + w * (PRZM_HALF / LZ_NW);
            /* eight digits at a time, their loads unconditional (positions below PRZM_BLOCK, whose
             * words lie inside in[]) so that they are in flight together */
            for (uint32_t c = 0; c < PRZM_HALF / LZ_NW; c += 8 * 64) {
                uint32_t dgs[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; u++) dgs[u] = hash4(S.in, wb + c + u * 64 + lane) >> 8;
#pragma unroll
                for (uint32_t u = 0; u < 8; u++)
                    if (wb + c + u * 64 + lane < end) atomicAdd(&S.sm.tab[w][dgs[u]], 1u);
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
            /* the digit of the next step is loaded one step ahead */
            uint32_t dnext = hash4(S.in, wb + lane) >> 8;
            for (uint32_t c = 0; c < PRZM_HALF / LZ_NW; c += 64) {
                const uint32_t p = wb + c + lane;
                const bool valid = p < end;
                const uint32_t dg = valid ? dnext : 0u;
                dnext = hash4(S.in, p + 64 < PRZM_BLOCK ? p + 64 : PRZM_BLOCK) >> 8;
                uint32_t cur;
                const uint64_t pm = peers_lds(S.sm.pbits[w], S.sm.tab[w], dg, valid, &cur);
                if (valid) {
                    S.big.stage[cur + (uint32_t)__popcll(pm & lt)] = (uint16_t)p;
                    if ((pm & lt) == 0) S.sm.tab[w][dg] = cur + (uint32_t)__popcll(pm);
                }
                lds_fence();
            }
            __syncthreads();
            const uint32_t cnt = S.bstart[H][256];
            uint16_t *T = H ? T1w : T0w;
            /* eight positions per 16-byte store (both arrays are 16-byte aligned), the tail one by one */
            const uint32_t c8 = cnt & ~7u;
            for (uint32_t i = tid * 8u; i < c8; i += LZ_NT * 8u) *(uint4 *)(T + i) = *(const uint4 *)(S.big.stage + i);
            for (uint32_t i = c8 + tid; i < cnt; i += LZ_NT) T[i] = S.big.stage[i];
            __syncthreads();
        }
        PTS(2);

        /* C. per bucket: group by the low byte, search; large buckets are searched by all waves */
        if (tid == 0) S.st[ST_BUCKET] = 0;
        __threadfence_block();
        __syncthreads();
        Pend pd = {0u, NONE32};
        {
            /* the next bucket is taken, and its first positions requested, before the current
             * one is worked (the chunk results held back in pd keep that wait short) */
            uint32_t dg = 0;
            if (lane == 0) dg = atomicAdd(&S.st[ST_BUCKET], 1u);
            dg = (uint32_t)__builtin_amdgcn_readlane((int)dg, 0);
            uint32_t pre[4];
            bucket_fetch(S, sc, dg, pre);
            while (dg < 256) {
                uint32_t dn = 0;
                if (lane == 0) dn = atomicAdd(&S.st[ST_BUCKET], 1u);
                dn = (uint32_t)__builtin_amdgcn_readlane((int)dn, 0);
                uint32_t nxt[4];
                bucket_fetch(S, sc, dn, nxt);
                const uint32_t m = (S.bstart[0][dg + 1] - S.bstart[0][dg]) + (S.bstart[1][dg + 1] - S.bstart[1][dg]);
                if (m != 0 && m <= WS_CAP) bucket_sort<true, B>(S, sc, w, dg, n, depth, false, pre, pd, ts ? ts + 27 : nullptr);
                else if (m != 0) {
                    const bool q = bucket_sort<false, B>(S, sc, w, dg, n, depth, S.bigq[w] == 0, pre, pd,
                                                      ts ? ts + 27 : nullptr);
                    if (q && lane == 0) S.bigq[w] = dg + 1;
                }
                dg = dn;
#pragma unroll
                for (int u = 0; u < 4; u++) pre[u] = nxt[u];
            }
        }
        if (pd.a != NONE32) sc.D[pd.a] = pd.v;
        if (ts && lane == 0 && (w == 0 || w == LZ_NW - 1)) ts[w == 0 ? 14 : 15] = __builtin_amdgcn_s_memrealtime();
        __threadfence_block();
        __syncthreads();
        PTS(13);
        for (uint32_t ow = 0; ow < LZ_NW; ow++) {
            const uint32_t dq = S.bigq[ow];
            if (!dq) continue;
            const uint32_t dg = dq - 1;
            const uint32_t m = (S.bstart[0][dg + 1] - S.bstart[0][dg]) + (S.bstart[1][dg + 1] - S.bstart[1][dg]);
            const uint8_t *gw = sc.gws + (uint64_t)ow * GWS_STRIDE;
            const uint16_t *SRT = (const uint16_t *)gw + PRZM_BLOCK;
            const uint32_t *GE = (const uint32_t *)(gw + 2u * PRZM_BLOCK * 2u);
            Pend pq = {0u, NONE32};
            if (depth > 64u) bucket_search_range<true, B>(S, sc, dg, n, depth, SRT, GE, w * 64, m, LZ_NW * 64, pq);
            else if (depth < B) bucket_search_range<false, 2>(S, sc, dg, n, depth, SRT, GE, w * 64, m, LZ_NW * 64, pq);
            else bucket_search_range<false, B>(S, sc, dg, n, depth, SRT, GE, w * 64, m, LZ_NW * 64, pq);
            if (pq.a != NONE32) sc.D[pq.a] = pq.v;
        }
        __threadfence_block();
        __syncthreads();
        PTS(3);

        /* E + parse, per quarter tile */
        for (uint32_t Q = 0; Q < 4; Q++) {
            const uint32_t a = Q * QP;
            if (a >= n) break;
            const uint32_t b = n < a + QP ? n : a + QP;
            /* the tile's match results [a, b] from the position-ordered results (positions
             * past n - 4 have none) */
            /* four positions per 16-byte load: thread t of round r holds positions 4t + 2048r ..
             * 4t + 2048r + 3 (a wave reads 1 KiB contiguous); the nine rounds go in two groups
             * whose loads are all issued before any is used, unconditionally (clamped into the
             * array; what lies past nh is masked), so the fill waits for two memory latencies */
            constexpr uint32_t FR4 = (QP + 1 + 4 * LZ_NT - 1) / (4 * LZ_NT), FB4 = 5;
#pragma unroll 1
            for (uint32_t r0 = 0; r0 < FR4; r0 += FB4) {
            uint4 dv[FB4];
#pragma unroll
            for (uint32_t u = 0; u < FB4; u++) {
                const uint32_t e = a + 4u * tid + (r0 + u) * 4u * LZ_NT;
                dv[u] = *(const uint4 *)(sc.D + (e < PRZM_BLOCK - 4u ? e : PRZM_BLOCK - 4u));
            }
#pragma unroll
            for (uint32_t u = 0; u < FB4; u++) {
                if (r0 + u >= FR4) break;
                const uint32_t i = 4u * tid + (r0 + u) * 4u * LZ_NT;
                const uint32_t vv[4] = {dv[u].x, dv[u].y, dv[u].z, dv[u].w};
                uint32_t nib = 0;
#pragma unroll
                for (uint32_t j = 0; j < 4; j++) {
                    const bool act = i + j <= b - a;
                    const uint32_t v = act && a + i + j < nh ? vv[j] : 0u;
                    if (act) {
                        S.big.q.dist[DIX(i + j)] = (uint16_t)v;
                        S.big.q.lc[LIX(i + j)] = (uint8_t)(v >> 16);
                    }
                    nib |= ((v & 0xFFFFu) != 0u ? 1u : 0u) << j;
                }
                /* eight lanes cover one 32-position segment: their nibbles make its match mask */
                uint32_t hv = nib << (4u * (lane & 7u));
                hv |= __shfl_xor(hv, 1);
                hv |= __shfl_xor(hv, 2);
                hv |= __shfl_xor(hv, 4);
                if ((lane & 7u) == 0 && i <= b - a) S.sm.ps.hm[i >> 5] = hv;
            }
            }
            __syncthreads();
            if (MATCH_OUT)
                for (uint32_t i = tid; i < b - a; i += LZ_NT)
                    dist_out[(uint64_t)blk * PRZM_BLOCK + a + i] = S.big.q.dist[DIX(i)];
            PTS(4 + 2 * Q);
            if (S.st[ST_ENTRY] < b) parse_tile(S, a, b, n, wsb, (ts && Q == 0) ? ts + 16 : nullptr);
            __syncthreads();
            PTS(5 + 2 * Q);
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
        PTS(12);
        __syncthreads();
    }
}

/* the encode pipeline's match finder and tokenizer, B chain candidates per search step */
#define PRZM_LZ_KERNEL(NAME, B)                                                                                       \
    extern "C" __global__ void __launch_bounds__(LZ_NT)                                                              \
    NAME(const uint8_t *__restrict__ src, const przm_block_t *__restrict__ blocks, uint32_t nblocks, uint32_t depth, \
         uint8_t *__restrict__ ws, przm_meta_t *__restrict__ meta, uint8_t *__restrict__ scratch,                     \
         uint64_t scratch_stride, uint32_t *__restrict__ next_block) {                                               \
        __shared__ LzShared S;                                                                                       \
        lz_body<false, B>(S, src, blocks, nblocks, depth, ws, meta, scratch, scratch_stride, nullptr, nullptr,       \
                          next_block);                                                                               \
    }
PRZM_LZ_KERNEL(przm_lz_kernel, 4)
PRZM_LZ_KERNEL(przm_lz_kernel_b2, 2)
PRZM_LZ_KERNEL(przm_lz_kernel_b8, 8)

/* the same, also writing every position's match distance (przm_gpu_match's descriptors) */
extern "C" __global__ void __launch_bounds__(LZ_NT)
przm_lz_match_kernel(const uint8_t *__restrict__ src, const przm_block_t *__restrict__ blocks, uint32_t nblocks,
                     uint32_t depth, uint8_t *__restrict__ ws, przm_meta_t *__restrict__ meta,
                     uint8_t *__restrict__ scratch, uint64_t scratch_stride, uint16_t *__restrict__ dist_out,
                     uint32_t *__restrict__ next_block) {
    __shared__ LzShared S;
    lz_body<true, 4>(S, src, blocks, nblocks, depth, ws, meta, scratch, scratch_stride, dist_out, nullptr, next_block);
}

#ifdef PRZM_DIAG
/* diagnostics builds: the encode kernel with 32 phase stamps per block */
extern "C" __global__ void __launch_bounds__(LZ_NT)
przm_lz_diag_kernel(const uint8_t *__restrict__ src, const przm_block_t *__restrict__ blocks, uint32_t nblocks,
                    uint32_t depth, uint8_t *__restrict__ ws, przm_meta_t *__restrict__ meta,
                    uint8_t *__restrict__ scratch, uint64_t scratch_stride, uint64_t *__restrict__ phase_ts,
                    uint32_t *__restrict__ next_block) {
    __shared__ LzShared S;
    lz_body<false, 8>(S, src, blocks, nblocks, depth, ws, meta, scratch, scratch_stride, nullptr, phase_ts, next_block);
}
#endif
