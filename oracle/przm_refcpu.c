/*
 * przm_refcpu.c -- scalar C restatement of the reference Go CPU compressor
 * (compressInternal at a level: compress.go:1036-1114), the reference CPU path
 * that the GPU encode path stands beside.
 *
 * TEST AND BASELINE INFRASTRUCTURE ONLY, like the rest of oracle/: the test
 * suite and the CPU-baseline leg of the benchmark load it; the product never
 * does.
 *
 * What it restates, per frame (one call = one compressInternal call):
 *   AnalyzeInput                compress.go:93-202 (entropy guard, stride, bMin, window log)
 *   the trellis                 compress.go:890-946 (solveBlockTrellis over 64 KiB horizons,
 *                               backtrackEdges), with every relaxation of compress.go:390-888:
 *                               literal, the four repeat registers, brackets, stride columns,
 *                               the hash-4 chain and the hash-8/12/16/24/32 tiers by level,
 *                               the static dictionary (core/dict/static_index.go:35-311, when
 *                               its tables are loaded), the difference lattice
 *   cost model                  compress.go:243-372 (float surrogates as Go computes them:
 *                               math.Log2 as frexp + log, truncating conversions)
 *   command emission            core/commands/commands.go:580-752 (EmitBlockCommands)
 *   Section 2                   SplitCommands with every opcode family (commands.go:18-202),
 *                               then PackCommandsMode8 / flat (orc_package_commands)
 *   Section 3                   PackageOffsetsSection (orc_package_offsets)
 *   Section 4                   PackageLiteralsSection with the sheaf trials
 *                               (orc_package_literals)
 *   frame                       assembleContainer / assembleCompressedFrame
 *                               (compress.go:962-1033), stride byte and flags
 *
 * Two rules are added on top of the reference, the same two as in the GPU mode's frames, so
 * that every frame decodes: a raw Section 4 that the decoders would read as a container makes
 * the frame uncompressed, and a distance-prefix container whose code stream is exactly as long
 * as its match count is written flat (orc_package_offsets).
 *
 * Parity with the Go bytes is unpinned (Go is absent here); the frames are checked by the
 * reference C99 decoder's round trip and by the invariants in tests/test_refcpu.py.
 */
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../przm_amd/csrc/przm_dist_freq.h"

/* the parts of the oracle this file builds on (przm_oracle.c, same library) */
typedef struct {
    uint8_t *cmd; uint32_t ncmd;
    uint8_t *off; uint32_t noff;
    uint8_t *lit; uint32_t nlit;
    uint8_t *op;  uint32_t nop;
    uint8_t *meta; uint32_t nmeta;
    uint8_t *var; uint32_t nvar;
    uint32_t ntok;
} orc_streams;
uint32_t orc_package_commands(const orc_streams *s, uint8_t *out, int choices[3]);
uint32_t orc_package_offsets(const uint8_t *off, uint32_t noff, uint8_t *out, int *mode, int *code_kind);
uint32_t orc_package_literals(const uint8_t *lits, uint32_t n, uint8_t *out, int *kind);
int orc_check_high_entropy(const uint8_t *src, uint32_t n);
uint32_t orc_window_log(uint32_t u);
void orc_dist_code(uint32_t d, uint8_t *code, uint32_t *extra, uint8_t *nb);

#define RC_HORIZON 65536u    /* MaxBlockHorizon */
#define RC_MAXM 1024         /* MaxMatchLength */
#define RC_INF 0x3FFFFFFFu   /* InfinityCost */
#define RC_M32 0x1E35A7BDu
#define RC_M64 0x9E3779B97F4A7C15ull
#define RC_C1 0xBF58476D1CE4E5B9ull
#define RC_C2 0x94D049BB133111EBull
#define RC_C3 0x3C79AC492BA7B653ull

static inline uint32_t rc_rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static inline uint64_t rc_rd64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
static inline int rc_min(int a, int b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ */
/* math.Log2 as Go computes it (frexp, then log of the fraction)       */

static double go_log2(double x) {
    int e;
    const double frac = frexp(x, &e);
    if (frac == 0.5) return (double)(e - 1);
    return log(frac) * 1.44269504088896340735992468100189214 + (double)e;
}

/* compress.go:243-258: the distance code costs, from the static Mode-12 tables */
static uint32_t rc_dist_cost[64];
static pthread_once_t rc_once = PTHREAD_ONCE_INIT;
static void rc_init_costs(void) {
    for (int c = 0; c < 64; c++) {
        int sum = 0;
        for (int ctx = 0; ctx < 4; ctx++) sum += przm_dist_freq[ctx][c];
        double avg = (double)sum / 4.0;
        if (avg <= 0) avg = 1.0;
        const double bits = go_log2(4096.0 / avg);
        const uint32_t cost = (uint32_t)(bits * 2.0 + 0.5);
        rc_dist_cost[c] = cost > 6 ? cost : 6;
    }
}

/* compress.go:291-302 */
static uint32_t calc_lit_cost(int f, double inv_n) {
    if (f == 0) return 36;
    const double bits = -go_log2((double)f * inv_n);
    const uint32_t cost = (uint32_t)(bits * 2.6 + 1.2);
    return cost < 3 ? 3 : cost;
}

/* compress.go:315-324 */
static uint32_t lit_flush_cost(int n) { return n == 0 ? 0 : n <= 8 ? 10 : 22; }

/* compress.go:326-341 */
static uint32_t cost_repeat(int L, int k, int bmin, int nlit) {
    uint32_t base = 36;
    if (L <= 3 + bmin && nlit <= 7) base = 14;
    else if (L <= 18 + bmin && nlit <= 15) base = 28;
    else if (nlit > 15) base += lit_flush_cost(nlit);
    const uint32_t bonus = (uint32_t)(4 - k);
    return base > bonus + 3 ? base - bonus : 3;
}

/* compress.go:343-372 */
static uint32_t cost_explicit(int L, uint32_t dist, int bmin, int nlit) {
    uint8_t code, nb;
    uint32_t extra;
    orc_dist_code(dist, &code, &extra, &nb);
    uint32_t base = 9;
    if (nlit >= 1 && nlit <= 3) base = 12;
    else if (nlit >= 4 && nlit <= 15) base = 16;
    else if (nlit > 15) base = 9 + lit_flush_cost(nlit);
    if (L > 14 + bmin) base += L <= 14 + bmin + 127 ? 3 : 7;
    uint32_t dc = rc_dist_cost[code] + 2u * nb;
    if (L == bmin) { if (dist > 512) dc += 8; }
    else if (L == bmin + 1) { if (dist > 4096) dc += 6; }
    return base + dc;
}

/* compress.go:260-289 */
static int pareto_sublengths(int L, int bmin, int *out) {
    int n = 0;
    out[n++] = L;
    for (int delta = 1; delta <= 4; delta++)
        if (L - delta >= bmin) out[n++] = L - delta;
    const int cb = 14 + bmin;
    if (cb >= bmin && cb < L - 4) out[n++] = cb;
    static const int steps[4] = {32, 24, 16, 12};
    for (int s = 0; s < 4; s++)
        if (steps[s] < L - 4 && steps[s] > 8 && steps[s] != cb) out[n++] = steps[s];
    for (int sub = 8; sub >= bmin; sub--)
        if (sub < L - 4 && sub != cb) out[n++] = sub;
    return n;
}

/* compress.go:374-387: eight bytes at a time, then the tail byte by byte */
static int match_len(const uint8_t *src, int p1, int p2, int max) {
    int l = 0;
    while (l + 8 <= max) {
        const uint64_t x = rc_rd64(src + p1 + l) ^ rc_rd64(src + p2 + l);
        if (x) return l + (__builtin_ctzll(x) >> 3);
        l += 8;
    }
    while (l < max && src[p1 + l] == src[p2 + l]) l++;
    return l;
}

/* core/distance/distance.go:479-504, 861-887 */
static void lru_repeat(uint32_t r[4], int k) {
    const uint32_t d = r[k];
    for (int j = k; j > 0; j--) r[j] = r[j - 1];
    r[0] = d;
}
static void lru_explicit(uint32_t r[4], uint32_t d) { r[3] = r[2]; r[2] = r[1]; r[1] = r[0]; r[0] = d; }
static uint32_t bracket_distance(int b) { return b <= 7 ? (uint32_t)(b + 1) : b == 8 ? 16u : 32u; }

/* compress.go:204-241 */
static uint16_t hash4(const uint8_t *s, int p) { return (uint16_t)((rc_rd32(s + p) * RC_M32) >> 16); }
static uint16_t hash8(const uint8_t *s, int p) { return (uint16_t)((rc_rd64(s + p) * RC_M64) >> 48); }
static uint16_t hash12(const uint8_t *s, int p) {
    return (uint16_t)(((rc_rd64(s + p) ^ (rc_rd64(s + p + 4) * RC_C1)) * RC_M64) >> 48);
}
static uint16_t hash16(const uint8_t *s, int p) {
    return (uint16_t)(((rc_rd64(s + p) ^ (rc_rd64(s + p + 8) * RC_C1)) * RC_M64) >> 48);
}
static uint16_t hash24(const uint8_t *s, int p) {
    const uint64_t h = rc_rd64(s + p) ^ (rc_rd64(s + p + 8) * RC_C1) ^ (rc_rd64(s + p + 16) * RC_C2);
    return (uint16_t)((h * RC_M64) >> 48);
}
static uint16_t hash32(const uint8_t *s, int p) {
    const uint64_t h = rc_rd64(s + p) ^ (rc_rd64(s + p + 8) * RC_C1) ^ (rc_rd64(s + p + 16) * RC_C2) ^
                       (rc_rd64(s + p + 24) * RC_C3);
    return (uint16_t)((h * RC_M64) >> 48);
}

/* ------------------------------------------------------------------ */
/* The static dictionary (core/dict/static_index.go, dict.go), built    */
/* from the reference decoder's own tables when they are loaded.        */

typedef struct { const char *str; uint8_t len; } rc_str;
static struct {
    int loaded;
    const uint16_t *word_base;        /* [22] */
    const uint32_t *dict_offsets;     /* [21] */
    const uint8_t (*transforms)[3];   /* [121] {prefix, op, suffix} */
    const uint16_t (*compact)[2];     /* [2048] {wid, tau} */
    const rc_str *ps;                 /* [50] */
    const rc_str *top;                /* [512] */
    const uint8_t *data;              /* 122784 */
} rc_tab;

typedef struct { uint32_t prefix; uint16_t p; uint8_t len; int32_t next; const uint8_t *word; } rc_entry;
static int32_t rc_head[65536];
static rc_entry *rc_ent;
static uint32_t rc_nent;
static uint8_t *rc_compact_words; /* transformed compact pairs, back to back */

/* LookupBrotliWord (core/internal/tables LookupBrotliWord): the word of id wid < 13504 */
static const uint8_t *brotli_word(uint32_t wid, int *len) {
    int lo = 0, hi = 21, best = 0;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        if (rc_tab.word_base[mid] <= wid) { best = mid; lo = mid + 1; }
        else hi = mid - 1;
    }
    *len = best + 4;
    return rc_tab.data + rc_tab.dict_offsets[best] + (size_t)(wid - rc_tab.word_base[best]) * (size_t)*len;
}

/* resolveWordByWid (dict.go:72-81) */
static const uint8_t *word_by_wid(uint32_t wid, int *len) {
    if (wid < 13504) return brotli_word(wid, len);
    const uint32_t i = wid - 13504;
    if (i >= 512) { *len = 0; return NULL; }
    *len = rc_tab.top[i].len;
    return (const uint8_t *)rc_tab.top[i].str;
}

/* insertStaticEntry (static_index.go:83-108): newest first in its bucket, duplicates dropped */
static void insert_static(const uint8_t *w, int len, uint16_t p) {
    const uint32_t prefix = rc_rd32(w), h = (prefix * RC_M32) >> 16;
    for (int32_t i = rc_head[h]; i >= 0; i = rc_ent[i].next)
        if (rc_ent[i].prefix == prefix && rc_ent[i].len == len && !memcmp(rc_ent[i].word, w, (size_t)len)) return;
    rc_ent[rc_nent] = (rc_entry){prefix, p, (uint8_t)len, rc_head[h], w};
    rc_head[h] = (int32_t)rc_nent++;
}

/* transformWord + ApplyTransform (dict.go:149-207) into dst (256 bytes); -1 on error */
static int apply_transform(uint8_t *dst, const uint8_t *w, int wl, uint32_t tau) {
    if (tau >= 121) return -1;
    const uint8_t *t = rc_tab.transforms[tau];
    const rc_str pre = rc_tab.ps[t[0]], suf = rc_tab.ps[t[2]];
    const uint8_t op = t[1];
    uint8_t scratch[256];
    const uint8_t *mw = w;
    int ml = wl;
    if (op >= 1 && op <= 9) { ml = (int)op >= wl ? 0 : wl - op; if ((int)op >= wl) mw = NULL; }
    else if (op == 10 || op == 11) {
        memcpy(scratch, w, (size_t)wl);
        for (int i = 0; i < wl; i++) {
            if (scratch[i] >= 'a' && scratch[i] <= 'z') scratch[i] -= 32;
            if (op == 10) break; /* capitalize: the first byte only */
        }
        mw = scratch;
    } else if (op >= 12 && op <= 20) {
        const int sk = op - 11;
        if (sk >= wl) { mw = NULL; ml = 0; }
        else { mw = w + sk; ml = wl - sk; }
    }
    const int n = pre.len + ml + suf.len;
    if (n > 256) return -1;
    memcpy(dst, pre.str, pre.len);
    if (ml) memcpy(dst + pre.len, mw, (size_t)ml);
    memcpy(dst + pre.len + ml, suf.str, suf.len);
    return n;
}

/* initStaticDictIndex (static_index.go:35-81) */
static void build_static_index(void) {
    for (int i = 0; i < 65536; i++) rc_head[i] = -1;
    rc_ent = (rc_entry *)calloc(16064, sizeof(rc_entry));
    rc_compact_words = (uint8_t *)malloc(2048 * 256);
    rc_nent = 0;
    for (uint32_t wid = 128; wid < 13504; wid++) {
        int l;
        const uint8_t *w = brotli_word(wid, &l);
        if (l >= 4) insert_static(w, l, (uint16_t)wid);
    }
    for (int i = 0; i < 512; i++)
        if (rc_tab.top[i].len >= 4) insert_static((const uint8_t *)rc_tab.top[i].str, rc_tab.top[i].len, (uint16_t)(13504 + i));
    size_t used = 0;
    for (int c = 0; c < 2048; c++) {
        int wl;
        const uint8_t *w = word_by_wid(rc_tab.compact[c][0], &wl);
        uint8_t tmp[256];
        const int n = apply_transform(tmp, w, wl, rc_tab.compact[c][1]);
        if (n < 4) continue;
        memcpy(rc_compact_words + used, tmp, (size_t)n);
        insert_static(rc_compact_words + used, n, (uint16_t)(14016 + c));
        used += (size_t)n;
    }
}

static pthread_mutex_t rc_dict_mu = PTHREAD_MUTEX_INITIALIZER;

/* Loads the static dictionary tables from a library built from the reference decoder's own
 * table header (oracle/ref_tables.c -> oracle/_ref/libprzm_ref_tables.so) and indexes them.
 * 0 when loaded (or already), -1 when the library or a symbol is missing: the trellis then
 * runs without the dictionary relaxation. */
int orc_ref_load_dict(const char *path) {
    pthread_mutex_lock(&rc_dict_mu);
    int rc = 0;
    if (!rc_tab.loaded) {
        void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
        const void *(*get)(int) = h ? (const void *(*)(int))dlsym(h, "przm_ref_table") : NULL;
        if (!get) rc = -1;
        else {
            rc_tab.word_base = (const uint16_t *)get(0);
            rc_tab.dict_offsets = (const uint32_t *)get(1);
            rc_tab.transforms = (const uint8_t(*)[3])get(2);
            rc_tab.compact = (const uint16_t(*)[2])get(3);
            rc_tab.ps = (const rc_str *)get(4);
            rc_tab.top = (const rc_str *)get(5);
            rc_tab.data = (const uint8_t *)get(6);
            build_static_index();
            rc_tab.loaded = 1;
        }
    }
    pthread_mutex_unlock(&rc_dict_mu);
    return rc;
}

int orc_ref_dict_entries(void) { return rc_tab.loaded ? (int)rc_nent : -1; }

typedef void (*dict_fn)(void *ctx, uint32_t wid, uint8_t tau, int len);

/* probeSuffixes (static_index.go:170-200) */
static void probe_suffixes(const uint8_t *s, int n, int l, uint32_t wid, dict_fn fn, void *ctx) {
    if (wid >= 13504 || l >= n) return;
    switch (s[l]) {
    case ' ': fn(ctx, wid, 1, l + 1); break;
    case '.':
        fn(ctx, wid, 20, l + 1);
        if (l + 1 < n && s[l + 1] == ' ') fn(ctx, wid, 31, l + 2);
        break;
    case ',':
        fn(ctx, wid, 76, l + 1);
        if (l + 1 < n && s[l + 1] == ' ') fn(ctx, wid, 14, l + 2);
        break;
    case '\n': fn(ctx, wid, 22, l + 1); break;
    case ':': fn(ctx, wid, 51, l + 1); break;
    case '(': fn(ctx, wid, 57, l + 1); break;
    case '"': fn(ctx, wid, 19, l + 1); break;
    case '\'': fn(ctx, wid, 36, l + 1); break;
    case ']': fn(ctx, wid, 24, l + 1); break;
    }
}

/* probeUppercaseFirstSuffixes (static_index.go:202-222) */
static void probe_upper_suffixes(const uint8_t *s, int n, int l, uint32_t wid, dict_fn fn, void *ctx) {
    if (wid >= 13504 || l >= n) return;
    switch (s[l]) {
    case ' ': fn(ctx, wid, 4, l + 1); break;
    case '.':
        fn(ctx, wid, 79, l + 1);
        if (l + 1 < n && s[l + 1] == ' ') fn(ctx, wid, 88, l + 2);
        break;
    case ',':
        fn(ctx, wid, 99, l + 1);
        if (l + 1 < n && s[l + 1] == ' ') fn(ctx, wid, 58, l + 2);
        break;
    case '(': fn(ctx, wid, 78, l + 1); break;
    }
}

/* probeSpacePrefixMatches / probeDotPrefixMatches (static_index.go:224-307): the word after a
 * leading ' ' or '.', with its listed suffixes */
static void probe_prefixed(const uint8_t *s, int n, uint8_t lead, dict_fn fn, void *ctx) {
    if (n < 5 || s[0] != lead) return;
    const uint8_t *sub = s + 1;
    const int sn = n - 1;
    const uint32_t prefix = rc_rd32(sub), h = (prefix * RC_M32) >> 16;
    int depth = 0;
    for (int32_t i = rc_head[h]; i >= 0 && depth < 32;) {
        const rc_entry *e = &rc_ent[i];
        i = e->next;
        depth++;
        if (e->prefix != prefix) continue;
        const int l = e->len;
        if (l > sn || memcmp(sub + 4, e->word + 4, (size_t)(l - 4))) continue;
        const uint32_t wid = e->p;
        if (wid >= 13504) continue;
        if (lead == ' ') {
            fn(ctx, wid, 6, 1 + l);
            if (1 + l < n) {
                switch (s[1 + l]) {
                case ' ': fn(ctx, wid, 2, 1 + l + 1); break;
                case '.':
                    fn(ctx, wid, 71, 1 + l + 1);
                    if (1 + l + 1 < n && s[1 + l + 1] == ' ') fn(ctx, wid, 52, 1 + l + 2);
                    break;
                case ',':
                    fn(ctx, wid, 103, 1 + l + 1);
                    if (1 + l + 1 < n && s[1 + l + 1] == ' ') fn(ctx, wid, 33, 1 + l + 2);
                    break;
                case '(': fn(ctx, wid, 89, 1 + l + 1); break;
                }
            }
        } else {
            fn(ctx, wid, 32, 1 + l);
            if (1 + l < n) {
                if (s[1 + l] == '(') fn(ctx, wid, 67, 1 + l + 1);
                else if (s[1 + l] == ' ') fn(ctx, wid, 77, 1 + l + 1);
            }
        }
    }
}

/* FindStaticDictMatches (static_index.go:110-168) over s[0, n) */
static void find_static_dict(const uint8_t *s, int n, dict_fn fn, void *ctx) {
    if (n < 4) return;
    const uint32_t prefix = rc_rd32(s), h = (prefix * RC_M32) >> 16;
    int depth = 0;
    for (int32_t i = rc_head[h]; i >= 0 && depth < 32;) {
        const rc_entry *e = &rc_ent[i];
        i = e->next;
        depth++;
        if (e->prefix != prefix) continue;
        const int l = e->len;
        if (l > n) continue;
        if (!memcmp(s + 4, e->word + 4, (size_t)(l - 4))) {
            fn(ctx, e->p, 0, l);
            probe_suffixes(s, n, l, e->p, fn, ctx);
        }
    }
    if (s[0] >= 'A' && s[0] <= 'Z') {
        const uint32_t lp = prefix | 0x20u, lh = (lp * RC_M32) >> 16;
        depth = 0;
        for (int32_t i = rc_head[lh]; i >= 0 && depth < 32;) {
            const rc_entry *e = &rc_ent[i];
            i = e->next;
            depth++;
            if (e->prefix != lp) continue;
            const int l = e->len;
            if (l > n) continue;
            if (s[0] == (uint8_t)(e->word[0] - 32) && !memcmp(s + 1, e->word + 1, (size_t)(l - 1)) && e->p < 13504) {
                fn(ctx, e->p, 9, l);
                probe_upper_suffixes(s, n, l, e->p, fn, ctx);
            }
        }
    }
    probe_prefixed(s, n, ' ', fn, ctx);
    probe_prefixed(s, n, '.', fn, ctx);
}

/* ------------------------------------------------------------------ */
/* The trellis (compress.go:390-946)                                    */

typedef struct { uint16_t from, len; uint32_t dist; uint8_t kind, arg; } rc_edge; /* commands.TrellisEdge */
typedef struct { uint32_t cost; uint32_t r[4]; rc_edge edge; uint16_t nlit; } rc_node;

typedef struct {
    int32_t head4[65536];
    int32_t head8[65536][8], head12[65536][16], head16[65536][16], head24[65536][8], head32[65536][8];
    int32_t *prev;      /* prevPos, window-sized */
    uint32_t prev_cap;
    rc_node nodes[RC_HORIZON + 1];
    uint32_t lit_cost[256];
    rc_edge edges[RC_HORIZON];
} rc_scratch;

typedef struct {
    const uint8_t *src;
    int n;                 /* len(src) */
    int bstart, bsize, maxdist, winmask, bmin, depth, level;
    uint32_t stride;
    rc_scratch *s;
} rc_ctx;

/* relax node i + l with a candidate edge (the update every relaxation shares) */
static inline void relax(rc_ctx *c, int i, int l, uint32_t cost, const uint32_t r[4], rc_edge e) {
    rc_node *t = &c->s->nodes[i + l];
    if (cost < t->cost) {
        t->cost = cost;
        memcpy(t->r, r, sizeof t->r);
        t->nlit = 0;
        t->edge = e;
    }
}

/* compress.go:390-400 */
static void relax_literal(rc_ctx *c, int i) {
    rc_node *nd = c->s->nodes;
    const uint8_t b = c->src[c->bstart + i];
    const uint32_t next = nd[i].cost + c->s->lit_cost[b];
    if (next < nd[i + 1].cost) {
        nd[i + 1].cost = next;
        memcpy(nd[i + 1].r, nd[i].r, sizeof nd[i].r);
        nd[i + 1].nlit = (uint16_t)(nd[i].nlit + 1);
        nd[i + 1].edge = (rc_edge){(uint16_t)i, 1, 0, 0, 0};
    }
}

/* compress.go:426-447 */
static void relax_repeat_sublengths(rc_ctx *c, int i, int k, uint32_t d, int L, uint32_t cur) {
    uint32_t rn[4];
    memcpy(rn, c->s->nodes[i].r, sizeof rn);
    lru_repeat(rn, k);
    int subs[32];
    const int ns = pareto_sublengths(L, c->bmin, subs);
    const int nlit = c->s->nodes[i].nlit;
    for (int j = 0; j < ns; j++) {
        const int l = subs[j];
        relax(c, i, l, cur + cost_repeat(l, k, c->bmin, nlit), rn, (rc_edge){(uint16_t)i, (uint16_t)l, d, 1, (uint8_t)k});
    }
}

/* compress.go:402-424 */
static int relax_repeat(rc_ctx *c, int i) {
    const int gp = c->bstart + i;
    const uint32_t cur = c->s->nodes[i].cost;
    int found = 0;
    for (int k = 0; k < 4; k++) {
        const uint32_t d = c->s->nodes[i].r[k];
        if (d == 0 || gp < (int64_t)d) continue;
        const int maxl = rc_min(c->bsize - i, RC_MAXM);
        if (maxl < c->bmin) continue;
        const int L = match_len(c->src, gp, gp - (int)d, maxl);
        if (L < c->bmin) continue;
        if (L > found) found = L;
        relax_repeat_sublengths(c, i, k, d, L, cur);
    }
    return found;
}

/* compress.go:491-509 */
static void relax_fixed_sublengths(rc_ctx *c, int i, uint8_t kind, uint8_t arg, uint32_t d, int L, uint32_t cur) {
    uint32_t rn[4];
    memcpy(rn, c->s->nodes[i].r, sizeof rn);
    lru_explicit(rn, d);
    int subs[32];
    const int ns = pareto_sublengths(L, c->bmin, subs);
    const int nlit = c->s->nodes[i].nlit;
    uint32_t base = cur + 25;
    if (nlit > 15) base += lit_flush_cost(nlit);
    for (int j = 0; j < ns; j++)
        relax(c, i, subs[j], base, rn, (rc_edge){(uint16_t)i, (uint16_t)subs[j], d, kind, arg});
}

/* compress.go:449-467 */
static void relax_brackets(rc_ctx *c, int i) {
    const int gp = c->bstart + i;
    const uint32_t cur = c->s->nodes[i].cost;
    for (int b = 0; b < 10; b++) {
        const uint32_t d = bracket_distance(b);
        if (gp < (int)d) continue;
        const int maxl = rc_min(c->bsize - i, 15 + c->bmin);
        if (maxl < c->bmin) continue;
        const int L = match_len(c->src, gp, gp - (int)d, maxl);
        if (L < c->bmin) continue;
        relax_fixed_sublengths(c, i, 2, (uint8_t)b, d, L, cur);
    }
}

/* compress.go:469-489 */
static void relax_stride(rc_ctx *c, int i) {
    if (c->stride <= 1) return;
    const int gp = c->bstart + i;
    const uint32_t cur = c->s->nodes[i].cost;
    for (int m = 1; m <= 8; m++) {
        const uint32_t d = (uint32_t)m * c->stride;
        if (gp < (int64_t)d) continue;
        const int maxl = rc_min(c->bsize - i, 15 + c->bmin);
        if (maxl < c->bmin) continue;
        const int L = match_len(c->src, gp, gp - (int)d, maxl);
        if (L < c->bmin) continue;
        relax_fixed_sublengths(c, i, 3, (uint8_t)m, d, L, cur);
    }
}

/* compress.go:812-855 */
static void relax_lattice_sublengths(rc_ctx *c, int i, uint8_t mode, int8_t delta, uint32_t d, int L, uint32_t cur) {
    uint32_t rn[4];
    memcpy(rn, c->s->nodes[i].r, sizeof rn);
    lru_explicit(rn, d);
    int subs[32];
    const int ns = pareto_sublengths(L, c->bmin, subs);
    const int nlit = c->s->nodes[i].nlit;
    const uint32_t flush = nlit > 6 ? lit_flush_cost(nlit) : 0;
    for (int j = 0; j < ns; j++) {
        const int l = subs[j];
        const uint32_t lat = cur + (l - c->bmin < 7 ? 30u : 40u) + flush;
        const uint32_t exp = cur + cost_explicit(l, d, c->bmin, nlit);
        if (lat >= exp) continue;
        relax(c, i, l, lat, rn, (rc_edge){(uint16_t)i, (uint16_t)l, (uint32_t)(uint8_t)delta, 6, mode});
    }
}

/* compress.go:709-749 */
static void relax_explicit_sublengths(rc_ctx *c, int i, uint32_t d, int L, uint32_t cur) {
    const uint32_t *r = c->s->nodes[i].r;
    for (int k = 0; k < 4; k++)
        if (d == r[k]) { relax_repeat_sublengths(c, i, k, d, L, cur); return; }
    if (d > 2048) {
        for (int mode = 0; mode < 3; mode++) {
            const int64_t diff = (int64_t)d - (int64_t)r[mode];
            if (diff >= -128 && diff <= 127 && diff != 0) relax_lattice_sublengths(c, i, (uint8_t)mode, (int8_t)diff, d, L, cur);
        }
        if (c->stride > 1) {
            const int64_t diff = (int64_t)d - (int64_t)r[0];
            if (diff % (int64_t)c->stride == 0) {
                const int64_t delta = diff / (int64_t)c->stride;
                if (delta >= -128 && delta <= 127 && delta != 0) relax_lattice_sublengths(c, i, 3, (int8_t)delta, d, L, cur);
            }
        }
    }
    uint32_t rn[4];
    memcpy(rn, r, sizeof rn);
    lru_explicit(rn, d);
    int subs[32];
    const int ns = pareto_sublengths(L, c->bmin, subs);
    const int nlit = c->s->nodes[i].nlit;
    for (int j = 0; j < ns; j++) {
        const int l = subs[j];
        relax(c, i, l, cur + cost_explicit(l, d, c->bmin, nlit), rn, (rc_edge){(uint16_t)i, (uint16_t)l, d, 4, 0});
    }
}

/* compress.go:541-679: the fixed-size tiers; W bytes hashed, the table's K slots, lengths >= W */
static void relax_tier(rc_ctx *c, int i, const int32_t *slots, int K, int W, uint32_t cur) {
    const int gp = c->bstart + i;
    const int maxl = rc_min(c->bsize - i, RC_MAXM);
    for (int k = 0; k < K; k++) {
        const int32_t cand = slots[k];
        if (cand < 0) break;
        const uint32_t d = (uint32_t)(gp - cand);
        if (d == 0 || d > (uint32_t)c->maxdist) continue;
        const int L = match_len(c->src, gp, cand, maxl);
        if (L >= W) relax_explicit_sublengths(c, i, d, L, cur);
    }
}

/* compress.go:511-539, 681-707 */
static void relax_hash_matches(rc_ctx *c, int i) {
    if (i + 4 > c->bsize) return;
    rc_scratch *s = c->s;
    const int gp = c->bstart + i;
    const uint32_t cur = s->nodes[i].cost;
    if (c->level >= 6 && i + 32 <= c->bsize && gp + 32 <= c->n) relax_tier(c, i, s->head32[hash32(c->src, gp)], 8, 32, cur);
    if (c->level >= 6 && i + 24 <= c->bsize && gp + 24 <= c->n) relax_tier(c, i, s->head24[hash24(c->src, gp)], 8, 24, cur);
    if (c->level >= 5 && i + 16 <= c->bsize && gp + 16 <= c->n) relax_tier(c, i, s->head16[hash16(c->src, gp)], 16, 16, cur);
    if (c->level >= 4 && i + 12 <= c->bsize && gp + 12 <= c->n) relax_tier(c, i, s->head12[hash12(c->src, gp)], 16, 12, cur);
    if (c->level >= 3 && i + 8 <= c->bsize && gp + 8 <= c->n) relax_tier(c, i, s->head8[hash8(c->src, gp)], 8, 8, cur);
    /* relaxHash4Matches: the chain through prevPos, depth by level */
    int32_t cand = s->head4[hash4(c->src, gp)];
    for (int depth = 0; cand >= 0 && depth < c->depth; depth++) {
        const uint32_t d = (uint32_t)(gp - cand);
        if (d == 0 || d > (uint32_t)c->maxdist) break;
        const int maxl = rc_min(c->bsize - i, RC_MAXM);
        const int L = match_len(c->src, gp, cand, maxl);
        if (L >= c->bmin) relax_explicit_sublengths(c, i, d, L, cur);
        cand = s->prev[cand & c->winmask];
    }
}

typedef struct { rc_ctx *c; int i; uint32_t cur, flush; } dict_cb;

/* the callback of relaxDictMatches (compress.go:762-787) */
static void dict_edge(void *ctx, uint32_t wid, uint8_t tau, int len) {
    const dict_cb *b = (const dict_cb *)ctx;
    rc_ctx *c = b->c;
    if (len < 5 || b->i + len > c->bsize) return;
    uint32_t base = len >= 8 ? 18u : 24u;
    if (tau) base += 6;
    const uint32_t cost = b->cur + base + b->flush;
    rc_node *t = &c->s->nodes[b->i + len];
    if (cost < t->cost) {
        t->cost = cost;
        memcpy(t->r, c->s->nodes[b->i].r, sizeof t->r);
        t->nlit = 0;
        t->edge = (rc_edge){(uint16_t)b->i, (uint16_t)len, wid, 5, tau};
    }
}

/* compress.go:751-787 */
static void relax_dict(rc_ctx *c, int i) {
    if (!rc_tab.loaded) return;
    const int gp = c->bstart + i;
    if (i + 4 > c->bsize || gp + 4 > c->n) return;
    const int nlit = c->s->nodes[i].nlit;
    dict_cb b = {c, i, c->s->nodes[i].cost, nlit > 6 ? lit_flush_cost(nlit) : 0};
    find_static_dict(c->src + gp, c->n - gp, dict_edge, &b);
}

/* compress.go:789-810 */
static void relax_lattice(rc_ctx *c, int i) {
    if (c->stride <= 1) return;
    const int gp = c->bstart + i;
    const uint32_t cur = c->s->nodes[i].cost;
    static const int8_t deltas[4] = {-2, -1, 1, 2};
    for (int j = 0; j < 4; j++) {
        /* ComputeDiffLatticeDist(3 << 6, delta, stride, r) with stride > 1 */
        const int64_t d = (int64_t)c->s->nodes[i].r[0] + (int64_t)deltas[j] * (int64_t)c->stride;
        if (d <= 0 || gp < d) continue;
        const int maxl = rc_min(c->bsize - i, 13 + c->bmin);
        if (maxl < c->bmin) continue;
        const int L = match_len(c->src, gp, gp - (int)d, maxl);
        if (L < c->bmin) continue;
        relax_lattice_sublengths(c, i, 3, deltas[j], (uint32_t)d, L, cur);
    }
}

/* compress.go:857-888 */
static void insert_hash(rc_ctx *c, int gp) {
    rc_scratch *s = c->s;
    const uint16_t h4 = hash4(c->src, gp);
    s->prev[gp & c->winmask] = s->head4[h4];
    s->head4[h4] = gp;
#define PUSH(tab, K, h) do { memmove(&s->tab[h][1], &s->tab[h][0], (K - 1) * sizeof(int32_t)); s->tab[h][0] = gp; } while (0)
    if (gp + 8 <= c->n) PUSH(head8, 8, hash8(c->src, gp));
    if (gp + 12 <= c->n) PUSH(head12, 16, hash12(c->src, gp));
    if (gp + 16 <= c->n) PUSH(head16, 16, hash16(c->src, gp));
    if (gp + 24 <= c->n) PUSH(head24, 8, hash24(c->src, gp));
    if (gp + 32 <= c->n) PUSH(head32, 8, hash32(c->src, gp));
#undef PUSH
}

/* ------------------------------------------------------------------ */
/* Command emission (core/commands/commands.go:580-752)                 */

typedef struct { uint8_t *b; size_t n, cap; } rc_buf;
static void rb_put(rc_buf *o, uint8_t v) {
    if (o->n == o->cap) { o->cap = o->cap ? 2 * o->cap : 4096; o->b = (uint8_t *)realloc(o->b, o->cap); }
    o->b[o->n++] = v;
}
static void rb_leb(rc_buf *o, uint32_t v) {
    while (v >= 0x80) { rb_put(o, (uint8_t)(v | 0x80)); v >>= 7; }
    rb_put(o, (uint8_t)v);
}
typedef struct { uint32_t *v; size_t n, cap; } rc_offs;
static void ro_put(rc_offs *o, uint32_t v) {
    if (o->n == o->cap) { o->cap = o->cap ? 2 * o->cap : 1024; o->v = (uint32_t *)realloc(o->v, o->cap * 4); }
    o->v[o->n++] = v;
}

/* EmitLiteralRun (commands.go:738-752): the command only; the bytes are already in lits */
static void emit_lit_run(int n, rc_buf *cmds) {
    if (n == 0) return;
    if (n <= 8) { rb_put(cmds, (uint8_t)(0xF2 + n)); return; }
    rb_put(cmds, 0xF0);
    rb_leb(cmds, (uint32_t)n);
}

static void emit_block(const rc_edge *edges, int ne, const uint8_t *block, int bmin, rc_buf *cmds, rc_buf *lits,
                       rc_offs *offs) {
    int pending = 0;
    for (int j = 0; j < ne; j++) {
        const rc_edge e = edges[j];
        if (e.kind == 0) {
            rb_put(lits, block[e.from]);
            pending++;
            continue;
        }
        int nlit = pending;
        const int L = e.len;
        switch (e.kind) {
        case 1: /* emitRepeatMatch, :667-689 */
            if (nlit <= 7 && L <= 3 + bmin) { rb_put(cmds, (uint8_t)((nlit << 4) | ((L - bmin) << 2) | e.arg)); break; }
            if (nlit > 15) { emit_lit_run(nlit, cmds); nlit = 0; }
            rb_put(cmds, (uint8_t)(0xC0 | e.arg));
            if (L - bmin < 15) rb_put(cmds, (uint8_t)((nlit << 4) | (L - bmin)));
            else { rb_put(cmds, (uint8_t)((nlit << 4) | 15)); rb_leb(cmds, (uint32_t)(L - bmin - 15)); }
            break;
        case 2: /* emitFixedMatch, :691-700 */
        case 3:
            if (nlit > 15) { emit_lit_run(nlit, cmds); nlit = 0; }
            rb_put(cmds, e.kind == 2 ? (uint8_t)(0xC4 + e.arg) : (uint8_t)(0xD0 + e.arg - 1));
            rb_put(cmds, (uint8_t)((nlit << 4) | (L - bmin)));
            break;
        case 4: /* emitExplicitMatch, :702-736 */
            if (nlit > 15) { emit_lit_run(nlit, cmds); nlit = 0; }
            ro_put(offs, e.dist);
            if (nlit <= 3) {
                if (L - bmin < 15) rb_put(cmds, (uint8_t)(0x80 | (nlit << 4) | (L - bmin)));
                else { rb_put(cmds, (uint8_t)(0x80 | (nlit << 4) | 15)); rb_leb(cmds, (uint32_t)(L - bmin - 15)); }
            } else {
                rb_put(cmds, 0xE9);
                if (L - bmin < 15) rb_put(cmds, (uint8_t)((nlit << 4) | (L - bmin)));
                else { rb_put(cmds, (uint8_t)((nlit << 4) | 15)); rb_leb(cmds, (uint32_t)(L - bmin - 15)); }
            }
            break;
        case 5: { /* emitRFCWordMatch, :627-645 */
            if (nlit > 6) { emit_lit_run(nlit, cmds); nlit = 0; }
            const uint8_t op = nlit > 0 ? (uint8_t)(0xEA + nlit - 1) : 0xCF;
            rb_put(cmds, op);
            if (e.arg) { rb_put(cmds, e.arg); rb_leb(cmds, e.dist); }
            else { rb_put(cmds, (uint8_t)(0x80 | ((e.dist >> 7) & 0x7F))); rb_put(cmds, (uint8_t)(e.dist & 0x7F)); }
            break;
        }
        case 6: { /* emitDiffLatticeMatch, :647-665 */
            if (nlit > 6) { emit_lit_run(nlit, cmds); nlit = 0; }
            const int diff = L - bmin;
            rb_put(cmds, 0xCE);
            rb_put(cmds, (uint8_t)((e.arg << 6) | (nlit << 3) | (diff < 7 ? diff : 7)));
            rb_put(cmds, (uint8_t)e.dist);
            if (diff >= 7) rb_leb(cmds, (uint32_t)(diff - 7));
            break;
        }
        }
        pending = 0;
    }
    emit_lit_run(pending, cmds);
}

/* SplitCommands with every opcode family (commands.go:18-202); 0, or -1 on a malformed stream */
static int split_varint(const uint8_t *s, size_t n, size_t *sp, uint8_t *v, uint32_t *nv) {
    for (int k = 0; k < 4; k++) {
        if (*sp >= n) return -1;
        const uint8_t b = s[(*sp)++];
        v[(*nv)++] = b;
        if (!(b & 0x80)) return 0;
    }
    return -1;
}
static int split_all(const uint8_t *s, size_t n, orc_streams *o) {
    size_t sp = 0;
    o->nop = o->nmeta = o->nvar = 0;
    while (sp < n) {
        const uint8_t op = s[sp++];
        o->op[o->nop++] = op;
        if (op < 0x80) continue;
        if (op < 0xC0) {
            if ((op & 15) == 15 && split_varint(s, n, &sp, o->var, &o->nvar)) return -1;
            continue;
        }
        if (op >= 0xF0) { /* splitSpecialOp */
            if (op == 0xF0 || op == 0xF1) { if (split_varint(s, n, &sp, o->var, &o->nvar)) return -1; continue; }
            if (op >= 0xF3 && op <= 0xFA) continue;
            return -1;
        }
        if (sp >= n) return -1;
        const uint8_t t = s[sp++];
        o->meta[o->nmeta++] = t;
        if (op <= 0xC3 || op == 0xE9) { /* long repeat, external dictionary */
            if ((t & 15) == 15 && split_varint(s, n, &sp, o->var, &o->nvar)) return -1;
        } else if (op <= 0xCD) { /* bracket */
        } else if (op == 0xCE) { /* difference lattice: descriptor, delta */
            if (sp >= n) return -1;
            o->meta[o->nmeta++] = s[sp++];
            if (((t >> 3) & 7) == 7 && split_varint(s, n, &sp, o->var, &o->nvar)) return -1;
            if ((t & 7) == 7 && split_varint(s, n, &sp, o->var, &o->nvar)) return -1;
        } else if (op == 0xCF || (op >= 0xEA && op <= 0xEF)) { /* RFC word */
            if (t >= 128) { if (sp >= n) return -1; o->var[o->nvar++] = s[sp++]; }
            else if (split_varint(s, n, &sp, o->var, &o->nvar)) return -1;
        } else if (op <= 0xD7) { /* stride */
        } else { /* delta mask, 0xD8-0xE8 */
            if (t == 255 && split_varint(s, n, &sp, o->var, &o->nvar)) return -1;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* compressInternal (compress.go:1036-1114) and the frame               */

typedef struct {
    uint32_t stride, bmin, wlog, uncompressed; /* AnalyzeInput */
    uint32_t kind;                             /* 0 compressed, 1 uncompressed (entropy / empty), 2 size, 3 S4 hazard */
    uint32_t ncmd, noff, nlit;
    uint32_t s2, s3, s4;
    uint32_t edges[7];                         /* edges per kind along the chosen path */
    uint32_t dict;                             /* 1 when the dictionary relaxation ran */
} orc_ref_info;

/* countStrideMatches / detectStride (compress.go:135-181) */
static uint32_t detect_stride(const uint8_t *src, int n) {
    if (n < 128) return 1;
    static const uint32_t cands[21] = {2, 3, 4, 5, 6, 7, 8, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64, 72, 80, 96, 128};
    const int offsets[4] = {0, n / 4, n / 2, (3 * n) / 4};
    const int chunk = 2048;
    uint32_t best = 1;
    int max_m = 0;
    for (int ci = 0; ci < 21; ci++) {
        const uint32_t s = cands[ci];
        int tm = 0, tl = 0;
        for (int k = 0; k < 4; k++) {
            int off = offsets[k];
            if (off + chunk > n) off = n - chunk > 0 ? n - chunk : 0;
            const int end = off + chunk < n ? off + chunk : n, len = end - off;
            if ((int)s >= len) continue;
            for (int i = (int)s; i < len; i++) tm += src[off + i] == src[off + i - (int)s];
            tl += len;
        }
        int thr = max_m;
        if (best > 1 && s % best == 0) thr = (max_m * 115) / 100;
        if (tl > 0 && tm > thr && tm > (tl * 20) / 100) { max_m = tm; best = s; }
    }
    return best;
}

static __thread rc_scratch *tl_scratch;

static uint32_t raw_frame(const uint8_t *src, uint32_t n, uint8_t *out) { /* compress.go:973-981 */
    memcpy(out, "PRZM", 4);
    out[4] = 0x04;
    out[5] = (uint8_t)n; out[6] = (uint8_t)(n >> 8); out[7] = (uint8_t)(n >> 16);
    if (n) memcpy(out + 8, src, n);
    return 8 + n;
}

/*
 * One compressInternal call: src[0, n) (n < 2^24) at `level` -> one frame in out (which holds
 * n + 64 bytes).  Returns the frame's length, or 0 when n is out of range.  cmd_out (optional,
 * cmd_cap bytes) receives the command stream before Section 2 packs it; info->ncmd its length.
 */
uint64_t orc_ref_compress_ex(const uint8_t *src, uint64_t n64, int level, uint8_t *out, orc_ref_info *info,
                             uint8_t *cmd_out, uint64_t cmd_cap) {
    orc_ref_info dummy;
    if (!info) info = &dummy;
    memset(info, 0, sizeof *info);
    if (n64 >= (1u << 24)) return 0;
    const int n = (int)n64;
    pthread_once(&rc_once, rc_init_costs);
    if (level <= 0) level = 6;
    /* AnalyzeInput (compress.go:93-111) */
    info->bmin = 4;
    if (n == 0) info->wlog = 16;
    else {
        info->uncompressed = (uint32_t)orc_check_high_entropy(src, (uint32_t)n);
        if (!info->uncompressed) {
            info->stride = detect_stride(src, n);
            if (info->stride > 1) info->bmin = 3;
            info->wlog = orc_window_log((uint32_t)n);
        }
    }
    if (info->uncompressed) { info->kind = 1; return raw_frame(src, (uint32_t)n, out); }
    rc_scratch *s = tl_scratch;
    if (!s) s = tl_scratch = (rc_scratch *)calloc(1, sizeof(rc_scratch));
    const uint32_t win = 1u << info->wlog;
    if (s->prev_cap < win) { free(s->prev); s->prev = (int32_t *)malloc((size_t)win * 4); s->prev_cap = win; }
    /* executeCompressionPipeline (compress.go:1060-1114) */
    for (int i = 0; i < 65536; i++) s->head4[i] = -1;
    memset(s->head8, 0xFF, sizeof s->head8);
    memset(s->head12, 0xFF, sizeof s->head12);
    memset(s->head16, 0xFF, sizeof s->head16);
    memset(s->head24, 0xFF, sizeof s->head24);
    memset(s->head32, 0xFF, sizeof s->head32);
    for (int i = 0; i < (n < (int)win ? n : (int)win); i++) s->prev[i] = -1;
    rc_ctx c = {src, n, 0, 0, (int)win, (int)win - 1, (int)info->bmin, 0, level, info->stride, s};
    c.depth = level <= 1 ? 2 : level <= 3 ? 4 : level <= 5 ? 8 : level <= 7 ? 16 : level <= 9 ? 32 : level <= 10 ? 64 : 128;
    info->dict = (uint32_t)rc_tab.loaded;
    uint32_t R[4] = {1, 4, 8, 16};
    if (info->stride > 1) { R[0] = 1; R[1] = info->stride; R[2] = 2 * info->stride; R[3] = 3 * info->stride; }
    rc_buf cmds = {0}, lits = {0};
    rc_offs offs = {0};
    for (int pos = 0; pos < n;) {
        const int bs = n - pos < (int)RC_HORIZON ? n - pos : (int)RC_HORIZON;
        c.bstart = pos;
        c.bsize = bs;
        /* initBlockNodes (compress.go:891-900) */
        for (int i = 0; i <= bs; i++) { s->nodes[i].cost = RC_INF; s->nodes[i].nlit = 0; }
        s->nodes[0].cost = 0;
        memcpy(s->nodes[0].r, R, sizeof R);
        /* solveBlockTrellis (compress.go:902-930) */
        int freq[256] = {0};
        for (int i = 0; i < bs; i++) freq[src[pos + i]]++;
        const double inv = 1.0 / (double)bs;
        for (int b = 0; b < 256; b++) s->lit_cost[b] = calc_lit_cost(freq[b], inv);
        for (int i = 0; i < bs; i++) {
            if (s->nodes[i].cost != RC_INF) {
                relax_literal(&c, i);
                const int rep = relax_repeat(&c, i);
                if (rep < rc_min(bs - i, RC_MAXM)) {
                    relax_brackets(&c, i);
                    relax_stride(&c, i);
                    relax_hash_matches(&c, i);
                    relax_dict(&c, i);
                    relax_lattice(&c, i);
                }
            }
            if (i + 4 <= bs) insert_hash(&c, pos + i);
        }
        /* backtrackEdges (compress.go:932-946) */
        int ne = 0;
        for (int cur = bs; cur > 0; cur = s->nodes[cur].edge.from) s->edges[ne++] = s->nodes[cur].edge;
        for (int a = 0, b = ne - 1; a < b; a++, b--) { const rc_edge t = s->edges[a]; s->edges[a] = s->edges[b]; s->edges[b] = t; }
        for (int j = 0; j < ne; j++) info->edges[s->edges[j].kind]++;
        emit_block(s->edges, ne, src + pos, (int)info->bmin, &cmds, &lits, &offs);
        memcpy(R, s->nodes[bs].r, sizeof R);
        pos += bs;
    }
    info->ncmd = (uint32_t)cmds.n; info->noff = (uint32_t)offs.n; info->nlit = (uint32_t)lits.n;
    if (cmd_out && cmds.n <= cmd_cap && cmds.n) memcpy(cmd_out, cmds.b, cmds.n);
    /* packageSections (compress.go:948-960) */
    uint8_t *s2 = NULL, *s3 = NULL, *s4 = NULL;
    uint32_t l2 = 0, l3 = 0, l4 = 0;
    uint64_t len = 0;
    if (cmds.n) {
        orc_streams st;
        memset(&st, 0, sizeof st);
        uint8_t *mem = (uint8_t *)malloc(3 * cmds.n + 16);
        st.cmd = cmds.b; st.ncmd = (uint32_t)cmds.n;
        st.op = mem; st.meta = mem + cmds.n + 4; st.var = mem + 2 * cmds.n + 8;
        s2 = (uint8_t *)malloc(16 + 2 * cmds.n + 3 * 640);
        /* PackageCommandsSection (commands.go:537-568): Mode 8 when it is shorter, else flat */
        if (split_all(cmds.b, cmds.n, &st) == 0) l2 = orc_package_commands(&st, s2, NULL);
        else { s2[0] = 0x09; memcpy(s2 + 1, cmds.b, cmds.n); l2 = 1 + (uint32_t)cmds.n; }
        free(mem);
    }
    if (offs.n) {
        rc_buf leb = {0};
        for (size_t j = 0; j < offs.n; j++) rb_leb(&leb, offs.v[j]);
        s3 = (uint8_t *)malloc(16 + 3 * leb.n + 700);
        int mode, ck;
        l3 = orc_package_offsets(leb.b, (uint32_t)leb.n, s3, &mode, &ck);
        free(leb.b);
    }
    int kind4 = 0;
    if (lits.n) {
        s4 = (uint8_t *)malloc(40 + 2400 + 16 + 2 * lits.n);
        l4 = orc_package_literals(lits.b, (uint32_t)lits.n, s4, &kind4);
    }
    info->s2 = l2; info->s3 = l3; info->s4 = l4;
    /* assembleContainer / assembleCompressedFrame (compress.go:962-1033) */
    const int ext = l2 > 65535 || l3 > 65535;
    const uint32_t dir = ext ? 6 : 4;
    uint64_t comp = 8ull + dir + 8 + l2 + l3 + l4 + (info->stride > 1 ? 1 : 0);
    const int hazard = !kind4 && l4 >= 8 &&
                       ((s4[0] == 'V' && s4[1] == 'A' && s4[2] == 0x01) || rc_rd32(s4 + 4) == (rc_rd32(s4) + 1) / 2);
    if (comp >= (uint64_t)n || hazard) {
        info->kind = comp >= (uint64_t)n ? 2 : 3;
        len = raw_frame(src, (uint32_t)n, out);
    } else {
        uint8_t flags = kind4 ? 0x80 : 0;
        if (info->stride > 1) flags |= 0x01;
        if (info->bmin == 3) flags |= 0x40;
        if (ext) flags |= 0x08;
        uint32_t p = 0;
        memcpy(out, "PRZM", 4);
        out[4] = flags;
        out[5] = (uint8_t)n; out[6] = (uint8_t)(n >> 8); out[7] = (uint8_t)(n >> 16);
        p = 8;
        if (info->stride > 1) out[p++] = (uint8_t)(info->stride - 1);
        if (ext) {
            out[p++] = (uint8_t)l2; out[p++] = (uint8_t)(l2 >> 8); out[p++] = (uint8_t)(l2 >> 16);
            out[p++] = (uint8_t)l3; out[p++] = (uint8_t)(l3 >> 8); out[p++] = (uint8_t)(l3 >> 16);
        } else {
            out[p++] = (uint8_t)l2; out[p++] = (uint8_t)(l2 >> 8);
            out[p++] = (uint8_t)l3; out[p++] = (uint8_t)(l3 >> 8);
        }
        memset(out + p, 0, 8);
        out[p + 1] = (uint8_t)info->wlog;
        p += 8;
        if (l2) memcpy(out + p, s2, l2);
        p += l2;
        if (l3) memcpy(out + p, s3, l3);
        p += l3;
        if (l4) memcpy(out + p, s4, l4);
        p += l4;
        len = p;
    }
    free(s2); free(s3); free(s4);
    free(cmds.b); free(lits.b); free(offs.v);
    return len;
}

uint64_t orc_ref_compress(const uint8_t *src, uint64_t n, int level, uint8_t *out, orc_ref_info *info) {
    return orc_ref_compress_ex(src, n, level, out, info, NULL, 0);
}

/* ---- the parts, for the known-answer tests ---- */

int orc_ref_sublengths(int L, int bmin, int *out) { return pareto_sublengths(L, bmin, out); }
uint32_t orc_ref_lit_cost(int f, int n) { return calc_lit_cost(f, 1.0 / (double)n); }
uint32_t orc_ref_dist_cost(int code) { pthread_once(&rc_once, rc_init_costs); return rc_dist_cost[code & 63]; }
uint32_t orc_ref_cost_explicit(int L, uint32_t d, int bmin, int nlit) {
    pthread_once(&rc_once, rc_init_costs);
    return cost_explicit(L, d, bmin, nlit);
}
uint32_t orc_ref_cost_repeat(int L, int k, int bmin, int nlit) { return cost_repeat(L, k, bmin, nlit); }
uint32_t orc_ref_detect_stride(const uint8_t *src, int n) { return detect_stride(src, n); }
uint32_t orc_ref_hash(int width, const uint8_t *s) {
    switch (width) {
    case 4: return hash4(s, 0);
    case 8: return hash8(s, 0);
    case 12: return hash12(s, 0);
    case 16: return hash16(s, 0);
    case 24: return hash24(s, 0);
    default: return hash32(s, 0);
    }
}

typedef struct { uint32_t *wid; uint8_t *tau; int32_t *len; int n, cap; } find_out;
static void find_collect(void *ctx, uint32_t wid, uint8_t tau, int len) {
    find_out *o = (find_out *)ctx;
    if (o->n < o->cap) { o->wid[o->n] = wid; o->tau[o->n] = tau; o->len[o->n] = len; }
    o->n++;
}
/* FindStaticDictMatches(s[0, n)) in callback order; the number of matches, -1 without tables */
int orc_ref_find_dict(const uint8_t *s, int n, uint32_t *wid, uint8_t *tau, int32_t *len, int cap) {
    if (!rc_tab.loaded) return -1;
    find_out o = {wid, tau, len, 0, cap};
    find_static_dict(s, n, find_collect, &o);
    return o.n;
}
