/*
 * przm_decode.c -- CPU decoder of PRZM frames (include/przm_decode.h).
 *
 * Written from the format description (docs/PRZM_SPEC.md:94-154 frame,
 * 457-647 commands, offsets and V-ANS; decompress.go:113-260 header and
 * section framing; core/commands/commands.go:205-258, 474-520, 755-789
 * Section 2; core/distance/distance.go:262-321, 525-690 Section 3;
 * core/literals/literals.go:601-664, 696-736 and sheaf.go Section 4;
 * core/commands/vans_meta.go:45-68
 * Meta8 contexts) for the features listed in the header.  Every read is
 * bounds-checked; a frame either decodes completely or returns an error.
 */
#include "../../include/przm_decode.h"

#include <stdlib.h>
#include <string.h>

#include "przm_dist_freq.h"
#include "przm_meta8_freq.h"

#define MAGIC 0x4D5A5250u    /* "PRZM" little-endian */
#define MAGIC_S 0x535A5250u  /* "PRZS": a multi-frame stream (decompress.go:25) */
#define STREAM_HDR 17u        /* decompress.go:849-866 */
#define MAX_FRAME (16u << 20) /* MaxBlockSize (api.go:22) */
#define F_STRIDE 0x01u
#define F_DELTA 0x02u
#define F_RAW 0x04u
#define F_EXT 0x08u
#define F_DICT 0x10u
#define F_NORM 0x20u
#define F_MIN3 0x40u

static uint32_t le32(const uint8_t *p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

/* ULEB128 of at most 4 bytes whose value fits a frame (decompress.go:113-134) */
static int uleb(const uint8_t *b, size_t n, size_t *pos, uint32_t *v) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) {
        if (*pos >= n) return PRZM_DEC_ERR_CORRUPT;
        const uint8_t c = b[(*pos)++];
        r |= (uint32_t)(c & 0x7F) << (7 * k);
        if (!(c & 0x80)) {
            if (r > MAX_FRAME) return PRZM_DEC_ERR_CORRUPT;
            *v = r;
            return 0;
        }
    }
    return PRZM_DEC_ERR_CORRUPT;
}

/* ---------------- 4-lane rANS ---------------- */

/* Decodes n symbols, lane i & 3 for symbol i, from the four states at st (16 bytes) and the
 * byte stream s[0, sn), with one table or one per Meta8 context of the previous symbol; every
 * lane must end at 65536 (the encoder's initial state) and every stream byte be used. */
typedef struct {
    const uint16_t *freq[8];  /* per context: frequency of each symbol */
    const uint16_t *cdf[8];
    const uint8_t *slot[8];   /* per context: symbol of each of the 4096 slots */
    int contexts;             /* 1: adaptive table; 8: Meta8 contexts of the previous symbol */
} rans_tables;

static int meta_ctx(uint8_t p) { /* vans_meta.go:45-68 */
    if (p == 0) return 0;
    if (p < 4) return 1;
    if (p < 16) return 2;
    if (p < 32) return 3;
    if (p < 64) return 4;
    if (p < 128) return 5;
    if (p < 192) return 6;
    return 7;
}

static int rans_run(const rans_tables *t, const uint8_t *st, const uint8_t *s, size_t sn, uint8_t *dst, uint32_t n) {
    uint32_t x[4];
    for (int j = 0; j < 4; j++) {
        x[j] = le32(st + 4 * j);
        if (x[j] < 65536u) return PRZM_DEC_ERR_CORRUPT;
    }
    size_t sp = 0;
    for (uint32_t i = 0; i < n; i++) {
        const int j = (int)(i & 3u);
        const int c = t->contexts == 8 && i > 0 ? meta_ctx(dst[i - 1]) : 0;
        const uint32_t slot = x[j] & 4095u;
        const uint8_t sym = t->slot[c][slot];
        dst[i] = sym;
        x[j] = (uint32_t)t->freq[c][sym] * (x[j] >> 12) + slot - t->cdf[c][sym];
        while (x[j] < 65536u) {
            if (sp >= sn) return PRZM_DEC_ERR_CORRUPT;
            x[j] = (x[j] << 8) | s[sp++];
        }
    }
    if (sp != sn) return PRZM_DEC_ERR_CORRUPT;
    for (int j = 0; j < 4; j++)
        if (x[j] != 65536u) return PRZM_DEC_ERR_CORRUPT;
    return 0;
}

/* cdf and slot table of a frequency table that sums to 4096 */
static int build_table(const uint16_t *freq, uint16_t *cdf, uint8_t *slot) {
    uint32_t c = 0;
    for (int s = 0; s < 256; s++) {
        cdf[s] = (uint16_t)c;
        if (c + freq[s] > 4096u) return PRZM_DEC_ERR_CORRUPT;
        memset(slot + c, s, freq[s]);
        c += freq[s];
    }
    return c == 4096u ? 0 : PRZM_DEC_ERR_CORRUPT;
}

/* from c[p]: the joint-nibble matrix (adaptive) or nothing (Meta8), four states and the stream:
 * n symbols into dst */
static int64_t vans_body(const uint8_t *c, size_t cn, size_t p, uint8_t *dst, uint32_t n, int meta8) {
    rans_tables t;
    memset(&t, 0, sizeof t);
    uint16_t freq[256], cdf[256];
    uint16_t (*mcdf)[256] = NULL;
    uint8_t *slots = NULL;
    int rc = 0;
    if (meta8) {
        mcdf = (uint16_t (*)[256])malloc(8 * 256 * sizeof(uint16_t));
        slots = (uint8_t *)malloc(8 * 4096);
        if (!mcdf || !slots) { rc = PRZM_DEC_ERR_CAPACITY; goto out; }
        for (int k = 0; k < 8; k++) {
            t.freq[k] = przm_meta8_freq[k];
            if ((rc = build_table(przm_meta8_freq[k], mcdf[k], slots + 4096 * k))) goto out;
            t.cdf[k] = mcdf[k];
            t.slot[k] = slots + 4096 * k;
        }
        t.contexts = 8;
    } else {
        /* row mask, then per row a column mask and 1- or 2-byte weights (docs/PRZM_SPEC.md:595-610) */
        memset(freq, 0, sizeof freq);
        if (p + 2 > cn) { rc = PRZM_DEC_ERR_CORRUPT; goto out; }
        const uint32_t rows = (uint32_t)c[p] | (uint32_t)c[p + 1] << 8;
        p += 2;
        for (int h = 0; h < 16; h++) {
            if (!(rows >> h & 1u)) continue;
            if (p + 2 > cn) { rc = PRZM_DEC_ERR_CORRUPT; goto out; }
            const uint32_t cols = (uint32_t)c[p] | (uint32_t)c[p + 1] << 8;
            p += 2;
            for (int l = 0; l < 16; l++) {
                if (!(cols >> l & 1u)) continue;
                if (p >= cn) { rc = PRZM_DEC_ERR_CORRUPT; goto out; }
                uint32_t w = c[p++];
                if (w & 0x80u) {
                    if (p >= cn) { rc = PRZM_DEC_ERR_CORRUPT; goto out; }
                    w = (w & 0x7Fu) << 8 | c[p++];
                }
                if (w == 0 || w > 4096u) { rc = PRZM_DEC_ERR_CORRUPT; goto out; }
                freq[h << 4 | l] = (uint16_t)w;
            }
        }
        slots = (uint8_t *)malloc(4096);
        if (!slots) { rc = PRZM_DEC_ERR_CAPACITY; goto out; }
        if ((rc = build_table(freq, cdf, slots))) goto out;
        t.freq[0] = freq;
        t.cdf[0] = cdf;
        t.slot[0] = slots;
        t.contexts = 1;
    }
    if (p + 16 > cn) { rc = PRZM_DEC_ERR_CORRUPT; goto out; }
    rc = rans_run(&t, c + p, c + p + 16, cn - p - 16, dst, n);
out:
    free(mcdf);
    free(slots);
    return rc ? rc : (int64_t)n;
}

/* 'V','A',1,flags, u32 n, then the body above: n symbols into dst (cap bytes) */
static int64_t vans_container(const uint8_t *c, size_t cn, uint8_t *dst, size_t cap, int meta8) {
    if (cn < 8) return PRZM_DEC_ERR_CORRUPT;
    const uint32_t n = le32(c + 4);
    if (n > cap) return PRZM_DEC_ERR_CAPACITY;
    return vans_body(c, cn, 8, dst, n, meta8);
}

/* ---------------- sheaf literal containers ---------------- */

static int sh_ws(uint8_t b) { return b == ' ' || b == '\t' || b == '\r' || b == '\n'; }
static int sh_syntax(uint8_t b) {
    switch (b) {
    case '{': case '}': case '[': case ']': case '(': case ')': case '<': case '>': case ':': case ';':
    case ',': case '.': case '"': case '\'': case '`': case '=': case '+': case '-': case '*': case '/':
    case '%': case '&': case '|': case '^': case '!': case '~': case '?': case '\\': case '#':
        return 1;
    default:
        return 0;
    }
}
static int sh_ident(uint8_t b) { return (b >= 'a' && b <= 'z') || b == '_'; }
static int sh_digit(uint8_t b) { return b >= '0' && b <= '9'; }

/* the fiber of the next literal after p2, p1 (sheaf.go:36-102) */
static int sh_class(uint8_t p2, uint8_t p1, int mode) {
    switch (mode) {
    case 0: return sh_ws(p1) ? 0 : sh_syntax(p1) ? 1 : sh_ident(p1) ? 2 : 3;
    case 1: return (p1 == '"' || p1 == ':') ? 0 : sh_ident(p1) ? 1 : sh_digit(p1) ? 2 : 3;
    case 2: return (sh_ws(p1) || sh_syntax(p1)) ? 0 : 1;
    default:
        if (sh_digit(p1)) return (sh_digit(p2) || p2 == '.' || p2 == '-') ? 0 : 1;
        if (sh_ident(p1)) return sh_ident(p2) ? 2 : 3;
        return sh_ws(p1) ? 0 : 1;
    }
}

/* 'V','A',1, 0x05 | mode << 4, u32 n, u32 fiber lengths l[4], u32 coded lengths e[4], then each
 * fiber raw (e == l) or as a V-ANS body of e bytes; the literals come back by walking the
 * contexts and taking each next byte from its fiber (literals.go:601-664, sheaf.go:125-202) */
static int64_t sheaf_container(const uint8_t *c, size_t cn, uint8_t *dst, uint32_t n) {
    if (cn < 40) return PRZM_DEC_ERR_CORRUPT;
    const int mode = c[3] >> 4;
    if (mode > 3) return PRZM_DEC_ERR_CORRUPT;
    uint32_t l[4], e[4];
    uint64_t sum = 0, pos = 40;
    for (int k = 0; k < 4; k++) {
        l[k] = le32(c + 8 + 4 * k);
        e[k] = le32(c + 24 + 4 * k);
        sum += l[k];
        if ((l[k] == 0 && e[k] != 0) || e[k] > l[k]) return PRZM_DEC_ERR_CORRUPT;
        pos += e[k];
    }
    if (sum != n || pos > cn || (mode == 2 && (l[2] || l[3]))) return PRZM_DEC_ERR_CORRUPT;
    uint8_t *fib = (uint8_t *)malloc((size_t)n + 16);
    if (!fib) return PRZM_DEC_ERR_CAPACITY;
    const uint8_t *f[4];
    size_t at = 40, fo = 0;
    int rc = 0;
    for (int k = 0; k < 4 && !rc; k++) {
        f[k] = fib + fo;
        if (l[k] == 0) continue;
        if (e[k] == l[k]) {
            memcpy(fib + fo, c + at, l[k]);
        } else {
            const int64_t r = vans_body(c + at, e[k], 0, fib + fo, l[k], 0);
            if (r < 0) rc = (int)r;
        }
        at += e[k];
        fo += l[k];
    }
    uint32_t used[4] = {0, 0, 0, 0};
    uint8_t p1 = 0, p2 = 0;
    for (uint32_t i = 0; i < n && !rc; i++) {
        const int k = sh_class(p2, p1, mode);
        if (used[k] >= l[k]) { rc = PRZM_DEC_ERR_CORRUPT; break; }
        dst[i] = f[k][used[k]++];
        p2 = p1;
        p1 = dst[i];
    }
    for (int k = 0; k < 4 && !rc; k++)
        if (used[k] != l[k]) rc = PRZM_DEC_ERR_CORRUPT;
    free(fib);
    return rc ? rc : (int64_t)n;
}

/* ---------------- sections ---------------- */

/* one Mode-8 substream (commands.go:474-520 unpackSubstreamOrVans); returns its decoded length */
static int64_t substream(const uint8_t *s, size_t n, uint8_t *dst, size_t cap) {
    if (n == 0) return 0;
    if (n >= 8 && s[0] == 'V' && s[1] == 'A' && s[2] == 0x01) {
        if (s[3] & 0x40) return vans_container(s, n, dst, cap, 1);
        if (s[3] == 0) return vans_container(s, n, dst, cap, 0);
    }
    if (s[0] == 0x09) {
        if (n - 1 > cap) return PRZM_DEC_ERR_CAPACITY;
        memcpy(dst, s + 1, n - 1);
        return (int64_t)(n - 1);
    }
    if (s[0] == 0x00) {
        if (n < 5) return PRZM_DEC_ERR_CORRUPT;
        const uint32_t l = le32(s + 1);
        if (l > cap || 5 + (size_t)l > n) return PRZM_DEC_ERR_CORRUPT;
        memcpy(dst, s + 5, l);
        return l;
    }
    if (n > cap) return PRZM_DEC_ERR_CAPACITY;
    memcpy(dst, s, n);
    return (int64_t)n;
}

/* copies one LEB128 value (its bytes) from v[*vp] to d[*dp] */
static int copy_leb(const uint8_t *v, size_t vn, size_t *vp, uint8_t *d, size_t dn, size_t *dp) {
    for (int k = 0; k < 4; k++) {
        if (*vp >= vn || *dp >= dn) return PRZM_DEC_ERR_CORRUPT;
        const uint8_t c = v[(*vp)++];
        d[(*dp)++] = c;
        if (!(c & 0x80)) return 0;
    }
    return PRZM_DEC_ERR_CORRUPT;
}

/* the command stream from op / meta / var (commands.go:205-258 for the opcodes read here) */
static int join(const uint8_t *op, size_t no, const uint8_t *meta, size_t nm, const uint8_t *var, size_t nv,
                uint8_t *dst, size_t want) {
    size_t i = 0, mp = 0, vp = 0, dp = 0;
    int rc;
    while (i < no && dp < want) {
        const uint8_t c = op[i++];
        dst[dp++] = c;
        if (c < 0x80) continue;
        if (c < 0xC0) {
            if ((c & 15) == 15 && (rc = copy_leb(var, nv, &vp, dst, want, &dp))) return rc;
            continue;
        }
        if (c <= 0xC3) {
            if (mp >= nm || dp >= want) return PRZM_DEC_ERR_CORRUPT;
            const uint8_t tok = meta[mp++];
            dst[dp++] = tok;
            if ((tok & 15) == 15 && (rc = copy_leb(var, nv, &vp, dst, want, &dp))) return rc;
            continue;
        }
        if (c == 0xF0 || c == 0xF1) {
            if ((rc = copy_leb(var, nv, &vp, dst, want, &dp))) return rc;
            continue;
        }
        if (c >= 0xF3 && c <= 0xFA) continue;
        return PRZM_DEC_ERR_UNSUPPORTED;
    }
    return dp == want ? 0 : PRZM_DEC_ERR_CORRUPT;
}

/* Section 2 -> the command stream (commands.go:755-789): *out points into s2 or buf */
static int section2(const uint8_t *s2, size_t n, uint8_t **buf, const uint8_t **out, size_t *on) {
    *out = NULL;
    *on = 0;
    if (n == 0) return 0;
    if (s2[0] == 0x09) { *out = s2 + 1; *on = n - 1; return 0; }
    if (s2[0] == 0x00) {
        if (n < 5) return PRZM_DEC_ERR_CORRUPT;
        const uint32_t l = le32(s2 + 1);
        if (5 + (size_t)l > n) return PRZM_DEC_ERR_CORRUPT;
        *out = s2 + 5;
        *on = l;
        return 0;
    }
    if (s2[0] != 0x08) { *out = s2; *on = n; return 0; }
    /* Mode 8: 0x08, u32 command bytes, u32 op section bytes, u32 meta section bytes, then the
     * three substreams (commands.go:409-444) */
    if (n < 13) return PRZM_DEC_ERR_CORRUPT;
    const uint32_t orig = le32(s2 + 1), lop = le32(s2 + 5), lme = le32(s2 + 9);
    if (orig > 4 * MAX_FRAME || (size_t)lop > n - 13 || (size_t)lme > n - 13 - lop) return PRZM_DEC_ERR_CORRUPT;
    const size_t lva = n - 13 - lop - lme;
    /* each substream holds at most the whole command stream */
    uint8_t *mem = (uint8_t *)malloc(4 * (size_t)orig + 16);
    if (!mem) return PRZM_DEC_ERR_CAPACITY;
    uint8_t *op = mem, *me = mem + orig + 4, *va = mem + 2 * (size_t)orig + 8, *cmd = mem + 3 * (size_t)orig + 12;
    const uint8_t *p = s2 + 13;
    int64_t no = substream(p, lop, op, orig);
    int64_t nm = no < 0 ? no : substream(p + lop, lme, me, orig);
    int64_t nv = nm < 0 ? nm : substream(p + lop + lme, lva, va, orig);
    int rc = no < 0 ? (int)no : nm < 0 ? (int)nm : nv < 0 ? (int)nv : 0;
    if (rc == PRZM_DEC_ERR_CAPACITY) rc = PRZM_DEC_ERR_CORRUPT;
    if (!rc) rc = join(op, (size_t)no, me, (size_t)nm, va, (size_t)nv, cmd, orig);
    if (rc) { free(mem); return rc; }
    *buf = mem;
    *out = cmd;
    *on = orig;
    return 0;
}

/* distance codes coded by one rANS state whose table is chosen by the class of the previous
 * code ([0x0C][u32 state][stream], distance.go:287-321); like the V-ANS containers, the
 * stream must end on the encoder's initial state with every byte used */
static int dist_codes_m12(const uint8_t *s, size_t sn, uint8_t *dst, uint32_t n) {
    if (sn < 5 || s[0] != 0x0C) return PRZM_DEC_ERR_CORRUPT;
    uint16_t cdf[4][64];
    uint8_t *slot = (uint8_t *)malloc(4 * 4096);
    if (!slot) return PRZM_DEC_ERR_CAPACITY;
    for (int c = 0; c < 4; c++) {
        uint32_t cum = 0;
        for (int k = 0; k < 64; k++) {
            cdf[c][k] = (uint16_t)cum;
            memset(slot + 4096 * c + cum, k, przm_dist_freq[c][k]);
            cum += przm_dist_freq[c][k];
        }
    }
    uint32_t x = le32(s + 1);
    size_t sp = 5;
    uint8_t prev = 0;
    int rc = x < 65536u ? PRZM_DEC_ERR_CORRUPT : 0;
    for (uint32_t i = 0; i < n && !rc; i++) {
        const int c = prev < 6 ? 0 : prev < 14 ? 1 : prev < 26 ? 2 : 3;
        const uint32_t sl = x & 4095u;
        const uint8_t sym = slot[4096 * c + sl];
        dst[i] = sym;
        prev = sym;
        x = (uint32_t)przm_dist_freq[c][sym] * (x >> 12) + sl - cdf[c][sym];
        while (x < 65536u && !rc) {
            if (sp >= sn) rc = PRZM_DEC_ERR_CORRUPT;
            else x = (x << 8) | s[sp++];
        }
    }
    free(slot);
    if (!rc && (sp != sn || x != 65536u)) rc = PRZM_DEC_ERR_CORRUPT;
    return rc;
}

/* a distance-prefix container (0x06 / 0x0A / 0x0B: distance.go:563-690, src/przm_dec.c:1057-1137)
 * -> its LEB128 varints in a new buffer *buf */
static int prefix_offsets(const uint8_t *s3, size_t n, uint8_t **buf, size_t *on) {
    uint32_t orig, count, cl;
    size_t h;
    if (s3[0] == 0x06) {
        if (n < 13) return PRZM_DEC_ERR_CORRUPT;
        orig = le32(s3 + 1); count = le32(s3 + 5); cl = le32(s3 + 9); h = 13;
    } else if (s3[0] == 0x0A) {
        if (n < 7) return PRZM_DEC_ERR_CORRUPT;
        orig = (uint32_t)s3[1] | (uint32_t)s3[2] << 8;
        count = (uint32_t)s3[3] | (uint32_t)s3[4] << 8;
        cl = (uint32_t)s3[5] | (uint32_t)s3[6] << 8;
        h = 7;
    } else {
        if (n < 4) return PRZM_DEC_ERR_CORRUPT;
        orig = s3[1]; count = s3[2]; cl = s3[3]; h = 4;
    }
    if (h + (size_t)cl > n || orig > MAX_FRAME || count > orig) return PRZM_DEC_ERR_CORRUPT;
    if (count == 0) return orig ? PRZM_DEC_ERR_CORRUPT : 0;
    const uint8_t *cs = s3 + h, *ex = s3 + h + cl;
    const size_t en = n - h - cl;
    uint8_t *codes = (uint8_t *)malloc((size_t)count + 16);
    uint8_t *out = (uint8_t *)malloc((size_t)orig + 16);
    int rc = codes && out ? 0 : PRZM_DEC_ERR_CAPACITY;
    /* the code stream (distance.go:600-631): raw when exactly count bytes, else Mode 12, an
     * adaptive V-ANS container or a flat stream */
    if (rc) {
    } else if (cl == count) {
        memcpy(codes, cs, count);
    } else if (cl > 0 && cs[0] == 0x0C) {
        rc = dist_codes_m12(cs, cl, codes, count);
    } else if (cl >= 8 && cs[0] == 'V' && cs[1] == 'A' && cs[2] == 0x01) {
        if (cs[3] != 0) rc = PRZM_DEC_ERR_UNSUPPORTED;
        else if (le32(cs + 4) != count) rc = PRZM_DEC_ERR_CORRUPT;
        else {
            const int64_t r = vans_container(cs, cl, codes, count, 0);
            rc = r < 0 ? (int)r : 0;
        }
    } else if (cl > 0 && cs[0] == 0x09) {
        if (cl - 1 < count) rc = PRZM_DEC_ERR_CORRUPT;
        else memcpy(codes, cs + 1, count);
    } else if (cl >= 5 && cs[0] == 0x00) {
        const uint32_t l = le32(cs + 1);
        if (l < count || 5 + (size_t)l > cl) rc = PRZM_DEC_ERR_CORRUPT;
        else memcpy(codes, cs + 5, count);
    } else {
        rc = PRZM_DEC_ERR_CORRUPT;
    }
    /* extra bits, most significant first (distance.go:262-285), and the varints */
    size_t ep = 0, op = 0;
    uint64_t acc = 0;
    uint32_t bits = 0;
    for (uint32_t i = 0; i < count && !rc; i++) {
        const uint32_t c = codes[i];
        if (c > 48) { rc = PRZM_DEC_ERR_CORRUPT; break; }
        const uint32_t nb = c < 4 ? 0 : c == 48 ? 24 : (c - 2) >> 1;
        const uint32_t base = c < 4 ? c + 1 : c == 48 ? 16777217u : ((2u + (c & 1u)) << nb) + 1u;
        while (bits < nb) {
            if (ep >= en) { rc = PRZM_DEC_ERR_CORRUPT; break; }
            acc = acc << 8 | ex[ep++];
            bits += 8;
        }
        if (rc) break;
        bits -= nb;
        uint32_t d = base + (nb ? (uint32_t)(acc >> bits) & ((1u << nb) - 1u) : 0u);
        if (d > MAX_FRAME) { rc = PRZM_DEC_ERR_RANGE; break; }
        do {
            if (op >= orig) { rc = PRZM_DEC_ERR_CORRUPT; break; }
            out[op++] = (uint8_t)(d >= 0x80 ? (d & 0x7Fu) | 0x80u : d);
            d >>= 7;
        } while (d);
    }
    if (!rc && op != orig) rc = PRZM_DEC_ERR_CORRUPT;
    free(codes);
    if (rc) { free(out); return rc; }
    *buf = out;
    *on = orig;
    return 0;
}

/* Section 3 -> the LEB128 offset stream (distance.go:525-545): *out points into s3 or *buf */
static int section3(const uint8_t *s3, size_t n, uint8_t **buf, const uint8_t **out, size_t *on) {
    *out = NULL;
    *on = 0;
    if (n == 0) return 0;
    if (s3[0] == 0x09) { *out = s3 + 1; *on = n - 1; return 0; }
    if (s3[0] == 0x00) {
        if (n < 5) return PRZM_DEC_ERR_CORRUPT;
        const uint32_t l = le32(s3 + 1);
        if (5 + (size_t)l > n) return PRZM_DEC_ERR_CORRUPT;
        *out = s3 + 5;
        *on = l;
        return 0;
    }
    if (s3[0] == 0x06 || s3[0] == 0x0A || s3[0] == 0x0B) {
        const int rc = prefix_offsets(s3, n, buf, on);
        *out = *buf;
        return rc;
    }
    /* any other mode byte is ErrUnsupportedMode in the reference too */
    return PRZM_DEC_ERR_UNSUPPORTED;
}

/* Section 4 -> the literal stream (literals.go:696-736 for the containers read here) */
static int section4(const uint8_t *s4, size_t n, uint8_t **buf, const uint8_t **out, size_t *on) {
    *out = s4;
    *on = n;
    if (n >= 8 && s4[0] == 'V' && s4[1] == 'A' && s4[2] == 0x01) {
        const uint32_t l = le32(s4 + 4);
        if (l > MAX_FRAME) return PRZM_DEC_ERR_CORRUPT;
        uint8_t *m = (uint8_t *)malloc((size_t)l + 1);
        if (!m) return PRZM_DEC_ERR_CAPACITY;
        /* the flag byte is not read past the sheaf bit: every other container is adaptive V-ANS */
        const int64_t r = (s4[3] & 0x04) ? sheaf_container(s4, n, m, l) : vans_container(s4, n, m, l, 0);
        if (r < 0) { free(m); return (int)r; }
        *buf = m;
        *out = m;
        *on = l;
        return 0;
    }
    if (n >= 8 && le32(s4 + 4) == (le32(s4) + 1) / 2) return PRZM_DEC_ERR_UNSUPPORTED; /* nibble planes */
    return 0;
}

/* ---------------- frame ---------------- */

int64_t przm_decode_size(const uint8_t *src, size_t n) {
    if (!src || n < 8) return PRZM_DEC_ERR_CORRUPT;
    if (le32(src) != MAGIC) return PRZM_DEC_ERR_MAGIC;
    return (int64_t)((uint32_t)src[5] | (uint32_t)src[6] << 8 | (uint32_t)src[7] << 16);
}

/* the command / offset / literal machine (docs/PRZM_SPEC.md:468-526) */
static int64_t execute(const uint8_t *cmd, size_t nc, const uint8_t *off, size_t nf, const uint8_t *lit, size_t nl,
                       uint32_t bmin, uint8_t *dst, uint32_t u) {
    uint32_t r[3] = {1, 4, 8};
    size_t cp = 0, op = 0, lp = 0, dp = 0;
    int rc;
#define LITS(L)                                                        \
    do {                                                               \
        const size_t L_ = (L);                                         \
        if (lp + L_ > nl) return PRZM_DEC_ERR_CORRUPT;                 \
        if (dp + L_ > u) return PRZM_DEC_ERR_SIZE;                     \
        memcpy(dst + dp, lit + lp, L_);                                \
        lp += L_;                                                      \
        dp += L_;                                                      \
    } while (0)
    while (cp < nc) {
        const uint8_t c = cmd[cp++];
        uint32_t L = 0, M = 0, d = 0, v = 0;
        int k = -1; /* register index, or -1 for an explicit distance */
        if (c < 0x80) {
            L = c >> 4 & 7u;
            M = (c >> 2 & 3u) + bmin;
            k = c & 3;
        } else if (c < 0xC0) {
            L = c >> 4 & 3u;
            M = (c & 15u) + bmin;
            if ((c & 15u) == 15u) {
                if ((rc = uleb(cmd, nc, &cp, &v))) return rc;
                M += v;
            }
            if ((rc = uleb(off, nf, &op, &d))) return rc;
        } else if (c <= 0xC3) {
            if (cp >= nc) return PRZM_DEC_ERR_CORRUPT;
            const uint8_t t = cmd[cp++];
            L = t >> 4;
            M = (t & 15u) + bmin;
            if ((t & 15u) == 15u) {
                if ((rc = uleb(cmd, nc, &cp, &v))) return rc;
                M += v;
            }
            k = c & 3;
        } else if (c == 0xF0) {
            if ((rc = uleb(cmd, nc, &cp, &v))) return rc;
            LITS(v);
            continue;
        } else if (c == 0xF1) {
            if ((rc = uleb(cmd, nc, &cp, &v))) return rc;
            if (lp >= nl) return PRZM_DEC_ERR_CORRUPT;
            if (dp + v > u) return PRZM_DEC_ERR_SIZE;
            memset(dst + dp, lit[lp++], v);
            dp += v;
            continue;
        } else if (c >= 0xF3 && c <= 0xFA) {
            LITS(c - 0xF2u);
            continue;
        } else {
            return PRZM_DEC_ERR_UNSUPPORTED;
        }
        if (k == 3) return PRZM_DEC_ERR_UNSUPPORTED; /* only three repeat registers are read here */
        LITS(L);
        /* the register move of spec 17.1: repeat 1 swaps r0 and r1, repeat 2 rotates, an
         * explicit distance shifts the ring and enters at r0 */
        if (k == 0) d = r[0];
        else if (k == 1) { d = r[1]; r[1] = r[0]; r[0] = d; }
        else if (k == 2) { d = r[2]; r[2] = r[1]; r[1] = r[0]; r[0] = d; }
        else { r[2] = r[1]; r[1] = r[0]; r[0] = d; }
        if (d == 0 || d > dp) return PRZM_DEC_ERR_RANGE;
        if (dp + M > u) return PRZM_DEC_ERR_SIZE;
        for (uint32_t i = 0; i < M; i++, dp++) dst[dp] = dst[dp - d]; /* overlapping copies replicate forward */
    }
#undef LITS
    return dp == u ? (int64_t)u : PRZM_DEC_ERR_SIZE;
}

int64_t przm_decode_frame(const uint8_t *src, size_t n, uint8_t *dst, size_t cap) {
    const int64_t us = przm_decode_size(src, n);
    if (us < 0) return us;
    const uint32_t u = (uint32_t)us;
    const uint8_t flags = src[4];
    if (u > cap) return PRZM_DEC_ERR_CAPACITY;
    if (u && !dst) return PRZM_DEC_ERR_CAPACITY;
    if (u == 0) return 0;
    if (flags & F_RAW) {
        if (n < 8 + (size_t)u) return PRZM_DEC_ERR_CORRUPT;
        memcpy(dst, src + 8, u);
        return u;
    }
    if (flags & (F_STRIDE | F_DELTA | F_DICT | F_NORM)) return PRZM_DEC_ERR_UNSUPPORTED;
    /* directory (decompress.go:215-233) and Section 1 (8 + 2 * byte 5 bytes, decompress.go:237-246) */
    size_t p = 8, lc, lo;
    if (flags & F_EXT) {
        if (p + 6 > n) return PRZM_DEC_ERR_CORRUPT;
        lc = (size_t)src[p] | (size_t)src[p + 1] << 8 | (size_t)src[p + 2] << 16;
        lo = (size_t)src[p + 3] | (size_t)src[p + 4] << 8 | (size_t)src[p + 5] << 16;
        p += 6;
    } else {
        if (p + 4 > n) return PRZM_DEC_ERR_CORRUPT;
        lc = (size_t)src[p] | (size_t)src[p + 1] << 8;
        lo = (size_t)src[p + 2] | (size_t)src[p + 3] << 8;
        p += 4;
    }
    const size_t l1 = p + 8 <= n ? 8 + 2 * (size_t)src[p + 5] : 0;
    const size_t o2 = p + l1, o3 = o2 + lc, o4 = o3 + lo;
    if (o4 > n) return PRZM_DEC_ERR_CORRUPT;
    uint8_t *b2 = NULL, *b3 = NULL, *b4 = NULL;
    const uint8_t *cmd, *off, *lit;
    size_t nc, nf, nl;
    int64_t rc = section2(src + o2, lc, &b2, &cmd, &nc);
    if (!rc) rc = section3(src + o3, lo, &b3, &off, &nf);
    if (!rc) rc = section4(src + o4, n - o4, &b4, &lit, &nl);
    if (!rc) rc = execute(cmd, nc, off, nf, lit, nl, (flags & F_MIN3) ? 3u : 4u, dst, u);
    free(b2);
    free(b3);
    free(b4);
    return rc;
}

int64_t przm_decode_frames(const uint8_t *src, size_t n, const uint32_t *sizes, size_t nframes, uint8_t *dst,
                           size_t cap) {
    if ((!src && n) || (!sizes && nframes)) return PRZM_DEC_ERR_CORRUPT;
    size_t sp = 0, dp = 0;
    for (size_t i = 0; i < nframes; i++) {
        if (sizes[i] > n - sp) return PRZM_DEC_ERR_CORRUPT;
        const int64_t r = przm_decode_frame(src + sp, sizes[i], dst ? dst + dp : NULL, cap - dp);
        if (r < 0) return r;
        sp += sizes[i];
        dp += (size_t)r;
    }
    return sp == n ? (int64_t)dp : PRZM_DEC_ERR_CORRUPT;
}

/* ---------------- PRZS stream ---------------- */

int64_t przm_decode_stream_size(const uint8_t *src, size_t n) {
    if (!src || n < 8) return PRZM_DEC_ERR_CORRUPT;
    if (le32(src) == MAGIC) return przm_decode_size(src, n);
    if (le32(src) != MAGIC_S) return PRZM_DEC_ERR_MAGIC;
    if (n < STREAM_HDR) return PRZM_DEC_ERR_CORRUPT; /* ParseStreamHeader, decompress.go:851-853 */
    const uint64_t total = (uint64_t)le32(src + 9) | (uint64_t)le32(src + 13) << 32;
    return total > (uint64_t)INT64_MAX ? PRZM_DEC_ERR_CORRUPT : (int64_t)total;
}

int64_t przm_decode_stream(const uint8_t *src, size_t n, uint8_t *dst, size_t cap) {
    if (!src || n < 8) return PRZM_DEC_ERR_CORRUPT;
    if (le32(src) == MAGIC) return przm_decode_frame(src, n, dst, cap);
    const int64_t total = przm_decode_stream_size(src, n);
    if (total < 0) return total;
    if (src[4] != 0) return PRZM_DEC_ERR_UNSUPPORTED; /* no stream flag is defined */
    if ((uint64_t)total > cap) return PRZM_DEC_ERR_CAPACITY;
    const uint32_t count = le32(src + 5);
    /* the chunks: u32 length, then that many bytes of one frame (cmd/przm-viz/analyzer.go:242-256) */
    size_t sp = STREAM_HDR, dp = 0;
    for (uint32_t i = 0; i < count; i++) {
        if (n - sp < 4) return PRZM_DEC_ERR_CORRUPT;
        const uint32_t len = le32(src + sp);
        sp += 4;
        if (len > n - sp) return PRZM_DEC_ERR_CORRUPT;
        const int64_t r = przm_decode_frame(src + sp, len, dst ? dst + dp : NULL, cap - dp);
        if (r < 0) return r;
        sp += len;
        dp += (size_t)r;
    }
    if (sp != n || dp != (uint64_t)total) return PRZM_DEC_ERR_SIZE;
    return (int64_t)dp;
}
