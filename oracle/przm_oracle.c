/*
 * przm_oracle.c -- scalar CPU restatement of the PRZM GPU-encode mode.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product (przm_amd/) never links, imports or calls it, and the product
 * path fails loudly when its HIP library is missing instead of falling back
 * here.
 *
 * What it restates ("mode G": the frame format of the reference CPU path,
 * produced by the reference GPU tokenizer's parse rules over 64 KiB blocks):
 *
 *   hash4 / search depth      compress.go:204-206, compress.go:34-51
 *   hash-chain candidates     compress.go:681-707 (walk), 857-862 (insert),
 *                             374-387 (match length), 1066-1085 (reset)
 *   best-candidate selection  longest length, ties keep the first (nearest)
 *                             candidate; length capped at min(n-p, 1024)
 *                             (compress.go:28, 699-700)
 *   tokenizer                 gpu/kernel/przm_tokenizer.cuh:115-466 for one
 *                             tile == the block, registers {1,4,8} (:162-164)
 *   command split             core/commands/commands.go:18-46, 174-202
 *   NormalizeFrequencies      core/literals/literals.go:78-140
 *   joint nibble matrix       core/literals/literals.go:143-184
 *   V-ANS 4x encoder          core/literals/literals.go:289-341, 368-384
 *   Meta8 encoder             core/commands/vans_meta.go:45-150
 *   substream / Mode 8        core/commands/commands.go:371-444, 537-568
 *   Section 3 flat LEB128     core/distance/distance.go:825-849 (flat branch)
 *   Section 3 prefix option   core/distance/distance.go:165-202, 217-247,
 *                             323-376, 692-849 (PackageOffsetsSection)
 *   Section 4                 core/literals/literals.go:664-693 without the
 *                             sheaf trials
 *   Section 4 sheaf option    core/literals/literals.go:344-355, 601-693,
 *                             sheaf.go:20-152 (PackageLiteralsSection)
 *   frame assembly            compress.go:113-132, 183-202, 962-1033
 *
 * Mode-G rule added on top of the reference: a raw Section 4 that either
 * reference decoder would misread as a container (it starts 'V','A',0x01, or
 * its first two u32 satisfy the nibble-plane sniff, literals.go:703-735,
 * src/przm_dec.c:854-878) makes the frame an uncompressed frame instead.
 *
 * Parity pins (tests/): the reference C99 decoder (oracle/_ref) decodes every
 * frame and every V-ANS / Meta8 payload produced here; brute-force chain
 * enumeration on small inputs; closed-form hash and histogram checks;
 * normalisation invariants and special cases; the reference split KAT.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../przm_amd/csrc/przm_dist_freq.h"
#include "../przm_amd/csrc/przm_meta8_freq.h"

#define ORC_BLOCK 65536u
#define ORC_MAX_MATCH 1024u
#define ORC_EXT_CAP 65530u
#define ORC_RLE_CAP 65535u

static inline uint32_t rd32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static inline void wr32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24);
}

/* compress.go:204-206 */
uint32_t orc_hash4(const uint8_t *p) { return (rd32(p) * 0x1E35A7BDu) >> 16; }

/* compress.go:34-51 */
int orc_search_depth(int level) {
    if (level <= 0) level = 6; /* compress.go:1038-1040 */
    if (level <= 1) return 2;
    if (level <= 3) return 4;
    if (level <= 5) return 8;
    if (level <= 7) return 16;
    if (level <= 9) return 32;
    if (level <= 10) return 64;
    return 128;
}

/* compress.go:374-387: plain byte loop, same result as the SWAR form */
static uint32_t lcp(const uint8_t *a, const uint8_t *b, uint32_t cap) {
    uint32_t l = 0;
    while (l < cap && a[l] == b[l]) l++;
    return l;
}

/*
 * Per-position best match of one block.  desc[p] = (len << 16) | dist with
 * len >= 4, or 0.  Candidates come from the hash-4 chain of positions
 * inserted before p (every q with q + 4 <= n), walked newest first for
 * `depth` steps (compress.go:681-707; insertion after the search,
 * compress.go:921-923).  stats (optional): [0] candidates visited.
 */
void orc_match_block(const uint8_t *src, uint32_t n, int depth, uint32_t *desc, uint64_t *stats) {
    int32_t *head = (int32_t *)malloc(65536 * sizeof(int32_t));
    int32_t *prev = (int32_t *)malloc((n ? n : 1) * sizeof(int32_t));
    uint64_t visited = 0;
    for (uint32_t i = 0; i < 65536; i++) head[i] = -1;
    for (uint32_t p = 0; p < n; p++) {
        desc[p] = 0;
        if (p + 4 > n) continue;
        uint32_t h = orc_hash4(src + p);
        uint32_t cap = n - p < ORC_MAX_MATCH ? n - p : ORC_MAX_MATCH;
        uint32_t best_l = 0, best_d = 0;
        int32_t cand = head[h];
        for (int k = 0; cand >= 0 && k < depth; k++) {
            uint32_t d = p - (uint32_t)cand;
            uint32_t l = lcp(src + p, src + cand, cap);
            if (l > best_l) { best_l = l; best_d = d; }
            cand = prev[cand];
            visited++;
        }
        if (best_l >= 4) desc[p] = (best_l << 16) | best_d;
        prev[p] = head[h];
        head[h] = (int32_t)p;
    }
    if (stats) stats[0] += visited;
    free(head);
    free(prev);
}

/* ------------------------------------------------------------------ */
/* Tokenizer: przm_tokenizer.cuh:115-466 restated for one tile.        */

typedef struct {
    uint8_t *cmd; uint32_t ncmd;
    uint8_t *off; uint32_t noff;
    uint8_t *lit; uint32_t nlit;
    uint8_t *op;  uint32_t nop;
    uint8_t *meta; uint32_t nmeta;
    uint8_t *var; uint32_t nvar;
    uint32_t ntok; /* match + RLE tokens */
} orc_streams;

static void put_leb(uint8_t *b, uint32_t *n, uint32_t v) { /* przm_tokenizer.cuh:33-55 */
    while (v >= 0x80) { b[(*n)++] = (uint8_t)(v | 0x80); v >>= 7; }
    b[(*n)++] = (uint8_t)v;
}

/* przm_tokenizer.cuh:57-76, with the split of commands.go:191-202 */
static void emit_lits(orc_streams *s, const uint8_t *src, uint32_t from, uint32_t count) {
    if (count == 0) return;
    if (count <= 8) {
        s->cmd[s->ncmd++] = (uint8_t)(0xF2 + count);
        s->op[s->nop++] = (uint8_t)(0xF2 + count);
    } else {
        s->cmd[s->ncmd++] = 0xF0;
        s->op[s->nop++] = 0xF0;
        put_leb(s->cmd, &s->ncmd, count);
        put_leb(s->var, &s->nvar, count);
    }
    memcpy(s->lit + s->nlit, src + from, count);
    s->nlit += count;
}

void orc_tokenize_block(const uint8_t *src, uint32_t n, const uint32_t *desc, orc_streams *s) {
    uint32_t pos = 0, lit_start = 0;
    int rep0 = 1, rep1 = 4, rep2 = 8; /* przm_tokenizer.cuh:162-164, tile 0 */
    s->ncmd = s->noff = s->nlit = s->nop = s->nmeta = s->nvar = s->ntok = 0;
    while (pos < n) {
        uint32_t mp = pos;
        while (mp < n && desc[mp] == 0) mp++;
        if (mp >= n) break;
        uint32_t mlen = desc[mp] >> 16, d = desc[mp] & 0xFFFF;
        /* lookahead, :196-216 */
        if (mp + 1 < n && desc[mp + 1] != 0 && (desc[mp + 1] >> 16) > mlen + 1) { pos = mp + 1; continue; }
        /* run-length, :218-286 */
        if (mp + 12 <= n) {
            uint32_t w4 = rd32(src + mp);
            uint8_t b0 = (uint8_t)w4;
            if (w4 == (uint32_t)b0 * 0x01010101u) {
                uint32_t max_r = n - mp < ORC_RLE_CAP ? n - mp : ORC_RLE_CAP;
                uint32_t r = 4;
                while (r < max_r && src[mp + r] == b0) r++;
                if (r >= 12 && r >= mlen) {
                    emit_lits(s, src, lit_start, mp - lit_start);
                    s->cmd[s->ncmd++] = 0xF1;
                    s->op[s->nop++] = 0xF1;
                    put_leb(s->cmd, &s->ncmd, r);
                    put_leb(s->var, &s->nvar, r);
                    s->lit[s->nlit++] = b0;
                    s->ntok++;
                    pos = mp + r;
                    lit_start = pos;
                    continue;
                }
            }
        }
        /* validate + extend, :288-320 */
        if (!(d > 0 && d <= mp && mlen >= 4 && rd32(src + mp) == rd32(src + mp - d))) { pos = mp + 1; continue; }
        if (mp + mlen > n) mlen = n - mp;
        while (mp + mlen < n && mlen < ORC_EXT_CAP && src[mp + mlen] == src[mp - d + mlen]) mlen++;
        /* emission, :328-427 */
        uint32_t lit_len = mp - lit_start;
        int k = -1;
        if ((int)d == rep0 && rep0 > 0) k = 0;
        else if ((int)d == rep1 && rep1 > 0) k = 1;
        else if ((int)d == rep2 && rep2 > 0) k = 2;
        if (k >= 0) {
            if (mlen <= 7) {
                if (lit_len > 7) { emit_lits(s, src, lit_start, lit_len - 7); lit_start = mp - 7; lit_len = 7; }
                uint8_t op = (uint8_t)((lit_len << 4) | ((mlen - 4) << 2) | (uint32_t)k);
                s->cmd[s->ncmd++] = op;
                s->op[s->nop++] = op;
            } else {
                if (lit_len > 15) { emit_lits(s, src, lit_start, lit_len - 15); lit_start = mp - 15; lit_len = 15; }
                uint32_t base = mlen - 4 > 14 ? 15 : mlen - 4;
                uint8_t tok = (uint8_t)((lit_len << 4) | base);
                s->cmd[s->ncmd++] = (uint8_t)(0xC0 | k);
                s->cmd[s->ncmd++] = tok;
                s->op[s->nop++] = (uint8_t)(0xC0 | k);
                s->meta[s->nmeta++] = tok;
                if (base == 15) { put_leb(s->cmd, &s->ncmd, mlen - 19); put_leb(s->var, &s->nvar, mlen - 19); }
            }
            memcpy(s->lit + s->nlit, src + lit_start, lit_len);
            s->nlit += lit_len;
            if (k == 1) { rep1 = rep0; rep0 = (int)d; }
            else if (k == 2) { rep2 = rep1; rep1 = rep0; rep0 = (int)d; }
        } else {
            if (lit_len > 3) { emit_lits(s, src, lit_start, lit_len - 3); lit_start = mp - 3; lit_len = 3; }
            uint32_t base = mlen - 4 > 14 ? 15 : mlen - 4;
            uint8_t op = (uint8_t)(0x80 | (lit_len << 4) | base);
            s->cmd[s->ncmd++] = op;
            s->op[s->nop++] = op;
            if (base == 15) { put_leb(s->cmd, &s->ncmd, mlen - 19); put_leb(s->var, &s->nvar, mlen - 19); }
            memcpy(s->lit + s->nlit, src + lit_start, lit_len);
            s->nlit += lit_len;
            put_leb(s->off, &s->noff, d);
            rep2 = rep1; rep1 = rep0; rep0 = (int)d;
        }
        s->ntok++;
        pos = mp + mlen;
        lit_start = pos;
    }
    emit_lits(s, src, lit_start, n - lit_start); /* :440-457 */
}

/* commands.go:18-46 with the mode-G opcode families (commands.go:76-202).
 * Returns 0, or -1 on an opcode mode G never emits / a truncated stream. */
static int leb_copy(const uint8_t *src, uint32_t n, uint32_t *sp, uint8_t *v, uint32_t *nv) {
    for (int i = 0; i < 4; i++) { /* commands.go:205-221 */
        if (*sp >= n) return -1;
        uint8_t b = src[(*sp)++];
        v[(*nv)++] = b;
        if (!(b & 0x80)) return 0;
    }
    return -1;
}

int orc_split_commands(const uint8_t *cmd, uint32_t n, uint8_t *op, uint32_t *nop, uint8_t *meta, uint32_t *nmeta,
                       uint8_t *var, uint32_t *nvar) {
    uint32_t sp = 0;
    *nop = *nmeta = *nvar = 0;
    while (sp < n) {
        uint8_t c = cmd[sp++];
        op[(*nop)++] = c;
        if (c < 0x80) continue;
        if (c < 0xC0) {
            if ((c & 15) == 15 && leb_copy(cmd, n, &sp, var, nvar)) return -1;
            continue;
        }
        if (c <= 0xC3) {
            if (sp >= n) return -1;
            uint8_t tok = cmd[sp++];
            meta[(*nmeta)++] = tok;
            if ((tok & 15) == 15 && leb_copy(cmd, n, &sp, var, nvar)) return -1;
            continue;
        }
        if (c == 0xF0 || c == 0xF1) {
            if (leb_copy(cmd, n, &sp, var, nvar)) return -1;
            continue;
        }
        if (c >= 0xF3 && c <= 0xFA) continue;
        return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Entropy coding.                                                      */

void orc_histogram(const uint8_t *src, uint32_t n, uint32_t counts[256]) {
    memset(counts, 0, 256 * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) counts[src[i]]++;
}

/* literals.go:78-140 */
void orc_normalize(const uint32_t counts[256], uint16_t freq[256]) {
    uint64_t total = 0;
    int unique = 0;
    memset(freq, 0, 256 * sizeof(uint16_t));
    for (int i = 0; i < 256; i++) if (counts[i]) { total += counts[i]; unique++; }
    if (!unique) return;
    uint64_t remaining = (uint64_t)(4096 - unique);
    uint32_t sum = 0;
    for (int i = 0; i < 256; i++) {
        if (!counts[i]) continue;
        uint32_t f = (uint32_t)(1 + ((uint64_t)counts[i] * remaining) / total);
        freq[i] = (uint16_t)f;
        sum += f;
    }
    int diff = 4096 - (int)sum;
    int64_t c[256];
    for (int i = 0; i < 256; i++) c[i] = counts[i];
    if (diff > 0) {
        for (int r = 0; r < diff; r++) {
            int best = -1;
            int64_t bc = -1;
            for (int s2 = 0; s2 < 256; s2++)
                if (c[s2] > bc && freq[s2] > 0) { bc = c[s2]; best = s2; }
            if (best >= 0) { freq[best]++; c[best] = -1; }
        }
    } else if (diff < 0) {
        for (int r = 0; r < -diff; r++)
            for (int s2 = 255; s2 >= 0; s2--)
                if (freq[s2] > 1) { freq[s2]--; break; }
    }
}

/* literals.go:143-184; returns bytes written */
uint32_t orc_joint_nibble(const uint16_t freq[256], uint8_t *out) {
    uint32_t n = 0;
    uint16_t rows = 0;
    for (int h = 0; h < 16; h++)
        for (int l = 0; l < 16; l++)
            if (freq[(h << 4) | l]) { rows |= (uint16_t)(1u << h); break; }
    out[n++] = (uint8_t)rows; out[n++] = (uint8_t)(rows >> 8);
    for (int h = 0; h < 16; h++) {
        if (!(rows & (1u << h))) continue;
        uint16_t cols = 0;
        for (int l = 0; l < 16; l++) if (freq[(h << 4) | l]) cols |= (uint16_t)(1u << l);
        out[n++] = (uint8_t)cols; out[n++] = (uint8_t)(cols >> 8);
        for (int l = 0; l < 16; l++) {
            if (!(cols & (1u << l))) continue;
            uint16_t w = freq[(h << 4) | l];
            if (w < 128) out[n++] = (uint8_t)w;
            else { out[n++] = (uint8_t)((w >> 8) | 0x80); out[n++] = (uint8_t)w; }
        }
    }
    return n;
}

/* literals.go:289-341 (container with flag byte 0x00, literals.go:297).
 * out must hold 8 + 546 + 16 + 2n bytes.  Returns the container length. */
uint32_t orc_vans4x_encode(const uint8_t *src, uint32_t n, const uint16_t freq[256], uint8_t *out) {
    uint16_t cdf[256];
    uint32_t cum = 0;
    for (int s = 0; s < 256; s++) { cdf[s] = (uint16_t)cum; cum += freq[s]; }
    uint8_t *stream = (uint8_t *)malloc(2 * (size_t)n + 16);
    uint32_t ns = 0;
    uint32_t x[4] = {65536, 65536, 65536, 65536};
    for (uint32_t ii = n; ii-- > 0;) {
        uint32_t lane = ii & 3, f = freq[src[ii]], c = cdf[src[ii]];
        while (x[lane] > 4096 * f - 1) { stream[ns++] = (uint8_t)x[lane]; x[lane] >>= 8; }
        x[lane] = (x[lane] / f) * 4096 + c + (x[lane] % f);
    }
    out[0] = 'V'; out[1] = 'A'; out[2] = 0x01; out[3] = 0x00;
    wr32(out + 4, n);
    uint32_t pos = 8 + orc_joint_nibble(freq, out + 8);
    for (int j = 0; j < 4; j++) { wr32(out + pos, x[j]); pos += 4; }
    for (uint32_t k = ns; k-- > 0;) out[pos++] = stream[k];
    free(stream);
    return pos;
}

/* literals.go:368-384 */
uint32_t orc_vans4x_adaptive(const uint8_t *src, uint32_t n, uint8_t *out) {
    uint32_t counts[256];
    uint16_t freq[256];
    if (!n) return 0;
    orc_histogram(src, n, counts);
    orc_normalize(counts, freq);
    return orc_vans4x_encode(src, n, freq, out);
}

/* vans_meta.go:45-68 */
static int meta_ctx(uint8_t p) {
    if (p == 0) return 0;
    if (p < 4) return 1;
    if (p < 16) return 2;
    if (p < 32) return 3;
    if (p < 64) return 4;
    if (p < 128) return 5;
    if (p < 192) return 6;
    return 7;
}

/* vans_meta.go:104-150, n > 0.  out holds 24 + 2n bytes. */
uint32_t orc_meta8_encode(const uint8_t *src, uint32_t n, uint8_t *out) {
    uint16_t cdf[8][256];
    for (int c = 0; c < 8; c++) {
        uint32_t cum = 0;
        for (int s = 0; s < 256; s++) { cdf[c][s] = (uint16_t)cum; cum += przm_meta8_freq[c][s]; }
    }
    uint8_t *stream = (uint8_t *)malloc(2 * (size_t)n + 16);
    uint32_t ns = 0;
    uint32_t x[4] = {65536, 65536, 65536, 65536};
    for (uint32_t ii = n; ii-- > 0;) {
        uint32_t lane = ii & 3;
        int ctx = ii > 0 ? meta_ctx(src[ii - 1]) : 0;
        uint32_t f = przm_meta8_freq[ctx][src[ii]], c = cdf[ctx][src[ii]];
        while (x[lane] > 4096 * f - 1) { stream[ns++] = (uint8_t)x[lane]; x[lane] >>= 8; }
        x[lane] = (x[lane] / f) * 4096 + c + (x[lane] % f);
    }
    out[0] = 'V'; out[1] = 'A'; out[2] = 0x01; out[3] = 0x41;
    wr32(out + 4, n);
    for (int j = 0; j < 4; j++) wr32(out + 8 + 4 * j, x[j]);
    for (uint32_t k = 0; k < ns; k++) out[24 + k] = stream[ns - 1 - k];
    free(stream);
    return 24 + ns;
}

/* commands.go:371-406.  out holds 625 + 2n bytes.  *choice: 0 flat,
 * 1 Meta8, 2 adaptive V-ANS. */
uint32_t orc_pack_substream(const uint8_t *sub, uint32_t n, uint8_t *out, int *choice) {
    uint32_t best = 1 + n;
    int ch = 0;
    uint8_t *tmp = (uint8_t *)malloc(8 + 600 + 16 + 2 * (size_t)n + 32);
    if (n >= 24) {
        uint32_t l = orc_meta8_encode(sub, n, tmp);
        if (l < best) { best = l; ch = 1; memcpy(out, tmp, l); }
    }
    if (n >= 16) {
        uint32_t l = orc_vans4x_adaptive(sub, n, tmp);
        if (l < best) { best = l; ch = 2; memcpy(out, tmp, l); }
    }
    free(tmp);
    if (ch == 0) { out[0] = 0x09; memcpy(out + 1, sub, n); }
    if (choice) *choice = ch;
    return best;
}

/* commands.go:409-444 + 537-568.  out holds 16 + cmd bytes... sized by
 * caller to 16 + 2*(ncmd) + 3*625. Returns the section length. */
uint32_t orc_package_commands(const orc_streams *s, uint8_t *out, int choices[3]) {
    if (!s->ncmd) return 0;
    uint8_t *buf = (uint8_t *)malloc(13 + 3 * 640 + 2 * ((size_t)s->nop + s->nmeta + s->nvar));
    uint32_t pos = 13;
    pos += orc_pack_substream(s->op, s->nop, buf + pos, choices ? &choices[0] : NULL);
    uint32_t l_op = pos - 13;
    uint32_t before = pos;
    pos += orc_pack_substream(s->meta, s->nmeta, buf + pos, choices ? &choices[1] : NULL);
    uint32_t l_m = pos - before;
    pos += orc_pack_substream(s->var, s->nvar, buf + pos, choices ? &choices[2] : NULL);
    buf[0] = 0x08;
    wr32(buf + 1, s->ncmd);
    wr32(buf + 5, l_op);
    wr32(buf + 9, l_m);
    uint32_t len;
    if (pos < s->ncmd) {
        memcpy(out, buf, pos);
        len = pos;
    } else {
        out[0] = 0x09;
        memcpy(out + 1, s->cmd, s->ncmd);
        len = 1 + s->ncmd;
        if (choices) choices[0] = choices[1] = choices[2] = -1;
    }
    free(buf);
    return len;
}

/* Section 2 of a whole command stream (commands.go:537-568): split, then Mode 8
 * or flat.  out holds 16 + 2n + 3*640 bytes; choices as orc_package_commands.
 * Returns the section length, 0 for an empty stream, or -1 when the stream
 * does not split (an opcode mode G never emits). */
int64_t orc_package_cmd_stream(const uint8_t *cmd, uint32_t n, uint8_t *out, int choices[3]) {
    orc_streams s;
    memset(&s, 0, sizeof s);
    uint8_t *mem = (uint8_t *)malloc(3 * (size_t)n + 16);
    s.cmd = (uint8_t *)cmd;
    s.ncmd = n;
    s.op = mem; s.meta = mem + n + 4; s.var = mem + 2 * (size_t)n + 8;
    if (orc_split_commands(cmd, n, s.op, &s.nop, s.meta, &s.nmeta, s.var, &s.nvar)) { free(mem); return -1; }
    int64_t len = orc_package_commands(&s, out, choices);
    free(mem);
    return len;
}

/* ------------------------------------------------------------------ */
/* Section 3 as a distance-prefix container (the encoder option that   */
/* brings PackageOffsetsSection of the reference CPU path to mode G).  */

/* distance.go:165-180: code, extra bits and their count of a distance d >= 1 */
void orc_dist_code(uint32_t d, uint8_t *code, uint32_t *extra, uint8_t *nb) {
    if (d <= 4) { *code = (uint8_t)(d - 1); *extra = 0; *nb = 0; return; }
    if (d <= 16777216u) {
        const uint32_t u = d - 1;
        const uint8_t b = (uint8_t)(32 - __builtin_clz(u) - 2);
        const uint32_t half = (u >> b) & 1u;
        *code = (uint8_t)(2 * b + 2 + half);
        *extra = d - (((2u + half) << b) + 1u);
        *nb = b;
        return;
    }
    *code = 48; *extra = d - 16777217u; *nb = 24;
}

/* distance.go:190-202 */
static int dist_ctx(uint8_t prev) { return prev < 6 ? 0 : prev < 14 ? 1 : prev < 26 ? 2 : 3; }

/* distance.go:323-376: the codes by one rANS state whose table is chosen by the class of the
 * previous code (0 before the first); [0x0C][u32 final state][stream].  out holds 5 + 2n. */
uint32_t orc_dist_m12_encode(const uint8_t *codes, uint32_t n, uint8_t *out) {
    uint16_t cdf[4][64];
    for (int c = 0; c < 4; c++) {
        uint32_t cum = 0;
        for (int s = 0; s < 64; s++) { cdf[c][s] = (uint16_t)cum; cum += przm_dist_freq[c][s]; }
    }
    uint8_t *stream = (uint8_t *)malloc(2 * (size_t)n + 16);
    uint32_t ns = 0, x = 65536;
    for (uint32_t i = n; i-- > 0;) {
        const uint8_t sym = codes[i];
        const int ctx = dist_ctx(i ? codes[i - 1] : 0);
        const uint32_t f = przm_dist_freq[ctx][sym], c = cdf[ctx][sym];
        while (x > 4096 * f - 1) { stream[ns++] = (uint8_t)x; x >>= 8; }
        x = (x / f) * 4096 + c + (x % f);
    }
    out[0] = 0x0C;
    wr32(out + 1, x);
    for (uint32_t k = 0; k < ns; k++) out[5 + k] = stream[ns - 1 - k];
    free(stream);
    return 5 + ns;
}

static uint32_t varint_len(uint32_t d) { /* distance.go:760-775 */
    return d <= 0x7F ? 1 : d <= 0x3FFF ? 2 : d <= 0x1FFFFF ? 3 : d <= 0x0FFFFFFF ? 4 : 5;
}

/* PackDistancePrefix with the mode chosen automatically and the coded code stream
 * (distance.go:692-748), SelectPrefixMode (750-758), assemblePrefixContainer (777-823).
 * d[i] in [1, 2^24].  out holds 13 + 5n + 700 bytes.  info (optional): [0] 0 Mode-12 /
 * 1 adaptive V-ANS code stream, [1] code stream bytes, [2] extra-bit bytes, [3] mode byte. */
uint32_t orc_pack_distance_prefix(const uint32_t *d, uint32_t n, uint8_t *out, uint32_t info[4]) {
    if (!n) return 0;
    uint8_t *codes = (uint8_t *)malloc((size_t)n + 16);
    uint8_t *extra = (uint8_t *)malloc(3 * (size_t)n + 16);
    uint8_t *cs = (uint8_t *)malloc(2 * (size_t)n + 700);
    uint8_t *alt = (uint8_t *)malloc(2 * (size_t)n + 700);
    uint32_t ne = 0, orig = 0, cnt = 0;
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; i++) { /* BitWriter, MSB first (distance.go:217-247) */
        uint8_t c, nb;
        uint32_t e;
        orc_dist_code(d[i], &c, &e, &nb);
        codes[i] = c;
        if (nb) {
            acc = (acc << nb) | ((uint64_t)e & ((1ull << nb) - 1));
            cnt += nb;
            while (cnt >= 8) { cnt -= 8; extra[ne++] = (uint8_t)(acc >> cnt); }
        }
        orig += varint_len(d[i]);
    }
    if (cnt) extra[ne++] = (uint8_t)(acc << (8 - cnt));
    uint32_t cl = orc_dist_m12_encode(codes, n, cs), kind = 0;
    if (n >= 64) {
        const uint32_t al = orc_vans4x_adaptive(codes, n, alt);
        if (al < cl) { cl = al; memcpy(cs, alt, al); kind = 1; }
    }
    uint32_t h;
    if (orig <= 255 && n <= 255 && cl <= 255) {
        out[0] = 0x0B; out[1] = (uint8_t)orig; out[2] = (uint8_t)n; out[3] = (uint8_t)cl; h = 4;
    } else if (orig <= 65535 && n <= 65535 && cl <= 65535) {
        out[0] = 0x0A;
        out[1] = (uint8_t)orig; out[2] = (uint8_t)(orig >> 8);
        out[3] = (uint8_t)n; out[4] = (uint8_t)(n >> 8);
        out[5] = (uint8_t)cl; out[6] = (uint8_t)(cl >> 8);
        h = 7;
    } else {
        out[0] = 0x06; wr32(out + 1, orig); wr32(out + 5, n); wr32(out + 9, cl); h = 13;
    }
    memcpy(out + h, cs, cl);
    memcpy(out + h + cl, extra, ne);
    if (info) { info[0] = kind; info[1] = cl; info[2] = ne; info[3] = out[0]; }
    free(codes); free(extra); free(cs); free(alt);
    return h + cl + ne;
}

/* PackageOffsetsSection (distance.go:825-849) for a block's LEB128 offset stream off[0, noff):
 * the prefix container when it is shorter than the varints, else 0x09 + the varints.
 * Mode-G rule added on top of the reference: a container whose code stream is exactly as
 * long as the match count is written flat too, since both reference decoders take such a
 * code stream for raw codes (distance.go:601-604, src/przm_dec.c:1097-1098).
 * out holds 16 + 3 noff + 700.  Returns the section length (0 for no offsets); *mode gets
 * the first byte, *code_kind 0 Mode-12 / 1 V-ANS (or -1 for a flat section). */
uint32_t orc_package_offsets(const uint8_t *off, uint32_t noff, uint8_t *out, int *mode, int *code_kind) {
    if (mode) *mode = 0;
    if (code_kind) *code_kind = -1;
    if (!noff) return 0;
    uint32_t *d = (uint32_t *)malloc(((size_t)noff + 1) * sizeof(uint32_t));
    uint32_t n = 0, raw = 0, v = 0, sh = 0;
    for (uint32_t i = 0; i < noff; i++) {
        v |= (uint32_t)(off[i] & 0x7F) << sh;
        sh += 7;
        if (!(off[i] & 0x80)) { d[n++] = v; raw += varint_len(v); v = 0; sh = 0; }
    }
    uint32_t info[4] = {0, 0, 0, 0};
    const uint32_t pl = orc_pack_distance_prefix(d, n, out, info);
    free(d);
    if (pl < raw && info[1] != n) {
        if (mode) *mode = out[0];
        if (code_kind) *code_kind = (int)info[0];
        return pl;
    }
    out[0] = 0x09;
    memcpy(out + 1, off, noff);
    if (mode) *mode = 0x09;
    return 1 + noff;
}

/* ------------------------------------------------------------------ */
/* Section 4 with the sheaf trials (the encoder option that brings     */
/* PackageLiteralsSection of the reference CPU path to mode G).        */

static int sheaf_ws(uint8_t b) { return b == ' ' || b == '\t' || b == '\r' || b == '\n'; }
static int sheaf_syntax(uint8_t b) { /* sheaf.go:21-25 */
    return b && strchr("{}[]()<>:;,.\"'`=+-*/%&|^!~?\\#", b) != NULL;
}
static int sheaf_ident(uint8_t b) { return (b >= 'a' && b <= 'z') || b == '_'; }
static int sheaf_digit(uint8_t b) { return b >= '0' && b <= '9'; }

/* the fiber of a literal whose two previous literals are p2, p1 (0 before the first):
 * ContextClassify modes 0-2 (sheaf.go:36-83) and ContextClassifyOrder2 mode 3 (86-102) */
int orc_sheaf_class(uint8_t p2, uint8_t p1, int mode) {
    switch (mode) {
    case 0: return sheaf_ws(p1) ? 0 : sheaf_syntax(p1) ? 1 : sheaf_ident(p1) ? 2 : 3;
    case 1: return (p1 == '"' || p1 == ':') ? 0 : sheaf_ident(p1) ? 1 : sheaf_digit(p1) ? 2 : 3;
    case 2: return (sheaf_ws(p1) || sheaf_syntax(p1)) ? 0 : 1;
    default:
        if (sheaf_digit(p1)) return (sheaf_digit(p2) || p2 == '.' || p2 == '-') ? 0 : 1;
        if (sheaf_ident(p1)) return sheaf_ident(p2) ? 2 : 3;
        return sheaf_ws(p1) ? 0 : 1;
    }
}

/* PackEtaleSheaf (literals.go:601-664) of lits[0, n) in one mode: DecomposeFibers
 * (sheaf.go:125-152), each fiber as EncodeVANS4x (the container without its 8-byte head,
 * literals.go:344-355) with its own normalised table when shorter than the fiber, and the
 * 40-byte header.  out holds 40 + 2n + 4 * 600 bytes.  Returns the container length. */
uint32_t orc_pack_sheaf(const uint8_t *lits, uint32_t n, int mode, uint8_t *out) {
    uint8_t *fib = (uint8_t *)malloc((size_t)n + 16);
    uint8_t *enc = (uint8_t *)malloc(2 * (size_t)n + 700);
    uint32_t l[4] = {0, 0, 0, 0}, start[4], e[4];
    uint8_t p1 = 0, p2 = 0;
    for (uint32_t i = 0; i < n; i++) {
        l[orc_sheaf_class(p2, p1, mode)]++;
        p2 = p1;
        p1 = lits[i];
    }
    start[0] = 0;
    for (int k = 1; k < 4; k++) start[k] = start[k - 1] + l[k - 1];
    uint32_t fill[4] = {start[0], start[1], start[2], start[3]};
    p1 = p2 = 0;
    for (uint32_t i = 0; i < n; i++) {
        fib[fill[orc_sheaf_class(p2, p1, mode)]++] = lits[i];
        p2 = p1;
        p1 = lits[i];
    }
    memcpy(out, "VA\x01", 3);
    out[3] = (uint8_t)(0x01 | 0x04 | ((mode & 0x0F) << 4));
    wr32(out + 4, n);
    uint32_t pos = 40;
    for (int k = 0; k < 4; k++) {
        e[k] = 0;
        if (!l[k]) continue;
        uint32_t counts[256];
        uint16_t freq[256];
        orc_histogram(fib + start[k], l[k], counts);
        orc_normalize(counts, freq);
        const uint32_t cl = orc_vans4x_encode(fib + start[k], l[k], freq, enc) - 8;
        if (cl < l[k]) { memcpy(out + pos, enc + 8, cl); e[k] = cl; }
        else { memcpy(out + pos, fib + start[k], l[k]); e[k] = l[k]; }
        pos += e[k];
    }
    for (int k = 0; k < 4; k++) { wr32(out + 8 + 4 * k, l[k]); wr32(out + 24 + 4 * k, e[k]); }
    free(fib);
    free(enc);
    return pos;
}

/* PackageLiteralsSection (literals.go:667-693): the shortest of the adaptive V-ANS container
 * and the four sheaf containers, tried in that order with strict <, when shorter than the
 * literals; else the literals.  out holds 40 + 2n + 4 * 600 bytes.  *kind: 0 raw, 1 V-ANS,
 * 2 + mode for a sheaf container (the frame then sets flag 0x80). */
uint32_t orc_package_literals(const uint8_t *lits, uint32_t n, uint8_t *out, int *kind) {
    *kind = 0;
    if (!n) return 0;
    uint8_t *tmp = (uint8_t *)malloc(40 + 2 * (size_t)n + 2400);
    uint32_t best = n;
    const uint32_t lv = orc_vans4x_adaptive(lits, n, tmp);
    if (lv < best) { best = lv; *kind = 1; memcpy(out, tmp, lv); }
    for (int m = 0; m < 4; m++) {
        const uint32_t ls = orc_pack_sheaf(lits, n, m, tmp);
        if (ls < best) { best = ls; *kind = 2 + m; memcpy(out, tmp, ls); }
    }
    if (*kind == 0) memcpy(out, lits, n);
    free(tmp);
    return best;
}

/* compress.go:113-132 */
int orc_check_high_entropy(const uint8_t *src, uint32_t n) {
    uint32_t sample = n < 512 ? n : 512;
    if (sample < 64) return 0;
    uint8_t seen[256] = {0};
    uint32_t unique = 0;
    for (uint32_t i = 0; i < sample; i++) if (!seen[src[i]]) { seen[src[i]] = 1; unique++; }
    uint32_t thr = sample >= 256 ? (256 * 88) / 100 : (sample * 88) / 100;
    return unique > thr;
}

/* compress.go:183-202 */
uint32_t orc_window_log(uint32_t u) {
    if (u <= 1024) return 10;
    if (u <= 4096) return 12;
    if (u <= 16384) return 14;
    if (u <= 65536) return 16;
    if (u <= 262144) return 18;
    if (u <= 1048576) return 20;
    if (u <= 4194304) return 22;
    return 24;
}

static uint32_t raw_frame(const uint8_t *src, uint32_t n, uint8_t *out) { /* compress.go:973-981 */
    memcpy(out, "PRZM", 4);
    out[4] = 0x04;
    out[5] = (uint8_t)n; out[6] = (uint8_t)(n >> 8); out[7] = (uint8_t)(n >> 16);
    if (n) memcpy(out + 8, src, n);
    return 8 + n;
}

/* Frame-level trace for tests. */
typedef struct {
    uint32_t kind;        /* 0 compressed, 1 uncompressed: entropy, 2: size guard, 3: S4 hazard */
    uint32_t ncmd, noff, nlit, nop, nmeta, nvar, ntok;
    uint32_t s2, s3, s4;
    int32_t sub_choice[3];
    uint32_t s4_vans;
    uint32_t s3_mode;     /* first byte of Section 3: 0x09 flat, 0x06 / 0x0A / 0x0B prefix; 0 none */
    int32_t s3_code;      /* prefix code stream: 0 Mode-12, 1 adaptive V-ANS; -1 flat or none */
    uint32_t s4_kind;     /* Section 4: 0 raw, 1 adaptive V-ANS, 2 + mode a sheaf container */
} orc_frame_info;

#define ORC_OPT_PREFIX 1u /* Section 3 as PackageOffsetsSection chooses it */
#define ORC_OPT_SHEAF 2u  /* Section 4 as PackageLiteralsSection chooses it (sheaf trials) */

/*
 * One block (n <= 65536) -> one frame.  out holds 8 + n bytes.
 * compress.go:962-1033 with the mode-G sections described above.
 */
uint32_t orc_encode_block_opts(const uint8_t *src, uint32_t n, int level, unsigned opts, uint8_t *out,
                               orc_frame_info *info) {
    orc_frame_info tmp_info;
    if (!info) info = &tmp_info;
    memset(info, 0, sizeof(*info));
    info->s3_code = -1;
    if (n == 0 || orc_check_high_entropy(src, n)) {
        info->kind = n ? 1 : 2;
        return raw_frame(src, n, out);
    }
    uint32_t *desc = (uint32_t *)malloc(n * sizeof(uint32_t));
    orc_match_block(src, n, orc_search_depth(level), desc, NULL);
    orc_streams s;
    size_t cap = 2 * (size_t)n + 64;
    uint8_t *mem = (uint8_t *)malloc(6 * cap);
    s.cmd = mem; s.off = mem + cap; s.lit = mem + 2 * cap; s.op = mem + 3 * cap; s.meta = mem + 4 * cap;
    s.var = mem + 5 * cap;
    orc_tokenize_block(src, n, desc, &s);
    free(desc);
    info->ncmd = s.ncmd; info->noff = s.noff; info->nlit = s.nlit;
    info->nop = s.nop; info->nmeta = s.nmeta; info->nvar = s.nvar; info->ntok = s.ntok;

    uint8_t *s2 = (uint8_t *)malloc(16 + 2 * (size_t)s.ncmd + 3 * 640);
    uint32_t l2 = orc_package_commands(&s, s2, info->sub_choice);
    /* Section 3: flat 0x09, or with ORC_OPT_PREFIX PackageOffsetsSection's choice */
    uint8_t *s3 = (uint8_t *)malloc(16 + 3 * (size_t)s.noff + 700);
    uint32_t l3;
    if (opts & ORC_OPT_PREFIX) {
        int mode, ck;
        l3 = orc_package_offsets(s.off, s.noff, s3, &mode, &ck);
        info->s3_mode = (uint32_t)mode;
        info->s3_code = ck;
    } else {
        l3 = s.noff ? 1 + s.noff : 0;
        if (l3) { s3[0] = 0x09; memcpy(s3 + 1, s.off, s.noff); }
        info->s3_mode = l3 ? 0x09 : 0;
    }
    /* Section 4: adaptive V-ANS or raw, or with ORC_OPT_SHEAF PackageLiteralsSection's choice */
    uint8_t *s4 = (uint8_t *)malloc(40 + 2400 + 16 + 2 * (size_t)s.nlit);
    uint32_t l4 = 0;
    int vans = 0;
    if (s.nlit && (opts & ORC_OPT_SHEAF)) {
        int k;
        l4 = orc_package_literals(s.lit, s.nlit, s4, &k);
        vans = k != 0;
        info->s4_kind = (uint32_t)k;
    } else if (s.nlit) {
        uint32_t lv = orc_vans4x_adaptive(s.lit, s.nlit, s4);
        if (lv < s.nlit) { l4 = lv; vans = 1; }
        else { memcpy(s4, s.lit, s.nlit); l4 = s.nlit; }
        info->s4_kind = (uint32_t)vans;
    }
    info->s2 = l2; info->s3 = l3; info->s4 = l4; info->s4_vans = (uint32_t)vans;
    int ext = l2 > 65535 || l3 > 65535;
    uint32_t dir = ext ? 6 : 4;
    uint32_t comp = 8 + dir + 8 + l2 + l3 + l4;
    int hazard = !vans && l4 >= 8 && ((s4[0] == 'V' && s4[1] == 'A' && s4[2] == 0x01) ||
                                      rd32(s4 + 4) == (rd32(s4) + 1) / 2);
    uint32_t len;
    if (comp >= n || hazard) {
        info->kind = comp >= n ? 2 : 3;
        len = raw_frame(src, n, out);
    } else {
        uint8_t flags = (uint8_t)((vans ? 0x80 : 0) | (ext ? 0x08 : 0));
        memcpy(out, "PRZM", 4);
        out[4] = flags;
        out[5] = (uint8_t)n; out[6] = (uint8_t)(n >> 8); out[7] = (uint8_t)(n >> 16);
        uint32_t p = 8;
        if (ext) {
            out[p++] = (uint8_t)l2; out[p++] = (uint8_t)(l2 >> 8); out[p++] = (uint8_t)(l2 >> 16);
            out[p++] = (uint8_t)l3; out[p++] = (uint8_t)(l3 >> 8); out[p++] = (uint8_t)(l3 >> 16);
        } else {
            out[p++] = (uint8_t)l2; out[p++] = (uint8_t)(l2 >> 8);
            out[p++] = (uint8_t)l3; out[p++] = (uint8_t)(l3 >> 8);
        }
        memset(out + p, 0, 8);
        out[p + 1] = (uint8_t)orc_window_log(n);
        p += 8;
        memcpy(out + p, s2, l2); p += l2;
        memcpy(out + p, s3, l3); p += l3;
        memcpy(out + p, s4, l4); p += l4;
        len = p;
    }
    free(s2); free(s3); free(s4); free(mem);
    return len;
}

uint32_t orc_encode_block(const uint8_t *src, uint32_t n, int level, uint8_t *out, orc_frame_info *info) {
    return orc_encode_block_opts(src, n, level, 0, out, info);
}

/* A buffer -> one frame per 64 KiB block (an empty buffer -> one empty
 * uncompressed frame, as compress.go:1036-1114 produces for u == 0).
 * frame_sizes gets one entry per frame; returns total bytes, or 0 when
 * out_cap / frames_cap is too small. */
uint64_t orc_encode_buffer_opts(const uint8_t *src, uint64_t n, int level, unsigned opts, uint8_t *out,
                                uint64_t out_cap, uint32_t *frame_sizes, uint64_t frames_cap, uint64_t *nframes) {
    uint64_t nb = n ? (n + ORC_BLOCK - 1) / ORC_BLOCK : 1;
    if (nb > frames_cap) return 0;
    uint64_t total = 0;
    uint8_t *tmp = (uint8_t *)malloc(ORC_BLOCK + 8);
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t off = b * ORC_BLOCK;
        uint32_t len = (uint32_t)(n - off < ORC_BLOCK ? n - off : ORC_BLOCK);
        if (!n) len = 0;
        uint32_t fl = orc_encode_block_opts(src + off, len, level, opts, tmp, NULL);
        if (total + fl > out_cap) { free(tmp); return 0; }
        memcpy(out + total, tmp, fl);
        frame_sizes[b] = fl;
        total += fl;
    }
    free(tmp);
    *nframes = nb;
    return total;
}

uint64_t orc_encode_buffer(const uint8_t *src, uint64_t n, int level, uint8_t *out, uint64_t out_cap,
                           uint32_t *frame_sizes, uint64_t frames_cap, uint64_t *nframes) {
    return orc_encode_buffer_opts(src, n, level, 0, out, out_cap, frame_sizes, frames_cap, nframes);
}

/* Stream-level entry point for tests: tokenize one block into caller
 * buffers (each >= 2n + 64 bytes); lens[0..6] = cmd, off, lit, op, meta,
 * var, tokens. */
void orc_tokenize_streams(const uint8_t *src, uint32_t n, int level, uint8_t *cmd, uint8_t *off, uint8_t *lit,
                          uint8_t *op, uint8_t *meta, uint8_t *var, uint32_t lens[7]) {
    uint32_t *desc = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    orc_match_block(src, n, orc_search_depth(level), desc, NULL);
    orc_streams s = {cmd, 0, off, 0, lit, 0, op, 0, meta, 0, var, 0, 0};
    orc_tokenize_block(src, n, desc, &s);
    free(desc);
    lens[0] = s.ncmd; lens[1] = s.noff; lens[2] = s.nlit; lens[3] = s.nop; lens[4] = s.nmeta; lens[5] = s.nvar;
    lens[6] = s.ntok;
}

/* Tokenize from caller-supplied descriptors (tests feed brute-force ones). */
void orc_tokenize_desc(const uint8_t *src, uint32_t n, const uint32_t *desc, uint8_t *cmd, uint8_t *off,
                       uint8_t *lit, uint8_t *op, uint8_t *meta, uint8_t *var, uint32_t lens[7]) {
    orc_streams s = {cmd, 0, off, 0, lit, 0, op, 0, meta, 0, var, 0, 0};
    orc_tokenize_block(src, n, desc, &s);
    lens[0] = s.ncmd; lens[1] = s.noff; lens[2] = s.nlit; lens[3] = s.nop; lens[4] = s.nmeta; lens[5] = s.nvar;
    lens[6] = s.ntok;
}
