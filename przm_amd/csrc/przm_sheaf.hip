/*
 * przm_sheaf.hip -- Section 4 with the sheaf trials on gfx950: the encoder
 * option that writes each frame's literals as PackageLiteralsSection chooses
 * them (core/literals/literals.go:667-693): the shortest of the raw literals,
 * the adaptive V-ANS container and the four sheaf containers.
 *
 *  przm_sheaf_split_kernel   one workgroup per (block, mode): each literal's
 *                            fiber from the class of the literals before it
 *                            (ContextClassify / ContextClassifyOrder2,
 *                            core/literals/sheaf.go:36-102); a counting pass
 *                            builds the fiber lengths and histograms, one wave
 *                            per fiber normalises its table (NormalizeFrequencies
 *                            and the matrix, literals.go:78-184), and a scatter
 *                            pass writes the fibers (DecomposeFibers,
 *                            sheaf.go:125-152), each 16-byte aligned.
 *  przm_sheaf_count_kernel   one lane quad per (block, mode, fiber): the exact
 *                            length of the fiber's V-ANS stream (the coder of
 *                            literals.go:289-341 without its stores), each lane
 *                            its own state; it stops once the fiber can no
 *                            longer be shorter coded than raw (literals.go:622).
 *  przm_sheaf_encode_kernel  one lane quad per (block, fiber): the section's
 *                            choice, in PackageLiteralsSection's order with
 *                            strict < (raw, V-ANS, modes 0..3), and the coded
 *                            fibers of the chosen mode (rans_quad).
 *
 * The frame kernels then write the chosen container (assembleSheafContainer,
 * literals.go:645-664).
 */

#define SHEAF_NT 256u

/* the class tables of modes 0-2, by the previous literal (sheaf.go:36-83) */
struct SheafTable {
    uint8_t c[3][256];
    static constexpr bool ws(int b) { return b == ' ' || b == '\t' || b == '\r' || b == '\n'; }
    static constexpr bool syntax(int b) {
        return b == '{' || b == '}' || b == '[' || b == ']' || b == '(' || b == ')' || b == '<' || b == '>' ||
               b == ':' || b == ';' || b == ',' || b == '.' || b == '"' || b == '\'' || b == '`' || b == '=' ||
               b == '+' || b == '-' || b == '*' || b == '/' || b == '%' || b == '&' || b == '|' || b == '^' ||
               b == '!' || b == '~' || b == '?' || b == '\\' || b == '#';
    }
    static constexpr bool ident(int b) { return (b >= 'a' && b <= 'z') || b == '_'; }
    static constexpr bool digit(int b) { return b >= '0' && b <= '9'; }
    constexpr SheafTable() : c() {
        for (int b = 0; b < 256; b++) {
            c[0][b] = ws(b) ? 0 : syntax(b) ? 1 : ident(b) ? 2 : 3;
            c[1][b] = (b == '"' || b == ':') ? 0 : ident(b) ? 1 : digit(b) ? 2 : 3;
            c[2][b] = (ws(b) || syntax(b)) ? 0 : 1;
        }
    }
};
__constant__ SheafTable przm_sheaf_lut = SheafTable();

/* the fiber of a literal after p2, p1 in mode m (mode 3: ContextClassifyOrder2, sheaf.go:86-102) */
__device__ __forceinline__ uint32_t sheaf_class(const uint8_t (*lut)[256], uint32_t p2, uint32_t p1, uint32_t m) {
    if (m < 3) return lut[m][p1];
    const bool d1 = p1 - '0' < 10u, i1 = (p1 - 'a' < 26u) || p1 == '_';
    if (d1) return ((p2 - '0' < 10u) || p2 == '.' || p2 == '-') ? 0u : 1u;
    if (i1) return ((p2 - 'a' < 26u) || p2 == '_') ? 2u : 3u;
    return (p1 == ' ' || p1 == '\t' || p1 == '\r' || p1 == '\n') ? 0u : 1u;
}

__device__ __forceinline__ uint32_t al16(uint32_t v) { return (v + 15u) & ~15u; }

/* byte i of the literals, 0 before the first */
__device__ __forceinline__ uint32_t lit_at(const uint8_t *L, int64_t i) { return i >= 0 ? L[i] : 0u; }

extern "C" __global__ void __launch_bounds__(SHEAF_NT)
przm_sheaf_split_kernel(const uint8_t *__restrict__ ws, const przm_meta_t *__restrict__ meta,
                        uint8_t *__restrict__ sws, uint32_t nblocks) {
    __shared__ uint8_t lut[3][256];
    __shared__ uint32_t cnt[4][256];
    __shared__ uint32_t key[4][256];
    __shared__ uint32_t frq[4][256];
    __shared__ uint32_t part[SHEAF_NT / 64u][2];
    const uint32_t blk = blockIdx.x >> 2, m = blockIdx.x & 3u, tid = threadIdx.x, w = tid >> 6;
    if (blk >= nblocks) return;
    const uint8_t *L = ws + (uint64_t)blk * PRZM_WS_BYTES + PRZM_WS_LIT;
    const uint32_t n = meta[blk].len[PRZM_S_LIT];
    uint8_t *sb = sws + (uint64_t)blk * PRZM_SWS_BYTES;
    przm_sheaf_t *H = (przm_sheaf_t *)(sb + PRZM_SWS_HDR);
    uint8_t *fib = sb + PRZM_SWS_FIB + (uint64_t)m * PRZM_SWS_FIB_CAP;
    for (uint32_t i = tid; i < 3 * 256; i += SHEAF_NT) (&lut[0][0])[i] = (&przm_sheaf_lut.c[0][0])[i];
    for (uint32_t i = tid; i < 4 * 256; i += SHEAF_NT) (&cnt[0][0])[i] = 0;
    __syncthreads();
    /* counting pass: each thread 16 literals a round */
    for (uint32_t base = 0; base < n; base += 16u * SHEAF_NT) {
        const uint32_t i0 = base + 16u * tid;
        uint32_t p2 = lit_at(L, (int64_t)i0 - 2), p1 = lit_at(L, (int64_t)i0 - 1);
        for (uint32_t t = 0; t < 16u && i0 + t < n; t++) {
            const uint32_t b = L[i0 + t];
            atomicAdd(&cnt[sheaf_class(lut, p2, p1, m)][b], 1u);
            p2 = p1;
            p1 = b;
        }
    }
    __syncthreads();
    /* fiber lengths (wave k sums fiber k's histogram) and tables */
    uint32_t lk = 0;
    for (uint32_t s = tid & 63u; s < 256; s += 64) lk += cnt[w][s];
    lk = wave_sum(lk);
    if ((tid & 63u) == 0) part[w][0] = lk;
    __syncthreads();
    uint32_t l[4], fs[4];
    for (uint32_t k = 0; k < 4; k++) l[k] = part[k][0];
    fs[0] = 0;
    for (uint32_t k = 1; k < 4; k++) fs[k] = fs[k - 1] + al16(l[k - 1]);
    if (tid == 0) {
        if (m == 0) H->n = n;
        for (uint32_t k = 0; k < 4; k++) {
            H->l[m][k] = l[k];
            if (!l[k]) H->mat[m][k] = 0; /* a fiber's own wave writes it otherwise */
        }
    }
    if (l[w] > 0)
        normalize_wave(cnt[w], key[w], frq[w], (przm_ftab_t *)(sb + PRZM_SWS_FTAB) + 4u * m + w, &H->mat[m][w]);
    __syncthreads();
    /* scatter pass: a literal's place in its fiber is the number of earlier literals of that
     * fiber; per round one workgroup scan of the four per-thread counts (16 bits each: a round
     * holds at most 4096 literals) */
    uint32_t run[4] = {0, 0, 0, 0};
    for (uint32_t base = 0; base < n; base += 16u * SHEAF_NT) {
        const uint32_t i0 = base + 16u * tid;
        uint32_t p2 = lit_at(L, (int64_t)i0 - 2), p1 = lit_at(L, (int64_t)i0 - 1);
        uint32_t cls[16], mine[4] = {0, 0, 0, 0};
        for (uint32_t t = 0; t < 16u; t++) {
            const uint32_t b = i0 + t < n ? L[i0 + t] : 0u;
            cls[t] = sheaf_class(lut, p2, p1, m);
            if (i0 + t < n) mine[cls[t]]++;
            p2 = p1;
            p1 = b;
        }
        const uint32_t lo = mine[0] | (mine[1] << 16), hi = mine[2] | (mine[3] << 16);
        const uint32_t lin = wave_scan_incl(lo), hin = wave_scan_incl(hi);
        if ((tid & 63u) == 63u) { part[w][0] = lin; part[w][1] = hin; }
        __syncthreads();
        uint32_t blo = 0, bhi = 0, tlo = 0, thi = 0;
        for (uint32_t k = 0; k < SHEAF_NT / 64u; k++) {
            if (k < w) { blo += part[k][0]; bhi += part[k][1]; }
            tlo += part[k][0];
            thi += part[k][1];
        }
        __syncthreads();
        const uint32_t elo = blo + lin - lo, ehi = bhi + hin - hi;
        uint32_t pos[4] = {fs[0] + run[0] + (elo & 0xFFFFu), fs[1] + run[1] + (elo >> 16),
                           fs[2] + run[2] + (ehi & 0xFFFFu), fs[3] + run[3] + (ehi >> 16)};
        for (uint32_t t = 0; t < 16u && i0 + t < n; t++) fib[pos[cls[t]]++] = L[i0 + t];
        run[0] += tlo & 0xFFFFu;
        run[1] += tlo >> 16;
        run[2] += thi & 0xFFFFu;
        run[3] += thi >> 16;
    }
}

/*
 * Stream bytes of the V-ANS coding of fiber src[0, n) with table tab, by one lane quad: lane
 * j runs state j over the symbols i = j (mod 4) from the top (literals.go:302-325), counting
 * the bytes its renormalisation emits.  Once the quad has emitted `limit` bytes the fiber
 * cannot be shorter coded than raw: PRZM_ENC_OVERFLOW.
 */
__device__ uint32_t vans_count_quad(const uint8_t *__restrict__ src, uint32_t n, const uint32_t *__restrict__ tab,
                                    uint32_t limit) {
    const uint32_t j = threadIdx.x & 3u;
    const uint4 *s4 = (const uint4 *)src;
    uint32_t x = 65536u, cnt = 0;
    const int32_t nc = (int32_t)((n + 15u) >> 4);
    const uint4 z = make_uint4(0, 0, 0, 0);
    uint4 cur = nc > 0 ? s4[nc - 1] : z;
    for (int32_t c = nc - 1; c >= 0; c--) {
        const uint4 nxt = c > 0 ? s4[c - 1] : z;
        const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
#pragma unroll
        for (int32_t t = 3; t >= 0; t--) {
            const uint32_t i = 16u * (uint32_t)c + 4u * (uint32_t)t + j;
            if (i >= n) continue;
            const uint32_t e = tab[(wd[t] >> (8u * j)) & 0xFFu];
            const uint32_t f = e & 0xFFFFu, lim = 4096u * f - 1u;
            if (x > lim) {
                cnt++;
                x >>= 8;
                if (x > lim) { cnt++; x >>= 8; }
            }
            uint32_t q, r;
            divmod(x, f, &q, &r);
            x = (q << 12) + (e >> 16) + r;
        }
        cur = nxt;
        uint32_t tot = cnt + QPERM(cnt, 0xB1);
        tot += QPERM(tot, 0x4E);
        if (tot >= limit) return PRZM_ENC_OVERFLOW;
    }
    uint32_t tot = cnt + QPERM(cnt, 0xB1);
    tot += QPERM(tot, 0x4E);
    return tot;
}

/* one quad per (block, mode, fiber): job = blockIdx.x * 16 + quad, block = job >> 4 */
extern "C" __global__ void __launch_bounds__(64)
przm_sheaf_count_kernel(uint8_t *__restrict__ sws, uint32_t nblocks) {
    __shared__ uint32_t vt[16][256];
    const uint32_t q = threadIdx.x >> 2, j = threadIdx.x & 3u;
    const uint32_t job = blockIdx.x * 16u + q, blk = job >> 4, m = (job >> 2) & 3u, k = job & 3u;
    const bool has = blk < nblocks;
    uint8_t *sb = sws + (uint64_t)(has ? blk : 0u) * PRZM_SWS_BYTES;
    przm_sheaf_t *H = (przm_sheaf_t *)(sb + PRZM_SWS_HDR);
    uint32_t n = 0, fs = 0, mat = 0;
    if (has) {
        n = H->l[m][k];
        for (uint32_t i = 0; i < k; i++) fs += al16(H->l[m][i]);
        mat = H->mat[m][k];
        if (n) {
            const przm_ftab_t *ft = (const przm_ftab_t *)(sb + PRZM_SWS_FTAB) + 4u * m + k;
            for (uint32_t s = j; s < 256; s += 4) vt[q][s] = ft->fc[s];
        }
    }
    __syncthreads();
    if (!has) return;
    /* coded = matrix + 4 states + stream, taken only when shorter than the fiber */
    const uint32_t head = mat + 16u;
    const uint32_t K = n > head ? vans_count_quad(sb + PRZM_SWS_FIB + (uint64_t)m * PRZM_SWS_FIB_CAP + fs, n, vt[q],
                                                  n - head)
                                : PRZM_ENC_OVERFLOW;
    if (j == 0) H->k[m][k] = K;
}

/* the section's length for choice c (2 + mode for a sheaf container) */
__device__ __forceinline__ uint32_t sheaf_len(const przm_sheaf_t *H, uint32_t m) {
    uint32_t len = 40;
    for (uint32_t k = 0; k < 4; k++) {
        const uint32_t K = H->k[m][k];
        len += (H->l[m][k] && K != PRZM_ENC_OVERFLOW) ? H->mat[m][k] + 16u + K : H->l[m][k];
    }
    return len;
}

/* one quad per (block, fiber): the section's choice (PackageLiteralsSection, literals.go:667-693)
 * and the chosen mode's coded fibers */
extern "C" __global__ void __launch_bounds__(64)
przm_sheaf_encode_kernel(const przm_meta_t *__restrict__ meta, uint8_t *__restrict__ sws, uint32_t nblocks) {
    __shared__ uint32_t vt[16][256];
    __shared__ __align__(16) uint8_t stg[16][RANS_STAGE + 16];
    const uint32_t q = threadIdx.x >> 2, j = threadIdx.x & 3u;
    const uint32_t job = blockIdx.x * 16u + q, blk = job >> 2, k = job & 3u;
    const bool has = blk < nblocks;
    uint8_t *sb = sws + (uint64_t)(has ? blk : 0u) * PRZM_SWS_BYTES;
    przm_sheaf_t *H = (przm_sheaf_t *)(sb + PRZM_SWS_HDR);
    uint32_t choice = 0, len = 0;
    if (has) {
        const przm_meta_t *mt = meta + blk;
        const uint32_t n = H->n;
        len = n;
        const uint32_t ev = mt->enc_len[PRZM_J_LIT];
        if (n && ev != PRZM_ENC_OVERFLOW && 8u + mt->mat_len[0] + 16u + ev < len) {
            len = 8u + mt->mat_len[0] + 16u + ev;
            choice = 1;
        }
        for (uint32_t m = 0; n && m < 4; m++) {
            const uint32_t sl = sheaf_len(H, m);
            if (sl < len) { len = sl; choice = 2 + m; }
        }
        if (k == 0 && j == 0) { H->choice = choice; H->len = len; }
    }
    /* the chosen mode's fiber k, when it is coded */
    const uint32_t m = choice >= 2 ? choice - 2 : 0u;
    const bool live = choice >= 2 && H->l[m][k] && H->k[m][k] != PRZM_ENC_OVERFLOW;
    const uint32_t n = live ? H->l[m][k] : 0u;
    uint32_t fs = 0, end = 0;
    if (choice >= 2) {
        for (uint32_t i = 0; i < k; i++) fs += al16(H->l[m][i]);
        for (uint32_t i = 0; i <= k; i++) end += al16(H->l[m][i] + 64u);
    }
    if (live) {
        const przm_ftab_t *ft = (const przm_ftab_t *)(sb + PRZM_SWS_FTAB) + 4u * m + k;
        for (uint32_t s = j; s < 256; s += 4) vt[q][s] = ft->fc[s];
    }
    uint32_t gmax = (n + 3u) >> 2;
    for (uint32_t o = 32; o >= 4; o >>= 1) {
        const uint32_t v = (uint32_t)__shfl_xor((int)gmax, (int)o, 64);
        gmax = v > gmax ? v : gmax;
    }
    __syncthreads();
    uint32_t x;
    const uint32_t K = rans_quad<false>(sb + PRZM_SWS_FIB + (uint64_t)m * PRZM_SWS_FIB_CAP + fs, n, vt[q], stg[q],
                                        sb + PRZM_SWS_OUT + end, n + 16u, gmax, &x);
    if (!live) return;
    if (j == 0) H->ek[k] = K;
    H->st[k][j] = x;
}
