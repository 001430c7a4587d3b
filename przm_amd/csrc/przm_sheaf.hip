/*
 * przm_sheaf.hip -- Section 4 with the sheaf trials on gfx950: the encoder
 * option that writes each frame's literals as PackageLiteralsSection chooses
 * them (core/literals/literals.go:667-693): the shortest of the raw literals,
 * the adaptive V-ANS container and the four sheaf containers.
 *
 *  przm_sheaf_split_kernel   one workgroup per block: each literal's fiber in
 *                            each mode from the class of the literals before it
 *                            (ContextClassify / ContextClassifyOrder2,
 *                            core/literals/sheaf.go:36-102); a counting pass
 *                            builds the 16 fiber lengths and histograms, the
 *                            waves normalise the fibers' tables
 *                            (NormalizeFrequencies and the matrix,
 *                            literals.go:78-184), and one scatter pass writes
 *                            the fibers of all four modes (DecomposeFibers,
 *                            sheaf.go:125-152), each 16-byte aligned.
 *  przm_sheaf_class_kernel,
 *  przm_sheaf_place_kernel   the counting jobs by fiber length.
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

#define SHEAF_ROUND (16u * SHEAF_NT) /* literals per round of the passes: 16 a thread */
#define SHEAF_STG (SHEAF_ROUND + 128u) /* LDS staging of one mode's fibers in a round */

/* each byte's kind for the context signature below: 0 white space, 1 '"' or ':', 2 other syntax
 * except '.' and '-', 3 identifier, 4 digit, 5 anything else, 6 '.' or '-' (the byte classes of
 * sheaf.go:36-102) */
struct SheafTable {
    uint8_t kind[256];
    static constexpr bool ws(int b) { return b == ' ' || b == '\t' || b == '\r' || b == '\n'; }
    static constexpr bool syntax(int b) {
        return b == '{' || b == '}' || b == '[' || b == ']' || b == '(' || b == ')' || b == '<' || b == '>' ||
               b == ':' || b == ';' || b == ',' || b == '.' || b == '"' || b == '\'' || b == '`' || b == '=' ||
               b == '+' || b == '-' || b == '*' || b == '/' || b == '%' || b == '&' || b == '|' || b == '^' ||
               b == '!' || b == '~' || b == '?' || b == '\\' || b == '#';
    }
    static constexpr bool ident(int b) { return (b >= 'a' && b <= 'z') || b == '_'; }
    static constexpr bool digit(int b) { return b >= '0' && b <= '9'; }
    constexpr SheafTable() : kind() {
        for (int b = 0; b < 256; b++)
            kind[b] = ws(b) ? 0 : (b == '"' || b == ':') ? 1 : (b == '.' || b == '-') ? 6 : syntax(b) ? 2
                    : ident(b) ? 3 : digit(b) ? 4 : 5;
    }
};
__constant__ SheafTable przm_sheaf_lut = SheafTable();

__device__ __forceinline__ uint32_t al16(uint32_t v) { return (v + 15u) & ~15u; }

/* byte i of the literals, 0 before the first */
__device__ __forceinline__ uint32_t lit_at(const uint8_t *L, int64_t i) { return i >= 0 ? L[i] : 0u; }

/* a fiber shorter than this is never coded: its coded form has a matrix of at least 5 bytes
 * (row mask, one column mask, one weight) and 16 bytes of states (literals.go:622) */
#define SHEAF_MIN_CODED 22u

/*
 * A literal's fibers in all four modes depend on its two previous literals only through one of
 * eight context signatures: 0 after white space, 1 after '"' or ':', 2 after other syntax, 3 / 4
 * after an identifier byte preceded by one / by anything else, 5 / 6 after a digit preceded by a
 * digit, '.' or '-' / by anything else, 7 after anything else.  Its fiber in mode m is
 * sheaf_fiber(m, signature): 2 bits per signature of the mode's table (mode 0 ws 0, syntax 1,
 * identifier 2, other 3; mode 1 '"' ':' 0, identifier 1, digit 2, other 3; mode 2 ws and syntax 0,
 * other 1; mode 3 ContextClassifyOrder2, sheaf.go:86-102).
 */
#define SHEAF_T0 0xFE94u /* 3 3 3 2 2 1 1 0 (signatures 7..0) */
#define SHEAF_T1 0xE973u /* 3 2 2 1 1 3 0 3 */
#define SHEAF_T2 0x5540u /* 1 1 1 1 1 0 0 0 */
#define SHEAF_T3 0x5394u /* 1 1 0 3 2 1 1 0 */
__host__ __device__ constexpr uint32_t sheaf_fiber(uint32_t m, uint32_t sig) {
    return ((m == 0 ? SHEAF_T0 : m == 1 ? SHEAF_T1 : m == 2 ? SHEAF_T2 : SHEAF_T3) >> (2u * sig)) & 3u;
}
/* the signature after a literal of kind k1 preceded by one of kind k2 */
__host__ __device__ constexpr uint32_t sheaf_sig(uint32_t k1, uint32_t k2) {
    return k1 == 6u ? 2u : k1 < 3u ? k1 : k1 == 3u ? (k2 == 3u ? 3u : 4u)
         : k1 == 4u ? ((k2 == 4u || k2 == 6u) ? 5u : 6u) : 7u;
}
/* the word of kinds (k1, k2), at sigw[8 k1 + k2]: bits 6m .. 6m + 5 hold 16 * the fiber in mode
 * m, bits 24 .. 29 8 * the signature (the shifts of packed 16-bit and 8-bit counters) */
__host__ __device__ constexpr uint32_t sheaf_word(uint32_t k1, uint32_t k2) {
    const uint32_t s = sheaf_sig(k1, k2);
    return (8u * s) << 24 | (16u * sheaf_fiber(0, s)) | (16u * sheaf_fiber(1, s)) << 6 |
           (16u * sheaf_fiber(2, s)) << 12 | (16u * sheaf_fiber(3, s)) << 18;
}

/* the 16 literals of a thread in a round (FULL: all of them live) and the table words of their
 * contexts; returns the number live */
template <bool FULL>
__device__ __forceinline__ uint32_t sheaf_load16(const uint8_t *__restrict__ L, uint32_t n, uint32_t i0,
                                                 const uint8_t *kind, const uint32_t *sigw, uint32_t wd[4],
                                                 uint32_t sw[16]) {
    const bool any = FULL || i0 < n;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (any) v = *(const uint4 *)(L + i0);
    wd[0] = v.x; wd[1] = v.y; wd[2] = v.z; wd[3] = v.w;
    uint32_t k2 = kind[any ? lit_at(L, (int64_t)i0 - 2) : 0u], k1 = kind[any ? lit_at(L, (int64_t)i0 - 1) : 0u];
#pragma unroll
    for (uint32_t t = 0; t < 16u; t++) {
        sw[t] = sigw[8u * k1 + k2];
        k2 = k1;
        k1 = kind[(wd[t >> 2] >> (8u * (t & 3u))) & 0xFFu];
    }
    return FULL ? 16u : any ? (n - i0 < 16u ? n - i0 : 16u) : 0u;
}

/* the counting pass's round: each live literal into the histogram of its signature */
template <bool FULL>
__device__ __forceinline__ void sheaf_hist_round(const uint8_t *__restrict__ L, uint32_t n, uint32_t i0,
                                                 const uint8_t *kind, const uint32_t *sigw, uint32_t *hs) {
    uint32_t wd[4], sw[16];
    const uint32_t live = sheaf_load16<FULL>(L, n, i0, kind, sigw, wd, sw);
#pragma unroll
    for (uint32_t t = 0; t < 16u; t++)
        if (FULL || t < live) atomicAdd(&hs[((sw[t] >> 24) << 5) + ((wd[t >> 2] >> (8u * (t & 3u))) & 0xFFu)], 1u);
}

/*
 * The per-fiber state of the scatter pass, in LDS: fdst[f] is the offset in its mode's fiber area
 * where fiber f's next 16-byte word goes, fcy[f] the number of its bytes carried over (below 16,
 * held by thread f), freg[f] / fnw[f] the round's staging region of the fiber (16-byte aligned in
 * its mode's staging area: the carried bytes, then the round's literals) and its whole words.
 */
struct SheafFib {
    uint32_t dst[16], cy[16], reg[16], nw[16];
};

/*
 * A round of the scatter pass, all four modes at once: each live literal goes to its fiber in
 * every mode in order.  A literal's place among the round's literals of its fiber is the number of
 * earlier ones of that fiber: the per-signature counts of the thread (8 bits each) are scanned
 * over the workgroup as 16-bit fields, and the fiber's count is the sum over the signatures that
 * map to it.  The literals are written to LDS staging, then the round's whole 16-byte words of
 * every fiber are stored; the last partial word of a fiber is carried into the next round.
 */
template <bool FULL>
__device__ __forceinline__ void sheaf_scatter_round(const uint8_t *__restrict__ L, uint32_t n, uint32_t base,
                                                    const uint8_t *kind, const uint32_t *sigw, uint8_t *stg,
                                                    uint32_t (*part)[4], SheafFib &F, uint8_t *__restrict__ fib,
                                                    uint4 &cw, uint32_t &my_dst, uint32_t &my_cy) {
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63u;
    uint32_t wd[4], sw[16];
    const uint32_t live = sheaf_load16<FULL>(L, n, base + 16u * tid, kind, sigw, wd, sw);
    uint64_t sc = 0;
#pragma unroll
    for (uint32_t t = 0; t < 16u; t++)
        if (FULL || t < live) sc += 1ull << (sw[t] >> 24);
    const uint32_t lo = (uint32_t)sc, hi = (uint32_t)(sc >> 32);
    const uint32_t c4[4] = {(lo & 0xFFu) | ((lo & 0xFF00u) << 8), ((lo >> 16) & 0xFFu) | ((lo >> 8) & 0xFF0000u),
                            (hi & 0xFFu) | ((hi & 0xFF00u) << 8), ((hi >> 16) & 0xFFu) | ((hi >> 8) & 0xFF0000u)};
    uint32_t in4[4];
#pragma unroll
    for (int i = 0; i < 4; i++) in4[i] = wave_scan_incl(c4[i]);
    if (lane == 63u) *(uint4 *)part[w] = make_uint4(in4[0], in4[1], in4[2], in4[3]);
    __syncthreads();
    uint32_t E[8], T[8];
    {
        uint32_t pre[4] = {0, 0, 0, 0}, tot[4] = {0, 0, 0, 0};
#pragma unroll
        for (uint32_t k = 0; k < SHEAF_NT / 64u; k++) {
            const uint4 pk = *(const uint4 *)part[k];
            const uint32_t a[4] = {pk.x, pk.y, pk.z, pk.w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                pre[i] += k < w ? a[i] : 0u;
                tot[i] += a[i];
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t e = pre[i] + in4[i] - c4[i];
            E[2 * i] = e & 0xFFFFu;
            E[2 * i + 1] = e >> 16;
            T[2 * i] = tot[i] & 0xFFFFu;
            T[2 * i + 1] = tot[i] >> 16;
        }
    }
    /* per mode: the fibers' round totals, staging regions and this thread's first places */
    uint64_t P[4];
#pragma unroll
    for (uint32_t m = 0; m < 4; m++) {
        uint32_t tf[4] = {0, 0, 0, 0}, ps[4] = {0, 0, 0, 0};
#pragma unroll
        for (uint32_t s = 0; s < 8; s++) {
            tf[sheaf_fiber(m, s)] += T[s];
            ps[sheaf_fiber(m, s)] += E[s];
        }
        uint32_t off = 0;
        P[m] = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t cy = F.cy[4u * m + k], r = off;
            P[m] |= (uint64_t)(r + cy + ps[k]) << (16u * k);
            off += al16(cy + tf[k]);
            if (tid == 4u * m + k) { /* the fiber's carry thread: its region, words and carried bytes */
                F.reg[tid] = r;
                F.nw[tid] = (cy + tf[k]) >> 4;
                const uint32_t cwd[4] = {cw.x, cw.y, cw.z, cw.w};
                uint8_t *d = stg + m * SHEAF_STG + r;
#pragma unroll
                for (uint32_t b = 0; b < 15u; b++)
                    if (b < cy) d[b] = (uint8_t)(cwd[b >> 2] >> (8u * (b & 3u)));
                my_cy = (cy + tf[k]) & 15u;
            }
        }
    }
#pragma unroll
    for (uint32_t t = 0; t < 16u; t++) {
        if (!(FULL || t < live)) continue;
        const uint8_t b = (uint8_t)(wd[t >> 2] >> (8u * (t & 3u)));
#pragma unroll
        for (uint32_t m = 0; m < 4; m++) {
            const uint32_t sh = (sw[t] >> (6u * m)) & 63u;
            stg[m * SHEAF_STG + ((uint32_t)(P[m] >> sh) & 0xFFFFu)] = b;
            P[m] += 1ull << sh;
        }
    }
    __syncthreads();
    /* the whole words of every fiber; fiber f's carry thread keeps its partial last word */
#pragma unroll
    for (uint32_t m = 0; m < 4; m++) {
        const uint32_t n0 = F.nw[4u * m], n1 = n0 + F.nw[4u * m + 1u], n2 = n1 + F.nw[4u * m + 2u],
                       W = n2 + F.nw[4u * m + 3u];
        for (uint32_t g = tid; g < W; g += SHEAF_NT) {
            const uint32_t k = (g >= n0 ? 1u : 0u) + (g >= n1 ? 1u : 0u) + (g >= n2 ? 1u : 0u);
            const uint32_t gb = k == 0 ? 0u : k == 1 ? n0 : k == 2 ? n1 : n2, f = 4u * m + k;
            const uint4 x = *(const uint4 *)(stg + m * SHEAF_STG + F.reg[f] + 16u * (g - gb));
            *(uint4 *)(fib + (uint64_t)m * PRZM_SWS_FIB_CAP + F.dst[f] + 16u * (g - gb)) = x;
        }
    }
    if (tid < 16u) {
        const uint32_t m = tid >> 2;
        if (my_cy) cw = *(const uint4 *)(stg + m * SHEAF_STG + F.reg[tid] + 16u * F.nw[tid]);
        my_dst = F.dst[tid] + 16u * F.nw[tid];
    }
    __syncthreads();
    if (tid < 16u) {
        F.dst[tid] = my_dst;
        F.cy[tid] = my_cy;
    }
}

/*
 * One workgroup per block, all four modes: a counting pass over the literals (16 a thread and
 * round, loaded as one 16-byte word) fills one histogram per context signature; the 16
 * histograms (mode, fiber) are sums of those rows, the fiber lengths come from them, and each
 * wave normalises the tables of four of them (the fibers long enough to be coded); a scatter
 * pass then writes the fibers of all four modes, each 16-byte aligned in its mode's area, the
 * bytes after its end up to the next 16-byte boundary zero.
 */
extern "C" __global__ void __launch_bounds__(SHEAF_NT)
przm_sheaf_split_kernel(const uint8_t *__restrict__ ws, const przm_meta_t *__restrict__ meta,
                        uint8_t *__restrict__ sws, uint32_t nblocks) {
    __shared__ uint8_t kind[256];
    __shared__ uint32_t sigw[64];
    /* hs[8][256] (per signature) and cnt[16][256] (per mode and fiber); the scatter pass's staging */
    __shared__ __align__(16) uint32_t pool[24u * 256u];
    __shared__ __align__(16) uint32_t part[SHEAF_NT / 64u][4];
    __shared__ uint32_t flen[16];
    __shared__ SheafFib F;
    static_assert(4u * SHEAF_STG <= sizeof(pool), "staging fits the histogram space");
    uint32_t *hs = pool;
    uint32_t (*cnt)[256] = (uint32_t (*)[256])(pool + 8u * 256u);
    const uint32_t blk = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lane = tid & 63u;
    if (blk >= nblocks) return;
    const uint8_t *L = ws + (uint64_t)blk * PRZM_WS_BYTES + PRZM_WS_LIT; /* 16-byte aligned */
    const uint32_t n = meta[blk].len[PRZM_S_LIT];
    uint8_t *sb = sws + (uint64_t)blk * PRZM_SWS_BYTES;
    przm_sheaf_t *H = (przm_sheaf_t *)(sb + PRZM_SWS_HDR);
    for (uint32_t i = tid; i < 256; i += SHEAF_NT) kind[i] = przm_sheaf_lut.kind[i];
    if (tid < 64u) sigw[tid] = sheaf_word(tid >> 3, tid & 7u);
    for (uint32_t i = tid; i < 8 * 256 / 4; i += SHEAF_NT) ((uint4 *)hs)[i] = make_uint4(0, 0, 0, 0);
    __syncthreads();
    /* counting pass: one histogram per context signature */
    for (uint32_t base = 0; base < n; base += SHEAF_ROUND) {
        const uint32_t i0 = base + 16u * tid;
        if (i0 + 16u <= n) sheaf_hist_round<true>(L, n, i0, kind, sigw, hs);
        else if (i0 < n) sheaf_hist_round<false>(L, n, i0, kind, sigw, hs);
    }
    __syncthreads();
    /* the (mode, fiber) histograms: thread s sums bin s over the signatures of each fiber */
    for (uint32_t sb_ = tid; sb_ < 256; sb_ += SHEAF_NT) {
        uint32_t acc[16] = {};
#pragma unroll
        for (uint32_t g = 0; g < 8; g++) {
            const uint32_t h = hs[g * 256u + sb_];
            acc[sheaf_fiber(0, g)] += h;
            acc[4u + sheaf_fiber(1, g)] += h;
            acc[8u + sheaf_fiber(2, g)] += h;
            acc[12u + sheaf_fiber(3, g)] += h;
        }
#pragma unroll
        for (uint32_t r = 0; r < 16; r++) cnt[r][sb_] = acc[r];
    }
    __syncthreads();
    /* fiber lengths: wave w sums rows w, w + 4, w + 8, w + 12 */
    for (uint32_t r = w; r < 16; r += SHEAF_NT / 64u) {
        uint32_t l = 0;
        for (uint32_t s = lane; s < 256; s += 64) l += cnt[r][s];
        l = wave_sum(l);
        if (lane == 0) flen[r] = l;
    }
    __syncthreads();
    if (tid < 16) {
        if (tid == 0) H->n = n;
        H->l[tid >> 2][tid & 3u] = flen[tid];
        if (flen[tid] < SHEAF_MIN_CODED) H->mat[tid >> 2][tid & 3u] = 0; /* the fiber's wave writes it otherwise */
        uint32_t fs = 0; /* the fiber's place in its mode's area: the 16-byte rounded lengths before it */
        for (uint32_t k = 0; k < (tid & 3u); k++) fs += al16(flen[(tid & ~3u) + k]);
        F.dst[tid] = fs;
        F.cy[tid] = 0;
    }
    for (uint32_t r = w; r < 16; r += SHEAF_NT / 64u)
        if (flen[r] >= SHEAF_MIN_CODED)
            normalize_wave(cnt[r], (przm_ftab_t *)(sb + PRZM_SWS_FTAB) + r, &H->mat[r >> 2][r & 3u]);
    __syncthreads(); /* the staging below reuses the histograms' space */
    /* scatter pass */
    uint8_t *stg = (uint8_t *)pool, *fib = sb + PRZM_SWS_FIB;
    uint4 cw = make_uint4(0, 0, 0, 0);
    uint32_t my_dst = 0, my_cy = 0;
    for (uint32_t base = 0; base < n; base += SHEAF_ROUND) {
        if (base + SHEAF_ROUND <= n)
            sheaf_scatter_round<true>(L, n, base, kind, sigw, stg, part, F, fib, cw, my_dst, my_cy);
        else
            sheaf_scatter_round<false>(L, n, base, kind, sigw, stg, part, F, fib, cw, my_dst, my_cy);
    }
    if (tid < 16u && my_cy) { /* the fiber's last partial word, zero after its end */
        uint32_t cwd[4] = {cw.x, cw.y, cw.z, cw.w};
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) {
            const uint32_t keep = my_cy >= 4u * q + 4u ? 4u : my_cy > 4u * q ? my_cy - 4u * q : 0u;
            cwd[q] &= keep == 4u ? 0xFFFFFFFFu : (1u << (8u * keep)) - 1u;
        }
        *(uint4 *)(fib + (uint64_t)(tid >> 2) * PRZM_SWS_FIB_CAP + my_dst) = make_uint4(cwd[0], cwd[1], cwd[2], cwd[3]);
    }
}

/*
 * The counting jobs (block, mode, fiber) by fiber length in 256-byte classes, longest first, so
 * that a wave of przm_sheaf_count_kernel steps as often as its longest job: one thread per job,
 * 256 jobs (16 blocks) a workgroup.  przm_sheaf_class_kernel adds the workgroup's class counts
 * into bins[0, 256); przm_sheaf_place_kernel takes the exclusive scan of those, reserves the
 * workgroup's run in each class with bins[256 + class], and places its jobs there.  bins start at
 * zero.  The order inside a class is whatever the atomics give: every job writes only its own
 * outputs, so the results do not depend on it.
 */
__device__ __forceinline__ uint32_t sheaf_job_class(const uint8_t *__restrict__ sws, uint32_t j) {
    const przm_sheaf_t *H = (const przm_sheaf_t *)(sws + (uint64_t)(j >> 4) * PRZM_SWS_BYTES + PRZM_SWS_HDR);
    const uint32_t c = H->l[(j >> 2) & 3u][j & 3u] >> 8;
    return 255u - (c < 255u ? c : 255u);
}

extern "C" __global__ void __launch_bounds__(256)
przm_sheaf_class_kernel(const uint8_t *__restrict__ sws, uint32_t nblocks, uint32_t *__restrict__ bins) {
    __shared__ uint32_t cnt[256];
    const uint32_t tid = threadIdx.x, j = blockIdx.x * 256u + tid;
    cnt[tid] = 0;
    __syncthreads();
    if (j < 16u * nblocks) atomicAdd(&cnt[sheaf_job_class(sws, j)], 1u);
    __syncthreads();
    if (cnt[tid]) atomicAdd(&bins[tid], cnt[tid]);
}

extern "C" __global__ void __launch_bounds__(256)
przm_sheaf_place_kernel(const uint8_t *__restrict__ sws, uint32_t nblocks, uint32_t *__restrict__ bins,
                        uint32_t *__restrict__ perm) {
    __shared__ uint32_t cnt[256], base[256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, j = blockIdx.x * 256u + tid;
    const bool has = j < 16u * nblocks;
    cnt[tid] = 0;
    __syncthreads();
    const uint32_t c = has ? sheaf_job_class(sws, j) : 0u;
    const uint32_t r = has ? atomicAdd(&cnt[c], 1u) : 0u; /* the job's rank in the workgroup's run */
    if (tid < 64u) { /* exclusive scan of the class counts, four per lane */
        uint32_t c4[4], t = 0;
        for (int k = 0; k < 4; k++) { c4[k] = bins[lane * 4 + k]; t += c4[k]; }
        uint32_t run = wave_scan_incl(t) - t;
        for (int k = 0; k < 4; k++) { base[lane * 4 + k] = run; run += c4[k]; }
    }
    __syncthreads();
    if (cnt[tid]) base[tid] += atomicAdd(&bins[256u + tid], cnt[tid]);
    __syncthreads();
    if (has) perm[base[c] + r] = j;
}

/* one counting step of quad lane j for step g with table entry e (freq | cdf << 16): the coder
 * step of literals.go:305-325 without its stores, counting the bytes its renormalisation emits in
 * cnt.  LIVE: the step may lie past the symbols below nl; such a step codes the entry freq 4096,
 * cdf 0, which leaves every state below 2^24 as it is and emits nothing, so no lane branches. */
template <bool LIVE>
__device__ __forceinline__ void count_step(uint32_t g, uint32_t e, uint32_t j, uint32_t nl, uint32_t &x,
                                           uint32_t &cnt) {
    if (LIVE) e = 4u * g + j < nl ? e : 4096u;
    const uint32_t f = e & 0xFFFFu, t = 4096u * f - 1u;
    const bool r1 = x > t;
    const uint32_t x1 = r1 ? x >> 8 : x;
    const bool r2 = r1 && x1 > t;
    const uint32_t xn = r2 ? x1 >> 8 : x1;
    cnt += (uint32_t)r1 + (uint32_t)r2;
    uint32_t qq, rr;
    divmod(xn, f, &qq, &rr);
    x = qq * 4096u + (e >> 16) + rr;
}

/* one turn of the chunk ring (RANS_PF chunks from cb down): the steps of the ring's chunks, the
 * table entries of each chunk's successor read while it is coded, the ring's slots reloaded */
template <bool LIVE>
__device__ __forceinline__ void count_turn(const uint4 *__restrict__ src4, int32_t own, int32_t cb,
                                           const uint32_t *__restrict__ tab, uint32_t j, uint32_t nl, uint4 (&buf)[RANS_PF],
                                           uint32_t (&e)[4], uint32_t &x, uint32_t &cnt) {
    auto ent = [&](uint32_t dw) -> uint32_t { return tab[(dw >> (8u * j)) & 0xFFu]; };
#pragma unroll
    for (uint32_t d = 0; d < RANS_PF; d++) {
        const int32_t c = cb - (int32_t)d;
        const uint4 nb = buf[(d + 1) % RANS_PF];
        const uint32_t en[4] = {ent(nb.w), ent(nb.z), ent(nb.y), ent(nb.x)};
        const uint32_t g0 = 4u * (uint32_t)c;
        count_step<LIVE>(g0 + 3, e[0], j, nl, x, cnt);
        count_step<LIVE>(g0 + 2, e[1], j, nl, x, cnt);
        count_step<LIVE>(g0 + 1, e[2], j, nl, x, cnt);
        count_step<LIVE>(g0, e[3], j, nl, x, cnt);
#pragma unroll
        for (int u = 0; u < 4; u++) e[u] = en[u];
        const int32_t cn = c - (int32_t)RANS_PF;
        buf[d] = *(cn >= 0 && cn < own ? src4 + cn : &przm_zero16);
    }
}

/*
 * Stream bytes of the V-ANS coding of fiber src[0, n) with table tab, by one lane quad: lane
 * j runs state j over the symbols i = j (mod 4) from the top (literals.go:302-325), counting
 * the bytes its renormalisation emits.  Once the quad has emitted `limit` bytes the fiber
 * cannot be shorter coded than raw: PRZM_ENC_OVERFLOW.
 *
 * The chunks are read as rans_quad reads them: RANS_PF 16-byte chunks in flight in a register
 * ring, every load issued (outside the fiber from przm_zero16), every quad of the wave stepping
 * as often as the wave's longest fiber (gmax, in 4-symbol groups) in uniform control flow.  A
 * turn of the ring below nmin, the wave's shortest fiber, has every step live and takes no
 * liveness test (a quad past its limit keeps stepping there: its result is overflow whatever
 * its state).  The limit is checked once per turn (the count only grows), and the wave stops
 * when all its quads are past it.
 */
__device__ uint32_t vans_count_quad(const uint8_t *__restrict__ src, uint32_t n, const uint32_t *__restrict__ tab,
                                    uint32_t limit, uint32_t gmax, uint32_t nmin) {
    const uint32_t j = threadIdx.x & 3u;
    const int32_t own = (int32_t)((n + 15u) >> 4);
    const int32_t NC = (int32_t)((((gmax + 3u) >> 2) + RANS_PF - 1u) / RANS_PF * RANS_PF);
    const uint4 *src4 = (const uint4 *)src;
    auto ent = [&](uint32_t dw) -> uint32_t { return tab[(dw >> (8u * j)) & 0xFFu]; };
    uint32_t x = 65536u, cnt = 0, nl = n;
    bool over = false;
    uint4 buf[RANS_PF];
#pragma unroll
    for (uint32_t d = 0; d < RANS_PF; d++) {
        const int32_t c = NC - 1 - (int32_t)d;
        buf[d] = *(c >= 0 && c < own ? src4 + c : &przm_zero16);
    }
    uint32_t e[4] = {ent(buf[0].w), ent(buf[0].z), ent(buf[0].y), ent(buf[0].x)};
    for (int32_t cb = NC - 1; cb >= 0; cb -= (int32_t)RANS_PF) {
        if (16u * (uint32_t)cb + 16u <= nmin)
            count_turn<false>(src4, own, cb, tab, j, nl, buf, e, x, cnt);
        else
            count_turn<true>(src4, own, cb, tab, j, nl, buf, e, x, cnt);
        uint32_t tot = cnt + QPERM(cnt, 0xB1);
        tot += QPERM(tot, 0x4E);
        if (tot >= limit) { over = true; nl = 0; }
        if (__all(over)) break;
    }
    uint32_t tot = cnt + QPERM(cnt, 0xB1);
    tot += QPERM(tot, 0x4E);
    return over || tot >= limit ? PRZM_ENC_OVERFLOW : tot;
}

/* one quad per (block, mode, fiber) job, 16 jobs a workgroup in the order perm gives
 * (przm_sheaf_place_kernel); job = block * 16 + mode * 4 + fiber */
extern "C" __global__ void __launch_bounds__(64)
przm_sheaf_count_kernel(uint8_t *__restrict__ sws, uint32_t nblocks, const uint32_t *__restrict__ perm) {
    __shared__ uint32_t vt[16][256];
    const uint32_t q = threadIdx.x >> 2, j = threadIdx.x & 3u;
    const uint32_t slot = blockIdx.x * 16u + q;
    const bool has = slot < 16u * nblocks;
    const uint32_t job = has ? perm[slot] : 0u, blk = job >> 4, m = (job >> 2) & 3u, k = job & 3u;
    uint8_t *sb = sws + (uint64_t)(has ? blk : 0u) * PRZM_SWS_BYTES;
    przm_sheaf_t *H = (przm_sheaf_t *)(sb + PRZM_SWS_HDR);
    uint32_t n = 0, fs = 0, mat = 0;
    if (has) {
        n = H->l[m][k];
        for (uint32_t i = 0; i < k; i++) fs += al16(H->l[m][i]);
        mat = H->mat[m][k];
        if (n >= SHEAF_MIN_CODED) {
            const przm_ftab_t *ft = (const przm_ftab_t *)(sb + PRZM_SWS_FTAB) + 4u * m + k;
            for (uint32_t s = j; s < 256; s += 4) vt[q][s] = ft->fc[s];
        }
    }
    /* coded = matrix + 4 states + stream, taken only when shorter than the fiber */
    const uint32_t head = mat + 16u;
    const bool coded = has && n >= SHEAF_MIN_CODED && n > head;
    /* the wave's longest and shortest coded fibers */
    uint32_t gmax = coded ? (n + 3u) >> 2 : 0u, nmin = coded ? n : 0xFFFFFFFFu;
    for (uint32_t o = 32; o >= 4; o >>= 1) {
        const uint32_t v = (uint32_t)__shfl_xor((int)gmax, (int)o, 64);
        const uint32_t u = (uint32_t)__shfl_xor((int)nmin, (int)o, 64);
        gmax = v > gmax ? v : gmax;
        nmin = u < nmin ? u : nmin;
    }
    __syncthreads();
    if (!has) return;
    /* every quad with a job steps, in uniform control flow: one whose fiber is not coded counts
     * an empty fiber with limit 0 (past it at once; in a turn without liveness tests it steps on
     * zero chunks with whatever its table slot holds, which touches no memory but that slot) */
    const uint32_t K = vans_count_quad(sb + PRZM_SWS_FIB + (uint64_t)m * PRZM_SWS_FIB_CAP + fs, coded ? n : 0u,
                                       vt[q], coded ? n - head : 0u, gmax, nmin);
    if (j == 0) H->k[m][k] = coded ? K : PRZM_ENC_OVERFLOW;
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
