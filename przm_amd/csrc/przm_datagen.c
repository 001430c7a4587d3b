/*
 * przm_datagen.c -- seeded synthetic inputs for the PRZM encode workloads.
 *
 * Every byte is a pure function of (seed, position), computed per 1 MiB
 * chunk (Markov source) or per 64 KiB block (mixture), so the output does
 * not depend on the thread count.
 *
 * PRNG: SplitMix64 (Steele, Lea, Flood 2014) seeding xoshiro256**
 * (Blackman, Vigna 2018).
 *
 *   przm_gen_markov   order-1 Markov chain over the 95 printable bytes
 *                     0x20..0x7E, each row a Zipf(s=1.1) law over a per-row
 *                     random permutation; then 30% of positions are
 *                     overwritten by earlier substrings: length 4 +
 *                     Geometric(1/21) (mean 24), distance log-uniform in
 *                     [1, 65536), copied forward byte by byte.
 *   przm_gen_mixture  per 64 KiB block, one of four contents in fixed
 *                     proportions 4:3:2:1 inside every group of 10 blocks:
 *                     JSON-like records (200-word vocabulary, ints in
 *                     [0, 1e6], N(0,1) floats printed "%.6f"), little-endian
 *                     int32 tables (8 columns, Laplace(b=3) deltas), random
 *                     bytes with repeated 16-byte motifs at 5% density, and
 *                     uniform random bytes.
 *   przm_gen_loguniform_sizes  sizes log-uniform in [lo, hi].
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static void rng_seed(rng_t *r, uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rng_next(rng_t *r) {
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return res;
}
static inline double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static inline uint64_t rng_below(rng_t *r, uint64_t n) { return n ? (uint64_t)(rng_unit(r) * (double)n) % n : 0; }

/* ---------------- Markov source ---------------- */

#define NSYM 95
typedef struct { uint8_t perm[NSYM][NSYM]; double cdf[NSYM][NSYM]; } markov_t;

static void markov_init(markov_t *m, uint64_t seed) {
    double w[NSYM], tot = 0;
    for (int k = 0; k < NSYM; k++) { w[k] = 1.0 / pow((double)(k + 1), 1.1); tot += w[k]; }
    rng_t r;
    rng_seed(&r, seed, 0xA11CEull);
    for (int row = 0; row < NSYM; row++) {
        for (int k = 0; k < NSYM; k++) m->perm[row][k] = (uint8_t)k;
        for (int k = NSYM - 1; k > 0; k--) {
            int j = (int)rng_below(&r, (uint64_t)k + 1);
            uint8_t t = m->perm[row][k]; m->perm[row][k] = m->perm[row][j]; m->perm[row][j] = t;
        }
        double c = 0;
        for (int k = 0; k < NSYM; k++) { c += w[k] / tot; m->cdf[row][k] = c; }
        m->cdf[row][NSYM - 1] = 1.0;
    }
}

static void markov_chunk(const markov_t *m, uint8_t *out, uint64_t n, uint64_t seed, uint64_t chunk) {
    rng_t r;
    rng_seed(&r, seed, 0x100000ull + chunk);
    int prev = (int)rng_below(&r, NSYM);
    for (uint64_t i = 0; i < n; i++) {
        double u = rng_unit(&r);
        int lo = 0, hi = NSYM - 1;
        while (lo < hi) { int mid = (lo + hi) >> 1; if (m->cdf[prev][mid] > u) hi = mid; else lo = mid + 1; }
        int sym = m->perm[prev][lo];
        out[i] = (uint8_t)(0x20 + sym);
        prev = sym;
    }
    uint64_t target = (n * 3) / 10, covered = 0;
    const double ln64k = log(65536.0);
    while (covered < target && n > 1) {
        uint64_t p = 1 + rng_below(&r, n - 1);
        uint64_t len = 4;
        while (rng_unit(&r) >= 1.0 / 21.0) len++;
        uint64_t d = (uint64_t)floor(exp(rng_unit(&r) * ln64k));
        if (d < 1) d = 1;
        if (d > p) d = p;
        if (d > 65536) d = 65536;
        if (p + len > n) len = n - p;
        for (uint64_t k = 0; k < len; k++) out[p + k] = out[p - d + k];
        covered += len;
    }
}

/* ---------------- mixture ---------------- */

#define VOCAB 200
static char g_vocab[VOCAB][12];
static pthread_once_t g_vocab_once = PTHREAD_ONCE_INIT;
static void vocab_init(void) {
    rng_t r;
    rng_seed(&r, 0x5EEDF00Dull, 0xC0CABull);
    for (int i = 0; i < VOCAB; i++) {
        int len = 3 + (int)rng_below(&r, 8);
        for (int k = 0; k < len; k++) g_vocab[i][k] = (char)('a' + rng_below(&r, 26));
        g_vocab[i][len] = 0;
    }
}

static double rng_normal(rng_t *r) { /* Box-Muller */
    double u1 = rng_unit(r), u2 = rng_unit(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static void gen_json(uint8_t *out, uint64_t n, rng_t *r) {
    uint64_t pos = 0;
    char rec[512];
    while (pos < n) {
        int len = 0, nf = 3 + (int)rng_below(r, 6);
        rec[len++] = '{';
        for (int f = 0; f < nf; f++) {
            if (f) rec[len++] = ',';
            len += snprintf(rec + len, sizeof(rec) - (size_t)len, "\"%s\":", g_vocab[rng_below(r, VOCAB)]);
            int kind = (int)rng_below(r, 3);
            if (kind == 0) len += snprintf(rec + len, sizeof(rec) - (size_t)len, "%u", (unsigned)rng_below(r, 1000001));
            else if (kind == 1) len += snprintf(rec + len, sizeof(rec) - (size_t)len, "%.6f", rng_normal(r));
            else len += snprintf(rec + len, sizeof(rec) - (size_t)len, "\"%s\"", g_vocab[rng_below(r, VOCAB)]);
        }
        rec[len++] = '}';
        rec[len++] = '\n';
        uint64_t take = (uint64_t)len < n - pos ? (uint64_t)len : n - pos;
        memcpy(out + pos, rec, take);
        pos += take;
    }
}

static void gen_table(uint8_t *out, uint64_t n, rng_t *r) {
    int32_t col[8];
    for (int j = 0; j < 8; j++) col[j] = (int32_t)rng_below(r, 1000000);
    uint64_t pos = 0;
    while (pos < n) {
        for (int j = 0; j < 8 && pos < n; j++) {
            double u = rng_unit(r) - 0.5;
            double lap = (u < 0 ? 1.0 : -1.0) * 3.0 * log(1.0 - 2.0 * fabs(u) + 1e-300);
            col[j] += (int32_t)lround(lap);
            uint32_t v = (uint32_t)col[j];
            for (int k = 0; k < 4 && pos < n; k++) out[pos++] = (uint8_t)(v >> (8 * k));
        }
    }
}

static void gen_random(uint8_t *out, uint64_t n, rng_t *r) {
    uint64_t i = 0;
    for (; i + 8 <= n; i += 8) { uint64_t v = rng_next(r); memcpy(out + i, &v, 8); }
    if (i < n) { uint64_t v = rng_next(r); memcpy(out + i, &v, n - i); }
}

static void gen_motif(uint8_t *out, uint64_t n, rng_t *r) {
    uint8_t motifs[64][16];
    gen_random(&motifs[0][0], sizeof(motifs), r);
    gen_random(out, n, r);
    uint64_t target = n / 20, covered = 0;
    while (covered < target && n >= 16) {
        uint64_t p = rng_below(r, n - 15);
        memcpy(out + p, motifs[rng_below(r, 64)], 16);
        covered += 16;
    }
}

/* kind of block b of a mixture stream: 4 JSON, 3 table, 2 motif, 1 random per group of 10 */
static int mixture_kind(uint64_t seed, uint64_t b) {
    static const int base[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
    int perm[10];
    memcpy(perm, base, sizeof(perm));
    rng_t r;
    rng_seed(&r, seed, 0x200000000ull + b / 10);
    for (int k = 9; k > 0; k--) {
        int j = (int)rng_below(&r, (uint64_t)k + 1);
        int t = perm[k]; perm[k] = perm[j]; perm[j] = t;
    }
    return perm[b % 10];
}

static void mixture_block(uint8_t *out, uint64_t n, uint64_t seed, uint64_t b) {
    rng_t r;
    rng_seed(&r, seed, 0x300000000ull + b);
    switch (mixture_kind(seed, b)) {
    case 0: gen_json(out, n, &r); break;
    case 1: gen_table(out, n, &r); break;
    case 2: gen_motif(out, n, &r); break;
    default: gen_random(out, n, &r); break;
    }
}

/* ---------------- threading ---------------- */

typedef struct {
    int kind; /* 0 markov, 1 mixture */
    uint8_t *out; uint64_t n, unit, seed;
    const markov_t *m;
    uint64_t next;
    pthread_mutex_t mu;
} job_t;

static void *worker(void *arg) {
    job_t *j = (job_t *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        uint64_t u = j->next++;
        pthread_mutex_unlock(&j->mu);
        uint64_t off = u * j->unit;
        if (off >= j->n) break;
        uint64_t len = j->n - off < j->unit ? j->n - off : j->unit;
        if (j->kind == 0) markov_chunk(j->m, j->out + off, len, j->seed, u);
        else mixture_block(j->out + off, len, j->seed, u);
    }
    return NULL;
}

static void run(job_t *j, int threads) {
    if (threads <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        threads = c > 16 ? 16 : (c < 1 ? 1 : (int)c);
    }
    uint64_t units = (j->n + j->unit - 1) / j->unit;
    if ((uint64_t)threads > units) threads = units ? (int)units : 1;
    pthread_mutex_init(&j->mu, NULL);
    j->next = 0;
    pthread_t tid[64];
    if (threads > 64) threads = 64;
    for (int t = 0; t < threads; t++) pthread_create(&tid[t], NULL, worker, j);
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    pthread_mutex_destroy(&j->mu);
}

void przm_gen_markov(uint8_t *out, uint64_t n, uint64_t seed, int threads) {
    markov_t *m = (markov_t *)malloc(sizeof(markov_t));
    markov_init(m, seed);
    job_t j = {0, out, n, 1u << 20, seed, m, 0, PTHREAD_MUTEX_INITIALIZER};
    run(&j, threads);
    free(m);
}

void przm_gen_mixture(uint8_t *out, uint64_t n, uint64_t seed, int threads) {
    pthread_once(&g_vocab_once, vocab_init);
    job_t j = {1, out, n, 65536, seed, NULL, 0, PTHREAD_MUTEX_INITIALIZER};
    run(&j, threads);
}

/* Kind of each 64 KiB block of a mixture stream (0 JSON, 1 table, 2 motif, 3 random). */
void przm_gen_mixture_kinds(uint64_t seed, uint64_t nblocks, uint8_t *kinds) {
    for (uint64_t b = 0; b < nblocks; b++) kinds[b] = (uint8_t)mixture_kind(seed, b);
}

void przm_gen_loguniform_sizes(uint64_t *sizes, uint64_t count, uint64_t lo, uint64_t hi, uint64_t seed) {
    rng_t r;
    rng_seed(&r, seed, 0x400000000ull);
    double a = log((double)lo), b = log((double)hi);
    for (uint64_t i = 0; i < count; i++) {
        uint64_t s = (uint64_t)floor(exp(a + (b - a) * rng_unit(&r)));
        if (s < lo) s = lo;
        if (s > hi) s = hi;
        sizes[i] = s;
    }
}

/* Batch of buffers laid out back to back; buffer i is a mixture stream with seed (seed, i). */
typedef struct { uint8_t *out; const uint64_t *offs; const uint64_t *sizes; uint64_t count, seed, next; pthread_mutex_t mu; } bjob_t;
static void *bworker(void *arg) {
    bjob_t *j = (bjob_t *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        uint64_t i = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (i >= j->count) break;
        uint64_t s = j->seed ^ (0x9E3779B97F4A7C15ull * (i + 1));
        for (uint64_t off = 0; off < j->sizes[i]; off += 65536) {
            uint64_t len = j->sizes[i] - off < 65536 ? j->sizes[i] - off : 65536;
            mixture_block(j->out + j->offs[i] + off, len, s, off / 65536);
        }
    }
    return NULL;
}
void przm_gen_batch(uint8_t *out, const uint64_t *offs, const uint64_t *sizes, uint64_t count, uint64_t seed,
                    int threads) {
    pthread_once(&g_vocab_once, vocab_init);
    if (threads <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        threads = c > 16 ? 16 : (c < 1 ? 1 : (int)c);
    }
    if (threads > 64) threads = 64;
    bjob_t j = {out, offs, sizes, count, seed, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t tid[64];
    for (int t = 0; t < threads; t++) pthread_create(&tid[t], NULL, bworker, &j);
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}
