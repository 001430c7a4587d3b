/*
 * A host emulation of the HIP device dialect the entropy and sheaf kernels are written in, so
 * that their sources compile with g++ and run on the CPU with every GPU thread a fiber
 * (ucontext).  Test infrastructure only: it lets a CPU test run a kernel's workgroups under
 * different thread schedules between synchronisations and compare the results, so a missing
 * __syncthreads or wave barrier shows up as a difference without a GPU.
 *
 * What is emulated:
 *  - workgroups run one at a time; __shared__ variables are function statics (one copy, as
 *    LDS is one copy per workgroup);
 *  - __syncthreads: a barrier over the workgroup's live threads;
 *  - wave operations (DPP moves, readlane, ballot, all, shuffles, the wave barrier): a
 *    rendezvous of the wave's live lanes, in the order each lane reaches them.  They must be
 *    reached by every live lane of the wave (uniform control flow per wave), which holds for
 *    the kernels this is used on;
 *  - the scheduler runs the fibers by sweeps (each live fiber resumed once per sweep, in
 *    ascending, descending or seeded random order: the waves move about in step) or greedily
 *    (always the first fiber, in ascending or descending thread order, that can go on: a wave
 *    runs ahead until a barrier stops it); a fiber runs until it blocks at a synchronisation or
 *    exits, and a schedule that can make no progress is reported as a deadlock;
 *  - atomics are plain read-modify-writes (one OS thread runs every fiber).
 * DPP controls: quad_perm (0x00-0xFF), row_shr:1..15 (0x111-0x11F), wave_shr:1 (0x138),
 * row_bcast:15 (0x142), row_bcast:31 (0x143), with row masks; a lane whose source is invalid keeps `old` (bound_ctrl
 * false).  v_rcp_f32 is the correctly rounded reciprocal.
 */
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <ucontext.h>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __constant__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
/* one address space on the host: __attribute__((address_space(1))) becomes an empty attribute */
#define address_space(n)

struct uint4 {
    uint32_t x, y, z, w;
};
static inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }

namespace simt {

struct Dim {
    uint32_t x, y, z;
};

struct Fiber {
    ucontext_t ctx;
    std::vector<char> stack;
    uint32_t tid = 0;
    bool done = false;
    int wait = 0;      /* 0 running, 1 at a barrier, 2 at a wave operation */
    uint64_t wait_gen = 0;
};

struct Wave {
    uint32_t live = 0, arrived = 0;
    uint64_t gen = 0;
    uint64_t vals[2][64];
    uint64_t mask[2] = {0, 0};
};

struct Group {
    Dim block_idx{0, 0, 0}, grid_dim{1, 1, 1}, block_dim{1, 1, 1};
    std::vector<Fiber> fibers;
    std::vector<Wave> waves;
    uint32_t live = 0, bar_arrived = 0;
    uint64_t bar_gen = 0, events = 0;
    int cur = -1;
    ucontext_t sched;
    std::function<void()> body;
};

inline Group *g_grp = nullptr;

inline Fiber &me() { return g_grp->fibers[(size_t)g_grp->cur]; }
inline void yield() { swapcontext(&me().ctx, &g_grp->sched); }

inline Dim tidx() { return Dim{me().tid, 0, 0}; }
inline Dim bidx() { return g_grp->block_idx; }
inline Dim gdim() { return g_grp->grid_dim; }
inline Dim bdim() { return g_grp->block_dim; }

inline void barrier() {
    Group &G = *g_grp;
    const uint64_t gen = G.bar_gen;
    G.events++;
    if (++G.bar_arrived == G.live) {
        G.bar_arrived = 0;
        G.bar_gen++;
        return;
    }
    me().wait = 1;
    me().wait_gen = gen;
    while (G.bar_gen == gen) yield();
    me().wait = 0;
}

/* the wave's values of v, after every live lane of the wave has given its own; *mask: the lanes
 * that took part */
inline const uint64_t *wave_xchg(uint64_t v, uint64_t *mask = nullptr) {
    Group &G = *g_grp;
    const uint32_t t = me().tid;
    Wave &W = G.waves[t / 64];
    const uint64_t gen = W.gen;
    const int b = (int)(gen & 1u);
    if (W.arrived == 0) W.mask[b] = 0;
    W.vals[b][t & 63] = v;
    W.mask[b] |= 1ull << (t & 63);
    G.events++;
    if (++W.arrived == W.live) {
        W.arrived = 0;
        W.gen++;
    } else {
        me().wait = 2;
        me().wait_gen = gen;
        while (W.gen == gen) yield();
        me().wait = 0;
    }
    if (mask) *mask = W.mask[b];
    return W.vals[b];
}

/* fiber F can go on */
inline bool ready(const Group &G, const Fiber &F) {
    if (F.done) return false;
    if (F.wait == 1) return G.bar_gen != F.wait_gen;
    if (F.wait == 2) return G.waves[F.tid / 64].gen != F.wait_gen;
    return true;
}

/* run body on a grid of nblk workgroups of nthr threads.  order: 0 / 1 sweeps in ascending /
 * descending thread order, 2 / 3 greedy in ascending / descending thread order, 4 and above
 * sweeps in a random order per sweep (seeded by order) */
inline int launch(uint32_t nblk, uint32_t nthr, int order, const std::function<void()> &body) {
    std::mt19937 rng((uint32_t)(order > 3 ? order : 0));
    for (uint32_t b = 0; b < nblk; b++) {
        Group G;
        G.block_idx = Dim{b, 0, 0};
        G.grid_dim = Dim{nblk, 1, 1};
        G.block_dim = Dim{nthr, 1, 1};
        G.fibers.resize(nthr);
        G.waves.resize((nthr + 63) / 64);
        G.live = nthr;
        G.body = body;
        for (uint32_t t = 0; t < nthr; t++) G.waves[t / 64].live++;
        g_grp = &G;
        for (uint32_t t = 0; t < nthr; t++) {
            Fiber &F = G.fibers[t];
            F.tid = t;
            F.stack.resize(nthr > 512 ? 96 * 1024 : 256 * 1024);
            getcontext(&F.ctx);
            F.ctx.uc_stack.ss_sp = F.stack.data();
            F.ctx.uc_stack.ss_size = F.stack.size();
            F.ctx.uc_link = &G.sched;
            makecontext(&F.ctx, (void (*)())+[]() {
                g_grp->body();
                Fiber &Me = me();
                Me.done = true;
                Group &Gr = *g_grp;
                Gr.live--;
                Gr.events++;
                Wave &W = Gr.waves[Me.tid / 64];
                W.live--;
                /* an exit can complete a barrier or a wave operation the others wait at */
                if (Gr.live && Gr.bar_arrived == Gr.live) { Gr.bar_arrived = 0; Gr.bar_gen++; }
                if (W.live && W.arrived == W.live) { W.arrived = 0; W.gen++; }
            }, 0);
        }
        std::vector<uint32_t> ord(nthr);
        for (uint32_t t = 0; t < nthr; t++) ord[t] = (order == 1 || order == 3) ? nthr - 1 - t : t;
        while (G.live && (order == 2 || order == 3)) {
            uint32_t k = 0;
            while (k < nthr && !ready(G, G.fibers[ord[k]])) k++;
            if (k == nthr) {
                fprintf(stderr, "simt: deadlock in workgroup %u (%u threads live)\n", b, G.live);
                g_grp = nullptr;
                return -1;
            }
            G.cur = (int)ord[k];
            swapcontext(&G.sched, &G.fibers[ord[k]].ctx);
        }
        while (G.live) {
            if (order > 3) std::shuffle(ord.begin(), ord.end(), rng);
            const uint64_t ev = G.events;
            for (uint32_t k = 0; k < nthr; k++) {
                Fiber &F = G.fibers[ord[k]];
                if (F.done) continue;
                G.cur = (int)ord[k];
                swapcontext(&G.sched, &F.ctx);
            }
            if (G.live && G.events == ev) {
                fprintf(stderr, "simt: deadlock in workgroup %u (%u threads live)\n", b, G.live);
                g_grp = nullptr;
                return -1;
            }
        }
        g_grp = nullptr;
    }
    return 0;
}

} // namespace simt

#define threadIdx (simt::tidx())
#define blockIdx (simt::bidx())
#define gridDim (simt::gdim())
#define blockDim (simt::bdim())

static inline void __syncthreads() { simt::barrier(); }
static inline void __threadfence_block() {}
static inline uint32_t __lane_id() { return simt::me().tid & 63u; }

static inline void __builtin_amdgcn_wave_barrier() { (void)simt::wave_xchg(0); }

static inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
    const uint32_t l = simt::me().tid & 63u;
    const uint64_t *v = simt::wave_xchg((uint32_t)src);
    const uint32_t row = l >> 4;
    if (!((row_mask >> row) & 1) || !((bank_mask >> ((l & 15u) >> 2)) & 1)) return old;
    int s = -1;
    if (ctrl <= 0xFF) s = (int)((l & ~3u) | ((uint32_t)(ctrl >> (2 * (l & 3u))) & 3u));
    else if (ctrl >= 0x111 && ctrl <= 0x11F) s = (l & 15u) >= (uint32_t)(ctrl & 15) ? (int)l - (ctrl & 15) : -1;
    else if (ctrl == 0x138) s = l > 0 ? (int)l - 1 : -1; /* wave_shr:1 */
    else if (ctrl == 0x142) s = row > 0 ? (int)(16 * row - 1) : -1;
    else if (ctrl == 0x143) s = row >= 2 ? 31 : -1;
    else {
        fprintf(stderr, "simt: DPP control 0x%x not emulated\n", ctrl);
        abort();
    }
    if (s < 0) return bound_ctrl ? 0 : old;
    return (int)(uint32_t)v[s];
}
static inline int __builtin_amdgcn_readlane(int x, int lane) { return (int)(uint32_t)simt::wave_xchg((uint32_t)x)[lane & 63]; }
static inline uint64_t __ballot(int pred) {
    uint64_t m;
    const uint64_t *v = simt::wave_xchg(pred ? 1u : 0u, &m);
    uint64_t r = 0;
    for (int l = 0; l < 64; l++)
        if (((m >> l) & 1) && v[l]) r |= 1ull << l;
    return r;
}
static inline int __all(int pred) {
    uint64_t m;
    const uint64_t *v = simt::wave_xchg(pred ? 1u : 0u, &m);
    for (int l = 0; l < 64; l++)
        if (((m >> l) & 1) && !v[l]) return 0;
    return 1;
}
static inline int __any(int pred) {
    uint64_t m;
    const uint64_t *v = simt::wave_xchg(pred ? 1u : 0u, &m);
    for (int l = 0; l < 64; l++)
        if (((m >> l) & 1) && v[l]) return 1;
    return 0;
}
template <typename T> static inline T __shfl_xor(T x, int mask, int width = 64) {
    const uint32_t l = simt::me().tid & 63u;
    const uint64_t *v = simt::wave_xchg((uint64_t)(uint32_t)x);
    const uint32_t s = ((l ^ (uint32_t)mask) & (uint32_t)(width - 1)) | (l & ~(uint32_t)(width - 1));
    return (T)(uint32_t)v[s];
}
template <typename T> static inline T __shfl_up(T x, unsigned d, int width = 64) {
    const uint32_t l = simt::me().tid & 63u;
    const uint64_t *v = simt::wave_xchg((uint64_t)(uint32_t)x);
    const uint32_t base = l & ~(uint32_t)(width - 1);
    return (l - base) >= d ? (T)(uint32_t)v[l - d] : x;
}

template <typename T> static inline T __shfl(T x, int src, int width = 64) {
    const uint32_t l = simt::me().tid & 63u;
    const uint64_t *v = simt::wave_xchg((uint64_t)(uint32_t)x);
    const uint32_t base = l & ~(uint32_t)(width - 1);
    return (T)(uint32_t)v[base + ((uint32_t)src & (uint32_t)(width - 1))];
}
static inline void __builtin_amdgcn_sched_barrier(int) {}
static inline uint64_t __builtin_amdgcn_s_memrealtime() {
    static uint64_t t = 0;
    return ++t;
}
static inline uint32_t __builtin_amdgcn_alignbit(uint32_t hi, uint32_t lo, uint32_t s) {
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (s & 31u));
}
static inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
static inline uint32_t atomicMin(uint32_t *p, uint32_t v) { const uint32_t o = *p; *p = o < v ? o : v; return o; }
static inline uint32_t __builtin_amdgcn_sad_u8(uint32_t a, uint32_t b, uint32_t c) {
    for (int k = 0; k < 4; k++) {
        const int x = (int)((a >> (8 * k)) & 0xFF), y = (int)((b >> (8 * k)) & 0xFF);
        c += (uint32_t)(x > y ? x - y : y - x);
    }
    return c;
}
static inline uint32_t __builtin_amdgcn_alignbyte(uint32_t hi, uint32_t lo, uint32_t s) {
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (8 * (s & 3u)));
}
static inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
static inline uint32_t __umul24(uint32_t a, uint32_t b) { return (a & 0xFFFFFFu) * (b & 0xFFFFFFu); }
static inline int __popc(uint32_t x) { return __builtin_popcount(x); }
static inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }

static inline uint32_t atomicAdd(uint32_t *p, uint32_t v) { const uint32_t o = *p; *p = o + v; return o; }
static inline int atomicAdd(int *p, int v) { const int o = *p; *p = o + v; return o; }
static inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v) {
    const unsigned long long o = *p; *p = o + v; return o;
}
static inline uint32_t atomicOr(uint32_t *p, uint32_t v) { const uint32_t o = *p; *p = o | v; return o; }
static inline unsigned long long atomicOr(unsigned long long *p, unsigned long long v) {
    const unsigned long long o = *p; *p = o | v; return o;
}
