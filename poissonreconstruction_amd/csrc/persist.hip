// The CG of every instance of a batched level in one persistent launch.
//
// The per-pass batched solve (batch.hip) enqueues one launch per phase of a
// CG iteration, replays captured chunks of iterations and returns to the host
// whenever every running instance has stopped, to replace or confirm residuals
// and restart instances.  Here one cooperative launch does all of it: its
// workgroups (all co-resident) run the iterations of every instance in
// lockstep rounds and make the per-instance decisions on the device.
//
// Round k (global pass index k: ping-pong slot k & 1, ring slot k % CG_XM):
//   replacement phases, when some instance stopped in round k - 1:
//     mixed: pending x terms of the FP32 correction e added; the FP64
//     screening values of x + e per point; per node x2 = x + e and the FP64
//     true residual b - A x2, which replaces the FP32 residual in slot k & 1;
//     its norm and the new FP32 |r|^2 per instance; then the decision of the
//     single-tree mixed solve: converged, failed, or continue along the
//     current direction with the replaced residual (a reliable update).
//     FP64: pending x terms added, r = b - A x into slot k & 1, its norm and
//     the decision of the single-tree FP64 solve (converged, stalled, capped,
//     or continue).
//   iteration phases, for the running instances:
//     gather: x terms (instances with CG_XM pending) and the per-point
//     screening values of the pass's operand;
//     pass: the fused CG step over the instances' 16^3 bricks (the tile pass
//     of the per-pass solve, one brick at a time);
//     fold: per instance its four sums in a fixed order and the scalar step
//     fin_cg; the last workgroup to finish builds the next round's work lists.
// A grid barrier separates the phases.  Its wait is bounded: a wait that
// exceeds about four seconds sets an error flag, and every workgroup leaves
// at its next barrier, so the launch always ends.
#include <cstdlib>

#include "tile.hpp"
#include "cg_dev.hpp"

namespace prx {

// Control words are read with relaxed agent-scope atomics: the barrier's
// acquire fence has already made the other workgroups' writes visible, and an
// acquire per read would invalidate the XCD's L2 under the workgroups still
// working.
__device__ __forceinline__ int32_t ld_ctl(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_ctl(int32_t* p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier over all workgroups of the launch (gen: barriers passed, the
// same in every thread).  On arrival a workgroup releases its writes (an L2
// write-back, which leaves the lines of the workgroups still working in
// place) and counts itself in its group's counter; the last of a group counts
// the group at the top counter, and the last group opens the barrier.  The
// waiting poll is relaxed and throttled; one acquire fence follows it.  The
// two levels keep the serialised device-scope atomics per barrier at about
// G / PBAR_GROUPS + PBAR_GROUPS instead of G.  Returns false when the launch
// is to end (a wait exceeded its bound here or elsewhere).
constexpr uint64_t PERSIST_SPIN_LIMIT = 1ull << 22;
// Each wave waits for its own global stores to be acknowledged (s_waitcnt
// vmcnt(0)) before a workgroup barrier after which one thread releases the
// workgroup's writes at agent scope: the workgroup barrier does not wait for
// the stores of the other waves, and a store still in flight when the L2 is
// written back could reach another XCD's reader stale.
// (0x0F70 on gfx9-family targets: vmcnt = 0, expcnt = 7 and lgkmcnt = 15, i.e.
// no wait on those two counters.)
__device__ __forceinline__ void own_stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }
__device__ inline bool gsync(PCtl* c, uint32_t& gen, int* sbad) {
    own_stores_done();
    __syncthreads();
    const uint32_t target = gen + 1;
    gen = target;
    if (wg_tid() == 0) {
        int bad = 0;
        const uint32_t G = gridDim.x, w = blockIdx.x;
        const uint32_t ng = G < (uint32_t)PBAR_GROUPS ? G : (uint32_t)PBAR_GROUPS, g = w % ng;
        const uint32_t members = (G - g + ng - 1) / ng;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t* gc = &c->group[g][0];
        const uint32_t a = __hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a == members - 1) {
            // the group's writes happen before this one's next release
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // each counter's reset is performed before the barrier can open, so that no
            // workgroup's count for the next barrier lands before it and is lost
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const uint32_t t = __hip_atomic_fetch_add(&c->bar_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t == ng - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                __hip_atomic_store(&c->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(&c->bar_gen, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        uint64_t spins = 0;
        while (__hip_atomic_load(&c->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != target) {
            if (ld_ctl(&c->err)) { bad = 1; break; }
            __builtin_amdgcn_s_sleep(8);
            if (++spins > PERSIST_SPIN_LIMIT) {
                st_ctl(&c->err, 1);
                bad = 1;
                break;
            }
        }
        if (ld_ctl(&c->err)) bad = 1;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *sbad = bad;
    }
    __syncthreads();
    const bool ok = *sbad == 0;
    __syncthreads();
    return ok;
}

// the last list entry j < n with pref[j] <= v (entries may be empty)
__device__ inline int list_find(const uint32_t* pref, int n, uint32_t v) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// advance j to the entry holding v (v never decreases along a loop)
__device__ __forceinline__ int list_advance(const uint32_t* pref, int n, uint32_t v, int j) {
    while (j + 1 < n && pref[j + 1] <= v) ++j;
    return j;
}

// Exclusive block scan of v (256 threads); *total gets the sum.
__device__ inline uint32_t block_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const int lane = wg_tid() & 63, wave = wg_tid() >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < 4; ++w) {
        if (w < wave) before += tmp[w];
        all += tmp[w];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// The next round's work lists, by one workgroup: running instances, those to
// replace (confirm), those whose CG_XM x terms are due before their next pass.
template <typename T>
__device__ inline void build_lists(const PArgs<T>& a, uint32_t* tmp) {
    constexpr int VEC = 16 / (int)sizeof(T);
    const int K = a.tab.K;
    __shared__ uint32_t acc[3][5];
    if (wg_tid() < 15) (&acc[0][0])[wg_tid()] = 0;
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += 256) {
        const int k = k0 + wg_tid();
        uint32_t mem[3] = {0, 0, 0}, nb = 0, np = 0, nv = 0, n1 = 0;
        if (k < K) {
            const CgState& st = a.tab.st[k];
            const int fin = a.pi[k].fin;
            const int rep = a.flag[k];
            const int run = !fin && !rep && st.done == 0;
            mem[0] = run;
            mem[1] = rep ? 1u : 0u;
            mem[2] = (run && st.xn == CG_XM) ? 1u : 0u;
            const uint64_t nk = a.tab.off[k + 1] - a.tab.off[k];
            nb = (uint32_t)(nk >> 12);
            np = a.pt_off[k + 1] - a.pt_off[k];
            nv = (uint32_t)(nk / (256 * VEC));
            n1 = (uint32_t)(nk >> 10);
        }
        for (int l = 0; l < 3; ++l) {
            const PList pl = l == 0 ? a.run : (l == 1 ? a.repl : a.xt);
            uint32_t t0, t1, t2, t3, t4;
            const uint32_t e0 = block_scan(mem[l], tmp, &t0);
            const uint32_t e1 = block_scan(mem[l] * nb, tmp, &t1);
            const uint32_t e2 = block_scan(mem[l] * np, tmp, &t2);
            const uint32_t e3 = block_scan(mem[l] * nv, tmp, &t3);
            const uint32_t e4 = block_scan(mem[l] * n1, tmp, &t4);
            if (mem[l]) {
                const uint32_t j = acc[l][0] + e0;
                pl.inst[j] = k;
                pl.gbrick[j] = (uint32_t)(a.tab.off[k] >> 12);
                pl.brick[j] = acc[l][1] + e1;
                pl.pts[j] = acc[l][2] + e2;
                pl.vblk[j] = acc[l][3] + e3;
                pl.kblk[j] = acc[l][4] + e4;
            }
            __syncthreads();
            if (wg_tid() == 0) {
                acc[l][0] += t0;
                acc[l][1] += t1;
                acc[l][2] += t2;
                acc[l][3] += t3;
                acc[l][4] += t4;
            }
            __syncthreads();
        }
    }
    if (wg_tid() < 3) {
        const int l = wg_tid();
        const PList pl = l == 0 ? a.run : (l == 1 ? a.repl : a.xt);
        const uint32_t n = acc[l][0];
        pl.brick[n] = acc[l][1];
        pl.pts[n] = acc[l][2];
        pl.vblk[n] = acc[l][3];
        pl.kblk[n] = acc[l][4];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (l == 0) st_ctl(&a.ctl->nrun, (int32_t)n);
        else if (l == 1) st_ctl(&a.ctl->nrepl, (int32_t)n);
        else st_ctl(&a.ctl->nxt, (int32_t)n);
    }
    __syncthreads();
}

// A block of work items finished; the last of `total` builds the lists.
template <typename T>
__device__ inline void finish_item(const PArgs<T>& a, int total, int* slast, uint32_t* tmp) {
    own_stores_done();
    __syncthreads();
    if (wg_tid() == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t t = __hip_atomic_fetch_add(&a.ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *slast = (int)t == total - 1;
        if (*slast) {
            __hip_atomic_store(&a.ctl->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
    if (*slast) build_lists<T>(a, tmp);
}

// Fixed-order sum over n values at p[stride * i] by the block (256 threads).
__device__ inline double block_fold(const double* p, int64_t n, int stride, double* ws) {
#pragma clang fp contract(off)
    double s = 0.0;
    for (int64_t i = wg_tid(); i < n; i += 256) s += p[stride * i];
    return block_sum_256(s, ws);
}

// The run flags of the batched bodies: instances whose CG runs (done == 0),
// or those flagged for a replacement / confirmation.
template <typename T>
__device__ inline BatchTab tab_running(const PArgs<T>& a) {
    BatchTab t = a.tab;
    t.run = &a.tab.st[0].done;
    t.run_stride = (int)(sizeof(CgState) / sizeof(int32_t));
    t.run_when = 0;
    return t;
}
template <typename T>
__device__ inline BatchTab tab_flagged(const PArgs<T>& a) {
    BatchTab t = a.tab;
    t.run = a.flag;
    t.run_stride = 1;
    t.run_when = 1;
    return t;
}

// The XCD a workgroup runs on: the XCC_ID hardware register (bits 3:0 of
// hwreg 20 on gfx940 and later).
__device__ __forceinline__ uint32_t xcc_id() {
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u;
}
// The workgroup's XCD for the brick phases (set once per launch): one of nx
// groups (8, one per XCD; 1 when some XCD has no workgroup).
struct XcdSlot {
    uint32_t x, nx;
};

// The 16^3 bricks of a list, XCD-aware: chunks of 64 consecutive positions go
// to group chunk % nx, and the workgroups of that group take its positions one
// at a time from the group's counter (*take), in the order they finish their
// previous brick, as the hardware hands out the workgroups of a launch.  The
// bricks in progress thus stay a narrow window of consecutive positions under
// one L2, whose halos the neighbouring bricks read again; a fixed share per
// workgroup lets the workgroups drift apart over hundreds of bricks.  Thread 0
// takes the next position as the current brick starts (its wave waits for the
// atomic; the other waves start on the brick), so the position is in LDS at
// the next brick's barrier.  f(brick) is called by the whole workgroup with a
// uniform brick id.  Every brick processed is counted in *done.  The counters
// start at zero.
template <typename F>
__device__ inline void for_bricks(const PList& L, int n, const XcdSlot& xs, uint32_t* take, uint32_t* done, F f) {
    if (n <= 0) return;
    const uint32_t total = L.brick[n];
    constexpr uint32_t CH = 64;
    __shared__ uint32_t snext[2];
    uint32_t* const ctr = take + 32u * xs.x;
    if (wg_tid() == 0) snext[0] = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int j = 0;
    for (uint32_t i = 0;; ++i) {
        __syncthreads();  // snext[i & 1] is written; the previous brick's LDS reads are over
        const uint32_t p = snext[i & 1];
        const uint64_t chunk = xs.x + (uint64_t)xs.nx * (p / CH);
        if (chunk * CH >= total) break;
        if (wg_tid() == 0) snext[(i + 1) & 1] = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t g = chunk * CH + p % CH;
        if (g < total) {
            j = list_advance(L.brick, n, (uint32_t)g, j);
            const uint32_t bid = L.gbrick[j] + ((uint32_t)g - L.brick[j]);
            f(bid);
            if (wg_tid() == 0) __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}
// After a brick phase of round k (by one thread): every brick of the list
// processed exactly once, else err 4; the count is cleared for round k + 2.
__device__ inline void check_done(PCtl* c, int ph, int64_t k, uint32_t want) {
    uint32_t* d = &c->done[ph][k & 1][0];
    const uint32_t got = __hip_atomic_load(d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (got != want) {
        c->bad_round = k * 2 + ph;
        c->bad_got = got;
        c->bad_want = want;
        st_ctl(&c->err, 4);
    }
}

// The decision after a replacement (mixed): as the host loop of the per-pass
// batched mixed solve.  rt = |b - A x|^2, dt = the new FP32 |r|^2.
__device__ inline void decide_mixed(const PArgs<float>& a, int k, double rt, double dt) {
#pragma clang fp contract(off)
    CgState& h = a.tab.st[k];
    PInst& p = a.pi[k];
    p.rr = rt;
    p.outer += 1;
    h.xn = 0;
    h.xpend = 0;
    if (rt <= p.thr) { p.fin = 1; h.done = 1; p.status = PRX_OK; return; }
    if (h.status == PRX_ERR_BREAKDOWN) { p.fin = 1; h.done = 1; p.status = PRX_ERR_BREAKDOWN; return; }
    if (h.status == PRX_ERR_NOT_CONVERGED || h.it >= a.max_iters || p.outer >= a.max_outer) {
        p.fin = 1;
        h.done = 1;
        p.status = PRX_ERR_NOT_CONVERGED;
        return;
    }
    const bool drifted = !(dt <= 4.0 * h.delta);
    h.beta = (!drifted && h.delta_prev > 0.0) ? dt / h.delta_prev : 0.0;
    h.delta = dt;
    const double t = a.theta * a.theta * dt;
    h.thr = t > p.thr32 ? t : p.thr32;
    h.done = 0;
    h.status = 0;
    h.rpend = 0;
}
// The decision after a confirmation (FP64): as the host loop of the per-pass
// batched FP64 solve.  rt = |b - A x|^2.
__device__ inline void decide_fp64(const PArgs<double>& a, int k, double rt) {
#pragma clang fp contract(off)
    CgState& h = a.tab.st[k];
    PInst& p = a.pi[k];
    h.xn = 0;
    h.xpend = 0;
    p.rr = rt;
    const bool stalled = rt > 0.25 * p.last_rt;
    if (rt <= h.thr || h.it >= h.max_iters || stalled) {
        h.delta = rt;
        if (rt > h.thr) h.status = PRX_ERR_NOT_CONVERGED;
        p.fin = 1;
        p.status = h.status;
        return;
    }
    p.last_rt = rt;
    const bool drifted = !(rt <= 4.0 * h.delta);
    h.beta = (!drifted && h.delta_prev > 0.0) ? rt / h.delta_prev : 0.0;
    h.delta = rt;
    h.done = 0;
    h.status = 0;
    h.rpend = 0;
}

// block 0 charges the time since the last barrier to phase kind i
__device__ __forceinline__ void tick(PCtl* c, int i, uint64_t& last) {
    if (blockIdx.x == 0 && wg_tid() == 0) {
        const uint64_t now = wall_clock64();
        c->ticks[i] += now - last;
        last = now;
    }
}

template <typename T>
__device__ __forceinline__ void persist_main(const PArgs<T>* __restrict__ pa) {
    const PArgs<T>& a = *pa;
    constexpr bool MIXED = sizeof(T) == 4;
    constexpr int VEC = 16 / (int)sizeof(T);
    // the tile of the CG pass, shared with the residual passes (FP64 in the mixed solve)
    constexpr size_t LDS_CG = sizeof(TileLds<T, 16, OP_CG>);
    constexpr size_t LDS_RES = MIXED ? sizeof(TileLds<double, 16, OP_RESID_UPD>) : sizeof(TileLds<T, 16, OP_RESID>);
    __shared__ __align__(16) unsigned char lds_raw[LDS_CG > LDS_RES ? LDS_CG : LDS_RES];
    TileLds<T, 16, OP_CG>& lds = *reinterpret_cast<TileLds<T, 16, OP_CG>*>(lds_raw);
    __shared__ double ws[16];
    __shared__ uint32_t tmp[8];
    __shared__ int sflag[2];
    PCtl* c = a.ctl;
    const int64_t N = a.N;
    const uint32_t G = gridDim.x, w = blockIdx.x;
    uint32_t gen = 0;
    const uint64_t t0 = wall_clock64();
    uint64_t tl = t0;
    // the workgroup's XCD; every XCD takes one eighth of the chunks
    __shared__ uint32_t sx;
    if (wg_tid() == 0) {
        uint32_t x = xcc_id();
        if (x >= 8) x = w & 7u;
        sx = x;
        __hip_atomic_fetch_add(&c->xcd_count[x], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (x == (w & 7u)) __hip_atomic_fetch_add(&c->xcd_rr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (w == 0) build_lists<T>(a, tmp);
    if (!gsync(c, gen, &sflag[0])) return;
    tick(c, 3, tl);
    XcdSlot xs{sx, 8u};
    bool all8 = true;  // an XCD without workgroups would leave its chunks undone
    for (int i = 0; i < 8; ++i) all8 = all8 && __hip_atomic_load(&c->xcd_count[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0;
    if (!all8) {
        if (G >= 8 && !(G & 7u)) xs.x = w & 7u;  // by launch order, eight equal groups
        else xs = XcdSlot{0u, 1u};
    }
    const BatchTab trun = tab_running(a), trep = tab_flagged(a);
    for (int64_t k = 0;; ++k) {
        if (ld_ctl(&c->err)) return;
        int nrun = ld_ctl(&c->nrun);
        const int nrepl = ld_ctl(&c->nrepl);
        if (nrun == 0 && nrepl == 0) break;
        if (k >= a.max_rounds) {
            if (w == 0 && wg_tid() == 0) st_ctl(&c->err, 2);
            break;
        }
        T* const rk = a.r + (k & 1) * N;
        // the brick counters of round k + 1 (last used in round k - 1)
        if (w == 0 && wg_tid() < 16)
            __hip_atomic_store(&c->take[wg_tid() >> 3][(k + 1) & 1][wg_tid() & 7][0], 0u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (nrepl > 0) {
            if (w == 0 && wg_tid() == 0) c->repl_rounds++;
            // -- pending x terms of the instances to replace
            {
                const uint32_t tv = a.repl.vblk[nrepl];
                for (uint32_t v = w, j = w < tv ? (uint32_t)list_find(a.repl.vblk, nrepl, w) : 0u; v < tv; v += G) {
                    j = (uint32_t)list_advance(a.repl.vblk, nrepl, v, (int)j);
                    const int kk = a.repl.inst[j];
                    const uint32_t blk = (uint32_t)(a.tab.off[kk] / (256 * VEC)) + (v - a.repl.vblk[j]);
                    bx_terms_block<T, 1>(trep, a.x, XDirs<T>{}, a.ring, N, blk);
                }
            }
            if (!gsync(c, gen, &sflag[0])) return;
            tick(c, 0, tl);
            // -- per-point screening values of the solution (mixed: x + e)
            if (a.has_scr) {
                const uint32_t tp = a.repl.pts[nrepl], v0 = w * 256u + (uint32_t)wg_tid();
                int j = v0 < tp ? list_find(a.repl.pts, nrepl, v0) : 0;
                for (uint32_t v = v0; v < tp; v += G * 256u) {
                    j = list_advance(a.repl.pts, nrepl, v, j);
                    const int kk = a.repl.inst[j];
                    const int64_t ip = (int64_t)a.pt_off[kk] + (v - a.repl.pts[j]);
                    if constexpr (MIXED)
                        bscreen_gather_point<double, 1>(trep, a.P, a.cell, a.t64, a.pinst, nullptr, a.x64, (const float*)a.x, a.sbuf64,
                                                        nullptr, nullptr, nullptr, ip);
                    else
                        bscreen_gather_point<double, 0>(trep, a.P, a.cell, a.t64, a.pinst, (const double*)a.x, nullptr, nullptr,
                                                        a.sbuf64, nullptr, nullptr, nullptr, ip);
                }
            }
            if (!gsync(c, gen, &sflag[0])) return;
            tick(c, 0, tl);
            // -- the true residual into slot k & 1 (partial per brick)
            if constexpr (MIXED) {
                // x2 = x + e, r = (b - A x2) / h into slot k & 1, |b - A x2|^2 per brick: the per-pass
                // solve's FP64 residual update pass
                StencilIO<double> io;
                io.x64 = a.x64;
                io.e32 = (const float*)a.x;
                io.x_out = a.x2;
                io.b = a.b;
                io.r32 = (float*)rk;
                io.s = a.sbuf64;
                io.partial = a.partial;
                for_bricks(a.repl, nrepl, xs, &c->take[0][k & 1][0][0], &c->done[0][k & 1][0], [&](uint32_t bid) {
                    stencil_tile_brick<double, 16, OP_RESID_UPD, true, true>(
                        a.L64, io, 0, trep, bid, *reinterpret_cast<TileLds<double, 16, OP_RESID_UPD>*>(lds_raw));
                });
            } else {
                StencilIO<T> io;
                io.x = a.x;
                io.b = (const T*)a.b;
                io.y = rk;
                io.s = (const T*)a.sbuf64;
                io.partial = a.partial;
                for_bricks(a.repl, nrepl, xs, &c->take[0][k & 1][0][0], &c->done[0][k & 1][0], [&](uint32_t bid) {
                    stencil_tile_brick<T, 16, OP_RESID, true, true>(a.L, io, 0, trep, bid,
                                                                    *reinterpret_cast<TileLds<T, 16, OP_RESID>*>(lds_raw));
                });
            }
            if (!gsync(c, gen, &sflag[0])) return;
            tick(c, 0, tl);
            // (the replacement list is rebuilt after workgroup 0's first decision, which follows)
            if (w == 0 && wg_tid() == 0) check_done(c, 0, k, a.repl.brick[nrepl]);
            if constexpr (MIXED) {
                // -- per instance |b - A x2|^2; x = x2, e = 0; the new FP32 |r|^2 per 1024-node block
                for (int j = (int)w; j < nrepl; j += (int)G) {
                    const int kk = a.repl.inst[j];
                    const int64_t p0 = (int64_t)(a.tab.off[kk] >> 12), np = (int64_t)(a.tab.off[kk + 1] >> 12) - p0;
                    const double t = block_fold(a.partial + p0, np, 1, ws);
                    if (wg_tid() == 0) a.tab.st[kk].slot[1] = t;
                    __syncthreads();
                }
                const uint32_t tk = a.repl.kblk[nrepl];
                for (uint32_t v = w, j = w < tk ? (uint32_t)list_find(a.repl.kblk, nrepl, w) : 0u; v < tk; v += G) {
                    j = (uint32_t)list_advance(a.repl.kblk, nrepl, v, (int)j);
                    const int kk = a.repl.inst[j];
                    const uint32_t blk = (uint32_t)(a.tab.off[kk] >> 10) + (v - a.repl.kblk[j]);
                    bcommit_block(trep, a.x64, a.x2, (float*)a.x, blk);
                    bdot_block<float>(trep, (const float*)rk, (const float*)rk, a.partial2, blk, ws);
                    __syncthreads();
                }
                if (!gsync(c, gen, &sflag[0])) return;
            tick(c, 0, tl);
                // -- decisions
                for (int j = (int)w; j < nrepl; j += (int)G) {
                    const int kk = a.repl.inst[j];
                    const int64_t p0 = (int64_t)(a.tab.off[kk] >> 10), np = (int64_t)(a.tab.off[kk + 1] >> 10) - p0;
                    const double dt = block_fold(a.partial2 + p0, np, 1, ws);
                    if (wg_tid() == 0) {
                        a.tab.st[kk].slot[2] = dt;
                        decide_mixed(a, kk, a.tab.st[kk].slot[1], dt);
                        a.flag[kk] = 0;
                    }
                    finish_item<T>(a, nrepl, &sflag[1], tmp);
                }
            } else {
                for (int j = (int)w; j < nrepl; j += (int)G) {
                    const int kk = a.repl.inst[j];
                    const int64_t p0 = (int64_t)(a.tab.off[kk] >> 12), np = (int64_t)(a.tab.off[kk + 1] >> 12) - p0;
                    const double rt = block_fold(a.partial + p0, np, 1, ws);
                    if (wg_tid() == 0) {
                        a.tab.st[kk].slot[1] = rt;
                        decide_fp64(a, kk, rt);
                        a.flag[kk] = 0;
                    }
                    finish_item<T>(a, nrepl, &sflag[1], tmp);
                }
            }
            if (!gsync(c, gen, &sflag[0])) return;
            tick(c, 0, tl);
            nrun = ld_ctl(&c->nrun);
        }
        if (nrun > 0) {
            T* const qk = a.q + (k & 1) * N;
            T* const dold = a.ring + (k % CG_XM) * N;
            // -- gather: x terms due, and the operand's screening values per point
            {
                const int nxt = ld_ctl(&c->nxt);
                XDirs<T> dirs;
                for (int i = 0; i < CG_XM; ++i) dirs.d[i] = a.ring + ((k + 1 + i) % CG_XM) * N;
                // XB consecutive blocks per workgroup and step, the steps interleaved over the workgroups
                constexpr int XB = sizeof(T) == 8 ? 4 : 2;
                const uint32_t tv = nxt > 0 ? a.xt.vblk[nxt] : 0u;
                for (uint32_t v0 = w * XB, j = v0 < tv ? (uint32_t)list_find(a.xt.vblk, nxt, v0) : 0u; v0 < tv; v0 += G * XB) {
                    uint32_t blk[XB], ins[XB];
                    int nb = 0;
#pragma unroll
                    for (int b = 0; b < XB; ++b) {
                        const uint32_t v = v0 + b;
                        blk[b] = ins[b] = 0;
                        if (v < tv) {
                            j = (uint32_t)list_advance(a.xt.vblk, nxt, v, (int)j);
                            ins[b] = (uint32_t)a.xt.inst[j];
                            blk[b] = (uint32_t)(a.tab.off[ins[b]] / (256 * VEC)) + (v - a.xt.vblk[j]);
                            nb = b + 1;
                        }
                    }
                    bx_terms_due<T, XB>(a.tab.st, a.x, dirs, blk, ins, nb);
                }
                if (a.has_scr) {
                    // SP points per thread and step (the steps interleaved over the workgroups)
                    constexpr int SP = 2;
                    const uint32_t tp = a.run.pts[nrun], v0 = w * (256u * SP) + (uint32_t)wg_tid();
                    int j = v0 < tp ? list_find(a.run.pts, nrun, v0) : 0;
                    for (uint32_t v = v0; v < tp; v += G * 256u * SP) {
                        int64_t ip[SP];
                        uint32_t ik[SP];
                        int np = 0;
#pragma unroll
                        for (int n = 0; n < SP; ++n) {
                            const uint32_t vn = v + 256u * n;
                            ip[n] = 0;
                            ik[n] = 0;
                            if (vn < tp) {
                                j = list_advance(a.run.pts, nrun, vn, j);
                                ik[n] = (uint32_t)a.run.inst[j];
                                ip[n] = (int64_t)a.pt_off[ik[n]] + (vn - a.run.pts[j]);
                                np = n + 1;
                            }
                        }
                        bscreen_gather_run<T, SP>(trun, a.P, a.cell, a.tt, a.sbuf, rk, dold, qk, ip, ik, np);
                    }
                }
            }
            if (!gsync(c, gen, &sflag[0])) return;
            tick(c, 1, tl);
            // -- the fused CG pass
            const uint64_t tp0 = wall_clock64();
            {
                StencilIO<T> io;
                io.r = rk;
                io.q = qk;
                io.rout = a.r + ((k + 1) & 1) * N;
                io.dold = dold;
                io.dnew = a.ring + ((k + 1) % CG_XM) * N;
                io.y = a.q + ((k + 1) & 1) * N;
                io.partial = a.partial;
                io.s = a.sbuf;
                for_bricks(a.run, nrun, xs, &c->take[1][k & 1][0][0], &c->done[1][k & 1][0],
                           [&](uint32_t bid) { stencil_tile_brick<T, 16, OP_CG, true, true>(a.L, io, 0, trun, bid, lds); });
            }
            if (wg_tid() == 0)
                __hip_atomic_fetch_add(&c->pass_busy, wall_clock64() - tp0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!gsync(c, gen, &sflag[0])) return;
            tick(c, 2, tl);
            // (the run list is rebuilt after workgroup 0's first fold, which follows)
            if (w == 0 && wg_tid() == 0) check_done(c, 1, k, a.run.brick[nrun]);
            // -- fold and scalar step per instance; the last one builds the lists
            for (int j = (int)w; j < nrun; j += (int)G) {
                const int kk = a.run.inst[j];
                const int64_t p0 = (int64_t)(a.tab.off[kk] >> 12), np = (int64_t)(a.tab.off[kk + 1] >> 12) - p0;
                double v4[4];
                for (int q = 0; q < 4; ++q) {
                    v4[q] = block_fold(a.partial + 4 * p0 + q, np, 4, ws);
                    __syncthreads();
                }
                if (wg_tid() == 0) {
                    CgState* st = a.tab.st + kk;
                    fin_cg(st, v4[0], v4[1], v4[2], v4[3], 0);
                    // the pending terms' directions sit in the ring slots of passes k - xn + 1 .. k
                    st->ring0 = (int32_t)(((st->it - (k + 1)) % CG_XM + CG_XM) % CG_XM);
                    if (st->done) {
                        if (MIXED || (st->status == PRX_OK && st->it > 0)) {
                            a.flag[kk] = 1;
                        } else {
                            a.pi[kk].fin = 1;
                            a.pi[kk].status = st->status;
                        }
                    }
                }
                finish_item<T>(a, nrun, &sflag[1], tmp);
            }
            if (!gsync(c, gen, &sflag[0])) return;
            tick(c, 3, tl);
        }
        if (w == 0 && wg_tid() == 0) {
            c->rounds = k + 1;
            if (wall_clock64() - t0 > a.deadline) st_ctl(&c->err, 3);  // the others leave at their next barrier
        }
    }
}

// FP32 (the mixed solve): three workgroups per compute unit (its LDS holds
// the FP64 tile of the residual pass, 47 KB), 168 VGPRs with 42 spilled.  At
// two per compute unit nothing spills (199 VGPRs), but the CG pass phase of
// the batched mixed solve took 5.3 s against 4.3 s at three: the pass needs
// the waves to hide memory latency more than it needs the registers.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_persist32(const PArgs<float>* __restrict__ pa) {
    persist_main<float>(pa);
}
// FP64: its 46 KB tile allows three workgroups per compute unit at most; the
// one-pass kernel of the same pass runs two (175 VGPRs), and this one is held
// to two as well (237 VGPRs, no spills).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_persist64(const PArgs<double>* __restrict__ pa) {
    persist_main<double>(pa);
}

template <typename T>
void launch_persistent_cg(const PArgs<T>& a0, const NodeList* scr, void* dargs, int* blocks, hipStream_t s) {
    PArgs<T> a = a0;
    a.L = batch_args<T>(scr, 12);
    a.L64 = batch_args<double>(scr, 12);
    a.has_scr = (scr && scr->T > 0 && a.P > 0) ? 1 : 0;
    int dev = 0, ncu = 0, per = 0;
    PRX_HIP(hipGetDevice(&dev));
    PRX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const void* kern = sizeof(T) == 4 ? (const void*)k_persist32 : (const void*)k_persist64;
    PRX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 256, 0));
    if (per <= 0 || ncu <= 0) throw Error(PRX_ERR_DEVICE, "the persistent CG kernel does not fit on a compute unit");
    const int grid = per * ncu;
    int khz = 0;
    PRX_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    // 60 s; PRX_PERSIST_DEADLINE_MS sets another (tests force the deadline exit with it)
    uint64_t deadline_ms = 60000;
    if (const char* e = std::getenv("PRX_PERSIST_DEADLINE_MS")) {
        char* end = nullptr;
        const unsigned long long v = std::strtoull(e, &end, 10);
        if (end != e && *end == '\0') deadline_ms = v;
    }
    a.deadline = (uint64_t)(khz > 0 ? khz : 100000) * deadline_ms;
    if (grid % 8) throw Error(PRX_ERR_DEVICE, "persistent CG: workgroup count not a multiple of the XCD count");
    PRX_HIP(hipMemcpyAsync(dargs, &a, sizeof(PArgs<T>), hipMemcpyHostToDevice, s));
    const PArgs<T>* pa = (const PArgs<T>*)dargs;
    void* args[] = {&pa};
    PRX_HIP(hipLaunchCooperativeKernel(kern, dim3(grid), dim3(256), args, 0, s));
    PRX_HIP(hipStreamSynchronize(s));  // the argument block and a is on the host stack
    *blocks = grid;
}
template void launch_persistent_cg<float>(const PArgs<float>&, const NodeList*, void*, int*, hipStream_t);
template void launch_persistent_cg<double>(const PArgs<double>&, const NodeList*, void*, int*, hipStream_t);

}  // namespace prx
