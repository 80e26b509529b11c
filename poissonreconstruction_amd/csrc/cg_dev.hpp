// Device pieces of the CG iterations shared by the per-pass kernels and the
// persistent batched solve: the deterministic block and multi-stage folds,
// the scalar steps of solve_cg (src/solvers.rs:36-64) on the folded sums, and
// the batched per-vector passes (x terms, commit, dot, screening gather).
// Contraction to FMA is off in every function here (as in k_grid.hip and
// k_points.hip, where these pieces were first written), so every file that
// uses them rounds them alike.
#pragma once

#include "kernels.hpp"

namespace prx {

// The thread's index in its workgroup, passed through an empty asm statement:
// nothing computed from it can be hoisted out of a loop around the caller, so
// the persistent CG kernel's round loop does not carry these values (in
// registers the tile pass needs) from one phase to the next.  The value is
// threadIdx.x.
__device__ __forceinline__ int wg_tid() {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Deterministic block sum: 64-lane butterfly, then waves in order.
__device__ inline double block_sum_256(double v, double* wsum) {
#pragma clang fp contract(off)
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = wg_tid() & 63, wave = wg_tid() >> 6;
    if (lane == 0) wsum[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (wg_tid() == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += wsum[w];
    return s;
}

// beta of the next direction, with the breakdown and iteration-cap tests
__device__ inline void fin_beta(CgState* st, double beta) {
#pragma clang fp contract(off)
    if (beta != beta || beta == INFINITY || beta == -INFINITY) {
        st->done = 1;
        st->status = PRX_ERR_BREAKDOWN;
        st->iters = st->max_iters;
        return;
    }
    st->beta = beta;
    if (st->it >= st->max_iters) {
        st->done = 1;
        st->iters = st->max_iters;
        st->status = PRX_ERR_NOT_CONVERGED;
    }
}

// The scalar step of solve_cg (src/solvers.rs:36-64) on the folded sum.
// the vector update of an iteration has run: account for the lazy x term
// (dir: ping-pong index of the iteration's direction)
__device__ inline void lazy_step(CgState* st, bool full, int dir) {
#pragma clang fp contract(off)
    st->dir_buf = dir;
    if (full || !st->lazy || st->xpend) {
        st->xpend = 0;
    } else {
        st->xpend = 1;
        st->alpha_pend = st->alpha;
    }
}

__device__ inline void fin_step(CgState* st, double sum, int op, int slot) {
#pragma clang fp contract(off)
    switch (op) {
        case FIN_SLOT: st->slot[slot] = sum; break;
        case FIN_INIT_REF:  // delta0 = r.r ; stop when delta <= delta0*eps^2
        case FIN_INIT_TRUE: // stop when delta <= ||b||^2 eps^2
            st->delta = sum;
            st->delta0 = sum;
            st->it = 0;
            st->iters = 0;
            st->status = 0;
            st->refresh_due = 0;
            st->xpend = 0;
            st->thr = ((op == FIN_INIT_REF ? sum : st->bnorm2) * st->eps) * st->eps;
            st->done = (sum == 0.0 || sum <= st->thr) ? 1 : 0;
            break;
        case FIN_ALPHA:
            st->dq = sum;
            if (sum == 0.0) {
                st->done = 1;
                st->iters = st->it + 1;
            } else {
                st->alpha = st->delta / sum;
            }
            break;
        case FIN_BETA:
        case FIN_BETA_FULL: {
            lazy_step(st, op == FIN_BETA_FULL, slot);
            double dold = st->delta;
            st->delta_prev = dold;
            st->delta = sum;
            st->it += 1;
            if (sum <= st->thr) {
                st->done = 1;
                st->iters = st->it;
                break;
            }
            double beta = sum / dold;
            if (beta != beta || beta == INFINITY || beta == -INFINITY) {
                st->done = 1;
                st->status = PRX_ERR_BREAKDOWN;
                st->iters = st->max_iters;
                break;
            }
            st->beta = beta;
            if (st->it >= st->max_iters) {
                st->done = 1;
                st->iters = st->max_iters;
                st->status = PRX_ERR_NOT_CONVERGED;
            }
            break;
        }
        case FIN_PINIT:  // sum = r0.z0 ; slot[0] holds |r0|^2 ; stop when |r|^2 <= ||b||^2 eps^2
            st->delta = sum;
            st->delta0 = sum;
            st->delta_prev = sum;
            st->beta = 0.0;
            st->it = 0;
            st->iters = 0;
            st->status = 0;
            st->refresh_due = 0;
            st->xpend = 0;
            st->thr = (st->bnorm2 * st->eps) * st->eps;
            st->slot[1] = st->slot[0];
            st->done = (sum == 0.0 || st->slot[0] <= st->thr) ? 1 : 0;
            break;
        case FIN_PRES:  // sum = |r|^2 after the update of iteration it
            lazy_step(st, false, slot);
            st->slot[1] = sum;
            st->it += 1;
            if (sum != sum) {
                st->done = 1;
                st->status = PRX_ERR_BREAKDOWN;
                st->iters = st->max_iters;
            } else if (sum <= st->thr) {
                st->done = 1;
                st->iters = st->it;
            } else if (st->it >= st->max_iters) {
                st->done = 1;
                st->iters = st->max_iters;
                st->status = PRX_ERR_NOT_CONVERGED;
            }
            break;
        case FIN_PBETA: {  // sum = r.z of the new residual
            const double beta = sum / st->delta;
            if (beta != beta || beta == INFINITY || beta == -INFINITY) {
                st->done = 1;
                st->status = PRX_ERR_BREAKDOWN;
                st->iters = st->max_iters;
                break;
            }
            st->beta = beta;
            st->delta_prev = st->delta;
            st->delta = sum;
            break;
        }
        case FIN_FRESH:  // sum = |b - A x|^2 of the refreshed residual (x complete)
            st->rpend = 0;
            st->xn = 0;
            st->delta = sum;
            if (sum <= st->thr) {
                st->done = 1;
                st->iters = st->it;
                break;
            }
            fin_beta(st, sum / st->delta_prev);
            break;
    }
}

// The fused CG step (src/solvers.rs:36-64 with the two reductions of an
// iteration merged): the pass formed r_k (the pending update of the previous
// iteration), d_k = r_k + beta d_{k-1} and q_k = A d_k and summed d_k.q_k,
// r_k.q_k, q_k.q_k and r_k.r_k.  delta_k = r_k.r_k and alpha_k = delta_k /
// d_k.q_k as in the two-pass form; the update r_k - alpha_k q_k is left to
// the next pass, so the stopping test and beta_k take its |r|^2 from the
// expansion delta_k - 2 alpha_k r_k.q_k + alpha_k^2 q_k.q_k.  The next pass
// sums its own r.r again, so the expansion's rounding never accumulates over
// iterations (carried from one to the next, it would: its absolute error,
// about eps * the early |r|^2, swamps |r|^2 once the residual has fallen by
// ~1e-6).  x: CG_XM pending terms were added before the pass (xn == CG_XM
// -> 0), then alpha_k d_k is pending.  refresh: the caller recomputes r = b -
// A x next (FIN_FRESH decides).
__device__ inline void fin_cg(CgState* st, double dq, double rq, double qq, double rr, int refresh) {
#pragma clang fp contract(off)
    if (st->xn == CG_XM) st->xn = 0;
    st->dq = dq;
    st->delta = rr;
    if (dq == 0.0) {
        st->done = 1;
        st->iters = st->it + 1;
        return;
    }
    const double a = st->delta / dq;
    st->alpha = a;
    st->xa[st->xn] = a;
    st->xn += 1;
    st->rpend = 1;
    const double g = (st->delta - 2.0 * a * rq) + a * a * qq;
    st->delta_prev = st->delta;
    st->delta = g;
    st->it += 1;
    // a refresh of r follows: FIN_FRESH tests the refreshed residual first,
    // then forms beta and applies the iteration cap, as the two-pass form does
    // (src/solvers.rs:49-63)
    if (refresh) return;
    if (g <= st->thr) {
        st->done = 1;
        st->iters = st->it;
        return;
    }
    fin_beta(st, g / st->delta_prev);
}

// NTH threads (the block) fold np partials in a fixed order (16 strided
// accumulators, a fixed pairwise tree, then the block sum), so a reduction
// is reproducible run to run; 16 independent loads per thread are in flight
// per round.  STRIDE: the partials of one sum are every STRIDE-th value.
template <int STRIDE = 1, int NTH = 1024>
__device__ inline double fin_sum(const double* partial, int64_t np, double* ws) {
#pragma clang fp contract(off)
    double a[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) a[u] = 0.0;
    int64_t i = wg_tid();
    for (; i + 15 * NTH < np; i += 16 * NTH) {
#pragma unroll
        for (int u = 0; u < 16; ++u) a[u] += partial[STRIDE * (i + u * NTH)];
    }
    for (; i < np; i += NTH) a[0] += partial[STRIDE * i];
#pragma unroll
    for (int w = 8; w >= 1; w >>= 1)
#pragma unroll
        for (int u = 0; u < w; ++u) a[u] += a[u + w];
    return block_sum_256(a[0], ws);
}

// FIN_CG first stage: block b folds the four interleaved sums of bricks
// [b*chunk, min(np, (b+1)*chunk)) in a fixed order (per thread in brick
// order, then the 64-lane butterfly, then the waves in order) into
// out[4b .. 4b+3], so the single-block fold reads 4 values per block
// instead of 4 per brick.
constexpr int FOLD_SPLIT = 256;  // more bricks than this: two stages (at most FOLD_BLOCKS first-stage blocks)
__device__ __host__ inline int fold_blocks(int np) {
    const int g = (np + 1023) / 1024;
    return np > FOLD_SPLIT ? (g < FOLD_BLOCKS ? g : FOLD_BLOCKS) : 0;
}
__device__ inline void fold4_block(const double* __restrict__ partial, int np, int b, double* __restrict__ out, double* ws) {
#pragma clang fp contract(off)
    typedef double nd2 __attribute__((ext_vector_type(2)));
    const int g = fold_blocks(np), chunk = (np + g - 1) / g;
    const int b0 = b * chunk, b1 = min(np, b0 + chunk);
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = b0 + (int)wg_tid(); i < b1; i += 256) {
        const nd2 u = *(const nd2*)(partial + 4 * (int64_t)i);
        const nd2 v = *(const nd2*)(partial + 4 * (int64_t)i + 2);
        a[0] += u[0];
        a[1] += u[1];
        a[2] += v[0];
        a[3] += v[1];
    }
    const int lane = wg_tid() & 63, wave = wg_tid() >> 6;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        for (int o = 32; o > 0; o >>= 1) a[j] += __shfl_xor(a[j], o, 64);
        if (lane == 0) ws[j * 4 + wave] = a[j];
    }
    __syncthreads();
    if (wg_tid() < 4) {
        double sum = 0.0;
        for (int w = 0; w < 4; ++w) sum += ws[wg_tid() * 4 + w];
        out[4 * b + wg_tid()] = sum;
    }
}

// the four sums of FIN_CG, second stage (or the only one for small grids)
template <int NTH = 1024>
__device__ inline void fold4_final(const double* partial, int np, const double* first, double* ws, double v[4]) {
#pragma clang fp contract(off)
    const int g = fold_blocks(np);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[j] = g ? fin_sum<4, NTH>(first + j, g, ws) : fin_sum<4, NTH>(partial + j, np, ws);
        __syncthreads();  // ws is reused by the next sum
    }
}

// ------------------------------------------------------------------ batched per-vector passes
// Every batched kernel processes instance k only while its run flag says so.
__device__ inline bool b_runs(const BatchTab& tab, uint32_t k) { return tab.run[(int64_t)k * tab.run_stride] == tab.run_when; }

// Fused CG x terms per instance (x + xa[0] d[0] + ... in FP64, rounded once,
// as k_x_terms).  A block covers 256 vectors inside one 4096-node brick, so
// it belongs to one instance.
template <typename T, int MODE>
__device__ inline void bx_terms_block(const BatchTab& tab, T* __restrict__ x, const XDirs<T>& dd, const T* __restrict__ ring,
                                      int64_t Ntot, uint32_t blk) {
#pragma clang fp contract(off)
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T);
    const int64_t v0 = (int64_t)blk * 256;
    const uint32_t k = tab.brick_inst[(uint64_t)(v0 * VEC) >> 12];
    if (!b_runs(tab, k)) return;
    const CgState* st = tab.st + k;
    const int xn = st->xn;
    if (MODE == 0 ? xn != CG_XM : xn == 0) return;
    const T* dp[CG_XM];
#pragma unroll
    for (int j = 0; j < CG_XM; ++j) {
        if (MODE == 1) {
            const int64_t sl = ((st->it - (int64_t)st->ring0 - xn + j + 1) % CG_XM + CG_XM) % CG_XM;
            dp[j] = ring + sl * Ntot;
        } else {
            dp[j] = dd.d[j];
        }
    }
    double a[CG_XM];
#pragma unroll
    for (int j = 0; j < CG_XM; ++j) a[j] = st->xa[j];
    const int64_t g = (v0 + wg_tid()) * VEC;
    const NV xv = ld_nt(x + g);
    NV dv[CG_XM];
#pragma unroll
    for (int j = 0; j < CG_XM; ++j)
        if (j < xn && dp[j]) dv[j] = ld_nt(dp[j] + g);
    NV xo;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        double v = (double)xv[e];
#pragma unroll
        for (int j = 0; j < CG_XM; ++j)
            if (j < xn && dp[j]) v += a[j] * (double)dv[j][e];
        xo[e] = (T)v;
    }
    st_nt(x + g, xo);
}

// The due x terms (CG_XM of them) of NB blocks at once, for a loop that owns
// few waves per compute unit: every block's x and direction loads are issued
// before any sum, so NB times as many bytes are in flight per wave.  Block b
// (b < nb) is blk[b], of instance inst[b], which runs with xn == CG_XM; the
// directions are the same ring slots for every instance (dd).  Each block's
// result equals bx_terms_block<T, 0>'s bit for bit.
template <typename T, int NB>
__device__ inline void bx_terms_due(const CgState* __restrict__ st, T* __restrict__ x, const XDirs<T>& dd,
                                    const uint32_t (&blk)[NB], const uint32_t (&inst)[NB], int nb) {
#pragma clang fp contract(off)
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T);
    NV xv[NB], dv[NB][CG_XM];
    int64_t g[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        g[b] = ((int64_t)blk[b] * 256 + wg_tid()) * VEC;
        if (b < nb) {
            xv[b] = ld_nt(x + g[b]);
#pragma unroll
            for (int j = 0; j < CG_XM; ++j) dv[b][j] = ld_nt(dd.d[j] + g[b]);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (b >= nb) break;
        const CgState* sk = st + inst[b];
        NV xo;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            double v = (double)xv[b][e];
#pragma unroll
            for (int j = 0; j < CG_XM; ++j) v += sk->xa[j] * (double)dv[b][j][e];
            xo[e] = (T)v;
        }
        st_nt(x + g[b], xo);
    }
}

// x = x2 and e = 0 on the nodes of the running instances: the fold of a
// batched residual-update pass, which wrote x + e into x2 (a block covers
// 1024 nodes inside one 4096-node brick, so it belongs to one instance).
__device__ inline void bcommit_block(const BatchTab& tab, double* __restrict__ x, const double* __restrict__ x2, float* __restrict__ e,
                                     uint32_t blk) {
    typedef double nd2 __attribute__((ext_vector_type(2)));
    typedef float nf4 __attribute__((ext_vector_type(4)));
    const int64_t j = ((int64_t)blk * 256 + wg_tid()) * 4;
    const uint32_t k = tab.brick_inst[((uint64_t)blk * 1024) >> 12];
    if (!b_runs(tab, k)) return;
    const nd2 a = *(const nd2*)(x2 + j), b = *(const nd2*)(x2 + j + 2);
    __builtin_nontemporal_store(a, (nd2*)(x + j));
    __builtin_nontemporal_store(b, (nd2*)(x + j + 2));
    __builtin_nontemporal_store(nf4{0.f, 0.f, 0.f, 0.f}, (nf4*)(e + j));
}

template <typename T>
__device__ inline void bdot_block(const BatchTab& tab, const T* __restrict__ a, const T* __restrict__ b, double* partial, uint32_t blk,
                                  double* ws) {
#pragma clang fp contract(off)
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T);
    const int64_t j = ((int64_t)blk * 256 + wg_tid()) * VEC;
    const uint32_t k = tab.brick_inst[(uint64_t)blk * 256 * VEC >> 12];
    if (!b_runs(tab, k)) return;
    const NV av = ld_v(a + j), bv = ld_v(b + j);
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < VEC; ++e) acc += (double)av[e] * (double)bv[e];
    const double sm = block_sum_256(acc, ws);
    if (wg_tid() == 0) partial[blk] = sm;
}

// The same gather over a batch: point i belongs to instance pts.inst[i]
// (its grid size, wh3, first node and, in MODE 2, beta); it runs while the
// instance's run flag says so.
template <typename T, int MODE>
__device__ inline void bscreen_gather_point(const BatchTab& tab, int64_t P, const uint32_t* __restrict__ cell, const T* __restrict__ tt,
                                            const uint32_t* __restrict__ pinst, const T* __restrict__ x, const double* __restrict__ x64,
                                            const float* __restrict__ e32, T* s, const T* __restrict__ r, const T* __restrict__ dd,
                                            const T* __restrict__ qq, int64_t i) {
#pragma clang fp contract(off)
    if (i >= P) return;
    const uint32_t k = pinst[i];
    if (tab.run[(int64_t)k * tab.run_stride] != tab.run_when) return;
    const BCoef cf = tab.coef[k];
    const uint64_t ib = tab.off[k];
    const double bt = MODE >= 2 ? tab.st[k].beta : 0.0;
    const double al = MODE == 3 ? tab.st[k].alpha : 0.0;
    const int rp = MODE == 3 ? tab.st[k].rpend : 0;
    const uint32_t res = cf.res;
    const uint32_t c[3] = {cell[i], cell[P + i], cell[2 * P + i]};
    const T t[3] = {tt[i], tt[P + i], tt[2 * P + i]};
    // as k_screen_gather: all loads first (out-of-grid corners read the
    // instance's node 0 with weight 0), then the sum in corner order
    uint64_t m[8];
    T w[8], v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const uint32_t ox = c[0] + (q & 1), oy = c[1] + ((q >> 1) & 1), oz = c[2] + (q >> 2);
        const bool in = ox < res && oy < res && oz < res;
        const T bx = (q & 1) ? t[0] : T(1) - t[0];
        const T by = ((q >> 1) & 1) ? t[1] : T(1) - t[1];
        const T bz = (q >> 2) ? t[2] : T(1) - t[2];
        w[q] = in ? bx * by * bz : T(0);
        m[q] = ib + (in ? morton3(ox, oy, oz) : 0u);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (MODE == 0) v[q] = x[m[q]];
        else if (MODE == 1) v[q] = (T)(x64[m[q]] + (double)e32[m[q]]);
        else if (MODE == 2) v[q] = dir_form(bt, dd[m[q]], r[m[q]]);
        else v[q] = cg_form(bt, al, rp, dd[m[q]], r[m[q]], qq[m[q]]);  // q is read always and used when rp
    }
    T acc = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) acc += w[q] * v[q];
    const T wh3 = sizeof(T) == 8 ? (T)cf.wh3_64 : (T)cf.wh3_32;
    s[i] = wh3 * acc;
}

// bscreen_gather_point<T, 3> for NP points at once whose instances are known
// to run (kk[n] of point ip[n], n < np): all points' loads are issued before
// any sum, for a loop that owns few waves per compute unit.  Each s[ip] equals
// bscreen_gather_point<T, 3>'s bit for bit.
template <typename T, int NP>
__device__ inline void bscreen_gather_run(const BatchTab& tab, int64_t P, const uint32_t* __restrict__ cell,
                                          const T* __restrict__ tt, T* s, const T* __restrict__ r, const T* __restrict__ dd,
                                          const T* __restrict__ qq, const int64_t (&ip)[NP], const uint32_t (&kk)[NP], int np) {
#pragma clang fp contract(off)
    uint64_t m[NP][8];
    T w[NP][8], v[NP][8], wh3[NP];
    double bt[NP], al[NP];
    int rp[NP];
#pragma unroll
    for (int n = 0; n < NP; ++n) {
        if (n >= np) break;
        const uint32_t k = kk[n];
        const int64_t i = ip[n];
        const BCoef cf = tab.coef[k];
        const uint64_t ib = tab.off[k];
        bt[n] = tab.st[k].beta;
        al[n] = tab.st[k].alpha;
        rp[n] = tab.st[k].rpend;
        wh3[n] = sizeof(T) == 8 ? (T)cf.wh3_64 : (T)cf.wh3_32;
        const uint32_t res = cf.res;
        const uint32_t c[3] = {cell[i], cell[P + i], cell[2 * P + i]};
        const T t[3] = {tt[i], tt[P + i], tt[2 * P + i]};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t ox = c[0] + (q & 1), oy = c[1] + ((q >> 1) & 1), oz = c[2] + (q >> 2);
            const bool in = ox < res && oy < res && oz < res;
            const T bx = (q & 1) ? t[0] : T(1) - t[0];
            const T by = ((q >> 1) & 1) ? t[1] : T(1) - t[1];
            const T bz = (q >> 2) ? t[2] : T(1) - t[2];
            w[n][q] = in ? bx * by * bz : T(0);
            m[n][q] = ib + (in ? morton3(ox, oy, oz) : 0u);
        }
    }
#pragma unroll
    for (int n = 0; n < NP; ++n) {
        if (n >= np) break;
#pragma unroll
        for (int q = 0; q < 8; ++q) v[n][q] = cg_form(bt[n], al[n], rp[n], dd[m[n][q]], r[m[n][q]], qq[m[n][q]]);
    }
#pragma unroll
    for (int n = 0; n < NP; ++n) {
        if (n >= np) break;
        T acc = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) acc += w[n][q] * v[n][q];
        s[ip[n]] = wh3[n] * acc;
    }
}

}  // namespace prx
