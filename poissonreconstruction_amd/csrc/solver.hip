// Host side of libprx: the FEMTree handle (device buffers, stream, events),
// the cascadic solve with its CG / multicolour-GS / mixed-precision
// refinement drivers, and the extern "C" boundary declared in include/prx.h.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>

#include "tree.hpp"

namespace prx {

thread_local std::string g_err;

void set_device(prx_tree* t) { PRX_HIP(hipSetDevice(t->device)); }

// ------------------------------------------------------------------ operator helpers
template <typename T>
static T* sbuf_for(prx_tree* t);
template <>
double* sbuf_for<double>(prx_tree* t) { t->s64.ensure(std::max<int64_t>(t->P, 1)); return t->s64.p; }
template <>
float* sbuf_for<float>(prx_tree* t) { t->s32.ensure(std::max<int64_t>(t->P, 1)); return t->s32.p; }
// The face form's screening data: the cell lists (built once per splat) and
// one array of [0, N) q at the touched nodes (node-indexed; the other nodes
// stay 0, which the cell gather never reads), [T] the screening sums per
// touched node, [8C] the sums per cell corner.
static const CellLists& cells_for(prx_tree* t, const LevelView& L) {
    if (!t->cells_ok) {
        free_cell_lists(t->cells);
        build_cell_lists(t->cells, t->D, t->pts, *L.scr, t->stream);
        t->cells_ok = true;
    }
    return t->cells;
}
template <typename T>
static T* qt_for(prx_tree* t, const LevelView& L) {
    const CellLists& CL = cells_for(t, L);
    const int64_t n = L.N + std::max<int64_t>(L.scr->T, 1) + 8 * std::max<int64_t>(CL.C, 1);
    bool fresh;
    T* p;
    if constexpr (sizeof(T) == 8) { fresh = !t->qt64.p || t->qt64.n < n; t->qt64.ensure(n); p = (T*)t->qt64.p; }
    else { fresh = !t->qt32.p || t->qt32.n < n; t->qt32.ensure(n); p = (T*)t->qt32.p; }
    if (fresh || !t->qt_ok) {
        launch_fill<T>(p, T(0), n, t->stream);
        t->qt_ok = true;
    }
    return p;
}

// y = A x (OP_APPLY) or y = b - A x (OP_RESID); partial sums into t->partial.
template <typename T>
static int op_pass(prx_tree* t, const LevelView& L, int op, const T* x, const T* b, T* y, const int32_t* done, double* partial) {
    StencilIO<T> io;
    io.x = x;
    io.b = b;
    io.y = y;
    io.partial = partial;
    io.done = done;
    if (L.scr) {
        T* sb = sbuf_for<T>(t);
        launch_screen_gather<T>(L, x, nullptr, nullptr, sb, t->stream);
        io.s = sb;
    }
    int nb = 0;
    launch_stencil<T>(L, op, io, &nb, t->tile, t->stream);
    return nb;
}

// x_new = x + e32 ; r32 = (b - A x_new) * rscale ; partial = |b - A x_new|^2.
// x_new goes to the second solution buffer, which then becomes t->x.
static int op_resid_update(prx_tree* t, const LevelView& L, const float* e32, const double* b, float* r32, double rscale) {
    t->x2.ensure(t->N);
    const double* x64 = t->x.p;
    StencilIO<double> io;
    io.x64 = x64;
    io.x_out = t->x2.p;
    io.e32 = e32;
    io.b = b;
    io.r32 = r32;
    io.rscale = rscale;
    io.partial = t->partial.p;
    if (L.scr) {
        double* sb = sbuf_for<double>(t);
        launch_screen_gather<double>(L, nullptr, x64, e32, sb, t->stream);
        io.s = sb;
    }
    int nb = 0;
    launch_stencil<double>(L, OP_RESID_UPD, io, &nb, t->tile, t->stream);
    t->x.swap(t->x2);
    return nb;
}

// Multicolour GS sweeps: one brick-tiled pass per colour (depth >= 3), the
// per-node colour kernel on smaller grids.
template <typename T>
static void gs_sweeps(prx_tree* t, const LevelView& L, T* x, const T* b, int sweeps) {
    T* sb = L.scr ? sbuf_for<T>(t) : nullptr;
    for (int k = 0; k < sweeps; ++k)
        for (int c = 0; c < 8; ++c) {
            if (L.scr) launch_screen_gather<T>(L, x, nullptr, nullptr, sb, t->stream);
            if (L.D >= 3) {
                StencilIO<T> io;
                io.x = x;
                io.y = x;
                io.b = b;
                io.s = sb;
                io.color = c;
                int nb = 0;
                launch_stencil<T>(L, OP_GS, io, &nb, t->tile, t->stream);
            } else {
                launch_gs_color<T>(L, c, x, b, sb, t->stream);
            }
        }
}

// In-solve timing of the fused CG pass: the durations of the pairs recorded
// since the last read (their launches have completed: the caller has waited
// on the stream, or waits on each end event here).
static void ptime_read(prx_tree* t) {
    auto& P = t->ptime;
    for (int i = 0; i < P.used; ++i) {
        PRX_HIP(hipEventSynchronize(P.ev[2 * i + 1]));
        float ms = 0.f;
        PRX_HIP(hipEventElapsedTime(&ms, P.ev[2 * i], P.ev[2 * i + 1]));
        P.ms.push_back(ms);
    }
    P.used = 0;
}
// the event pair for the next timed launch (64 pairs in flight at most)
static hipEvent_t* ptime_pair(prx_tree* t) {
    auto& P = t->ptime;
    if (P.used == 64) ptime_read(t);
    if ((int)P.ev.size() < 2 * (P.used + 1)) {
        for (int i = 0; i < 2; ++i) {
            hipEvent_t e;
            PRX_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            P.ev.push_back(e);
        }
    }
    return &P.ev[2 * P.used++];
}

void read_state(prx_tree* t) {
    PRX_HIP(hipMemcpyAsync(t->st_host, t->st, sizeof(CgState), hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
}

static void init_state(prx_tree* t, double eps, uint64_t max_iters, double bnorm2) {
    CgState h{};
    h.eps = eps;
    h.max_iters = (int64_t)max_iters;
    h.bnorm2 = bnorm2;
    h.lazy = 1;
    *t->st_host = h;
    PRX_HIP(hipMemcpyAsync(t->st, t->st_host, sizeof(CgState), hipMemcpyHostToDevice, t->stream));
}

// Dot product of two device vectors, folded on the device into st->slot[k].
template <typename T>
static double dev_dot(prx_tree* t, const T* a, const T* b, int64_t n) {
    int nb = 0;
    launch_dot<T>(a, b, n, t->partial.p, &nb, t->stream);
    launch_fin(t->st, t->partial.p, nb, FIN_SLOT, 0, t->stream);
    read_state(t);
    return t->st_host->slot[0];
}

// solve_cg (src/solvers.rs:20-69) on level L, with the update d = r + beta d
// fused into the next operator pass (OP_DIR_APPLY).  The fused kernel reads
// the previous direction and writes the new one, so the two live in a
// ping-pong pair: iteration k reads dd[k&1] and writes dd[(k+1)&1].  Every
// kernel of an iteration returns at once when the device flag `done` is set,
// so chunks of iterations are enqueued without a host round-trip and the
// iteration count is exact.  Refresh r = b - A x on iterations refresh,
// 2*refresh, ... as the reference does every 10th (refresh 0 = never).
//
// Fused form (tiled levels, when second r and q buffers are given): one
// operator pass per iteration (OP_CG) also forms the previous iteration's
// residual update r - alpha q, and one fold (FIN_CG) yields alpha, the new
// |r|^2 and beta.  r and q then live in ping-pong pairs: pass k reads
// rr[k&1], qq[k&1] and writes rr[(k+1)&1], qq[(k+1)&1]; the directions live
// in a ring of CG_XM: pass k reads dd[k % CG_XM] and writes dd[(k+1) % CG_XM]
// (d_j sits in slot (j+1) % CG_XM).  x takes CG_XM terms at once, before the
// pass that overwrites the oldest direction.  Per iteration this reads r, q,
// d and writes r, d, q (24 bytes per node in FP32) plus 6 for x, where the
// two-pass form moves 28 plus 8.  In face form (FP32, depth >= 4) qq[k&1]
// hold the bricks' faces of q instead of q, and a pass moves 16 bytes per
// node plus 3 of faces (launch_cg_faces).
template <typename T>
struct Cg {
    prx_tree* t;
    LevelView L;
    T* x;
    const T* b;
    T* r;
    T* dd[CG_XM] = {};
    T* q;
    T* rr[2] = {nullptr, nullptr};
    T* qq[2] = {nullptr, nullptr};
    bool fused = false;
    bool faces = false;  // fused pass in face form: qq hold the bricks' faces of q, not q
    T* qt = nullptr;                // face form: q at the touched nodes (node-indexed)
    T* sc = nullptr;                // face form: the screening sums per touched node
    T* sig = nullptr;               // face form: the screening sums per cell corner
    const CellLists* cells = nullptr;
    int xn = 0;  // fused: x terms pending on the device (as FIN_CG counts them)
    int refresh = 0;
    uint64_t it = 0;  // completed iterations (the device's count after run())
    uint64_t applies = 0;

    // the residual the next iteration starts from (where a replacement goes)
    T* rcur() { return fused ? rr[it & 1] : r; }

    // r = b - A x (or r = b when x is known to be 0); delta0 = r.r
    void start(bool x_zero, int mode, double eps, uint64_t max_iters, double bnorm2) {
        const hipStream_t s = t->stream;
        init_state(t, eps, max_iters, bnorm2);
        int nb;
        if (x_zero) {
            PRX_HIP(hipMemcpyAsync(r, b, sizeof(T) * L.N, hipMemcpyDeviceToDevice, s));
            launch_dot<T>(r, r, L.N, t->partial.p, &nb, s);
        } else {
            nb = op_pass<T>(t, L, OP_RESID, x, b, r, nullptr, t->partial.p);
            applies++;
        }
        launch_fin(t->st, t->partial.p, nb, mode == PRX_RESID_TRUE ? FIN_INIT_TRUE : FIN_INIT_REF, 0, s);
        launch_fill<T>(dd[0], T(0), L.N, s);
        it = 0;
        xn = 0;
    }
    // start from a residual already in r (x = 0 implied): delta0 = r.r
    void start_from_residual(double eps, uint64_t max_iters) {
        const hipStream_t s = t->stream;
        init_state(t, eps, max_iters, 0.0);
        int nb;
        launch_dot<T>(r, r, L.N, t->partial.p, &nb, s);
        launch_fin(t->st, t->partial.p, nb, FIN_INIT_REF, 0, s);
        launch_fill<T>(dd[0], T(0), L.N, s);
        it = 0;
        xn = 0;
    }
    // the directions of the xn terms pending after `done_it` iterations
    XDirs<T> pending_dirs(uint64_t done_it, int n_terms) const {
        XDirs<T> d{};
        for (int i = 0; i < n_terms; ++i) d.d[i] = dd[(done_it - (uint64_t)n_terms + (uint64_t)i + 1) % CG_XM];
        return d;
    }
    void enqueue_fused(uint64_t k) {
        const hipStream_t s = t->stream;
        const int32_t* done = &t->st->done;
        T* dold = dd[k % CG_XM];  // d_{k-1}
        T* dnew = dd[(k + 1) % CG_XM];  // holds d_{k-CG_XM} until this pass writes d_k
        if (xn == CG_XM) launch_x_terms_full<T>(x, pending_dirs(k, CG_XM), t->st, L.N, s);
        StencilIO<T> io;
        io.r = rr[k & 1];
        io.q = qq[k & 1];
        io.rout = rr[(k + 1) & 1];
        io.dold = dold;
        io.dnew = dnew;
        io.y = qq[(k + 1) & 1];
        io.cst = t->st;
        io.partial = t->partial.p;
        io.done = done;
        int nb = 0;
        // the finest level's pass, timed on request
        hipEvent_t* tev = (t->ptime.on && L.D == t->D) ? ptime_pair(t) : nullptr;
        if (faces) {
            if (L.scr) {
                launch_cell_screening<T>(L, *cells, io.r, dold, qt, t->st, done, sig, sc, s);
                io.sc = sc;
                io.qt = qt;
            }
            if (tev) PRX_HIP(hipEventRecord(tev[0], s));
            launch_cg_faces<T>(L, io, &nb, s);
        } else {
            if (L.scr) {
                T* sb = sbuf_for<T>(t);
                launch_screen_gather<T>(L, nullptr, nullptr, nullptr, sb, s, io.r, dold, nullptr, io.q, t->st);
                io.s = sb;
            }
            if (tev) PRX_HIP(hipEventRecord(tev[0], s));
            launch_stencil<T>(L, OP_CG, io, &nb, t->tile, s);
        }
        if (tev) PRX_HIP(hipEventRecord(tev[1], s));
        applies++;
        const bool fresh = refresh > 0 && (k + 1) % (uint64_t)refresh == 0;
        launch_fin(t->st, t->partial.p, nb, FIN_CG, fresh ? 1 : 0, s);
        xn = xn == CG_XM ? 1 : xn + 1;
        if (fresh) {  // x completed, then r = b - A x replaces the pending update
            launch_x_flush<T>(x, pending_dirs(k + 1, xn), t->st, L.N, true, s);
            nb = op_pass<T>(t, L, OP_RESID, x, b, io.rout, done, t->partial.p);
            applies++;
            launch_fin(t->st, t->partial.p, nb, FIN_FRESH, 0, s);
            xn = 0;
        }
    }
    void enqueue(uint64_t k) {
        if (fused) {
            enqueue_fused(k);
            return;
        }
        const hipStream_t s = t->stream;
        const int32_t* done = &t->st->done;
        T* dold = dd[k & 1];
        T* dnew = dd[(k + 1) & 1];
        StencilIO<T> io;
        io.r = r;
        io.dold = dold;
        io.dnew = dnew;
        io.beta = &t->st->beta;
        io.y = q;
        io.partial = t->partial.p;
        io.done = done;
        if (L.scr) {
            T* sb = sbuf_for<T>(t);
            launch_screen_gather<T>(L, nullptr, nullptr, nullptr, sb, s, r, dold, &t->st->beta);
            io.s = sb;
        }
        int nb = 0;
        launch_stencil<T>(L, OP_DIR_APPLY, io, &nb, t->tile, s);
        applies++;
        launch_fin(t->st, t->partial.p, nb, FIN_ALPHA, 0, s);
        int fop = FIN_BETA;
        if (refresh > 0 && (k + 1) % (uint64_t)refresh == 0) {
            launch_axpy_alpha<T>(x, dold, dnew, t->st, L.N, s);
            nb = op_pass<T>(t, L, OP_RESID, x, b, r, done, t->partial.p);
            applies++;
            fop = FIN_BETA_FULL;
        } else {
            launch_cg_update<T>(x, r, dold, dnew, q, t->st, L.N, t->partial.p, &nb, s);
        }
        launch_fin(t->st, t->partial.p, nb, fop, (int)((k + 1) & 1), s);
    }
    // Iterate until the device flag is set or `cap` iterations are complete;
    // x is complete on return (a pending lazy term is flushed).
    void run(uint64_t cap) {
        const int chunk = L.N >= (1 << 21) ? 8 : 32;
        uint64_t k = it;
        // The host enqueues chunk n + 1 before it waits for the state behind
        // chunk n (a pinned snapshot copied after each chunk), so the device
        // does not idle while the host looks.  The chunk enqueued past
        // convergence returns at once: every kernel of an iteration tests the
        // done flag.
        CgState* snap = t->state_snapshots();
        hipEvent_t* ev = t->snapshot_events();
        for (int n = 0;; ++n) {
            for (int c = 0; c < chunk && k < cap; ++c) enqueue(k++);
            PRX_HIP(hipMemcpyAsync(snap + (n & 1), t->st, sizeof(CgState), hipMemcpyDeviceToHost, t->stream));
            PRX_HIP(hipEventRecord(ev[n & 1], t->stream));
            if (k >= cap) break;
            if (n >= 1) {
                PRX_HIP(hipEventSynchronize(ev[(n - 1) & 1]));
                if (snap[(n - 1) & 1].done) break;
            }
        }
        read_state(t);
        if (t->ptime.used) ptime_read(t);
        it = (uint64_t)t->st_host->it;
        if (fused && t->st_host->xn) {
            launch_x_flush<T>(x, pending_dirs(it, t->st_host->xn), t->st, L.N, false, t->stream);
            launch_clear_xpend(t->st, t->stream);
            t->st_host->xn = 0;
        } else if (t->st_host->xpend) {
            launch_flush_x<T>(x, dd[0], dd[1], t->st, L.N, t->stream);
            launch_clear_xpend(t->st, t->stream);
            t->st_host->xpend = 0;
        }
        xn = 0;
    }
    // Replace the recursive residual by `delta_true` = |r|^2 of a true
    // residual already written to r.  While the two residuals still agree
    // (within 4x) CG continues along its current direction with beta =
    // delta_true / delta_prev; otherwise it restarts from the true residual.
    void replace_residual(double delta_true, double thr) {
        CgState h = *t->st_host;
        const bool drifted = !(delta_true <= 4.0 * h.delta);
        h.beta = (!drifted && h.delta_prev > 0.0) ? delta_true / h.delta_prev : 0.0;
        h.delta = delta_true;
        h.thr = thr;
        h.done = 0;
        h.status = 0;
        h.rpend = 0;  // the replacement is the residual of the next iteration
        h.xn = 0;
        *t->st_host = h;
        PRX_HIP(hipMemcpyAsync(t->st, t->st_host, sizeof(CgState), hipMemcpyHostToDevice, t->stream));
    }
};

// d: 2 direction buffers of L.N, or CG_XM for the fused form on tiled levels,
// which also needs r1, q1 (second residual and product buffers)
template <typename T>
static Cg<T> make_cg(prx_tree* t, const LevelView& L, T* x, const T* b, T* r, T* d, T* q, int refresh, int nd, T* r1 = nullptr,
                     T* q1 = nullptr) {
    Cg<T> c;
    c.t = t;
    c.L = L;
    c.x = x;
    c.b = b;
    c.r = r;
    for (int i = 0; i < CG_XM; ++i) c.dd[i] = i < nd ? d + (int64_t)i * L.N : nullptr;
    c.q = q;
    c.refresh = refresh;
    c.fused = r1 && q1 && nd >= CG_XM && L.D >= 3 && t->cg_form != PRX_CG_TWO_PASS;
    // the face form: FP32, 16^3 bricks, and the screening's point data when the level has screening
    // by default from depth 8 (measured: 1.5-3% faster at depth 8, 7% at depth 9, 5-11% slower at depths
    // 6-7, where its two extra kernels per iteration weigh more than the bytes it saves)
    c.faces = c.fused && L.D >= 4 && t->tile == 16 &&
              ((t->cg_form == PRX_CG_FUSED && sizeof(T) == 4 && L.D >= 8) || t->cg_form == PRX_CG_FUSED_FACES) &&
              (!L.scr || (L.scr->spc && L.pts && L.pts->P > 0));
    if (c.faces && L.scr) {
        c.qt = qt_for<T>(t, L);
        c.sc = c.qt + L.N;
        c.sig = c.sc + std::max<int64_t>(L.scr->T, 1);
        c.cells = &t->cells;
    }
    c.rr[0] = r;
    c.rr[1] = r1;
    c.qq[0] = q;
    c.qq[1] = q1;
    return c;
}

// Faithful solve_cg: stop when delta <= delta0*eps^2 (REFERENCE) or, in TRUE
// mode, when the recursive residual meets eps*||b|| and the true residual
// confirms it (otherwise continue from the true residual).
template <typename T>
static CgOut run_cg(Cg<T>& cg, uint64_t max_iters, double eps, int mode, double bnorm2) {
    CgOut out;
    prx_tree* t = cg.t;
    cg.start(false, mode, eps, max_iters, bnorm2);
    double last_rt = INFINITY;
    bool stalled = false;
    for (;;) {
        cg.run(max_iters);
        const CgState h = *t->st_host;
        if (mode != PRX_RESID_TRUE || !h.done || h.status != PRX_OK || h.iters == 0) break;
        int nb = op_pass<T>(t, cg.L, OP_RESID, cg.x, cg.b, cg.rcur(), nullptr, t->partial.p);
        cg.applies++;
        launch_fin(t->st, t->partial.p, nb, FIN_SLOT, 1, t->stream);
        read_state(t);
        const double rt = t->st_host->slot[1];
        // stop when the true residual meets the target, or stops improving
        // (the attainable accuracy of the arithmetic has been reached)
        stalled = rt > 0.25 * last_rt;
        if (rt <= h.thr || cg.it >= max_iters || stalled) {
            t->st_host->delta = rt;
            if (rt > h.thr) t->st_host->status = PRX_ERR_NOT_CONVERGED;
            break;
        }
        last_rt = rt;
        cg.replace_residual(rt, h.thr);
    }
    const CgState& h = *t->st_host;
    out.iters = h.done ? (uint64_t)h.iters : cg.it;
    if (h.done && h.status == PRX_OK && mode == PRX_RESID_TRUE) out.iters = cg.it;
    out.status = h.done ? h.status : PRX_ERR_NOT_CONVERGED;
    const double den = mode == PRX_RESID_TRUE ? bnorm2 : h.delta0;
    out.rel = den > 0.0 ? std::sqrt(h.delta / den) : 0.0;
    out.applies = cg.applies;
    return out;
}

// ------------------------------------------------------------------ cascadic solve
struct SolveCfg {
    uint32_t base;
    int gs;
    uint64_t cg_iters;
    double eps;
    prx_solve_opts o;
};

// Coarsest level: FP64 CG.  The reference's test is ||r|| <= eps ||r0|| with
// r0 the residual after the GS sweeps (src/solvers.rs:36-58), which an eps
// near machine precision cannot reach once GS has already reduced r0; in
// TRUE residual mode the test is ||b - A x|| <= eps ||b|| with the same
// stagnation stop as the fine level.
static void coarse_solve(prx_tree* t, uint32_t cdepth, int gs, uint64_t iters, double eps, int mode, prx_solve_report& rep) {
    const hipStream_t s = t->stream;
    const int64_t nc = (int64_t)1 << (3 * cdepth);
    const double hc = 1.0 / (double)(1u << cdepth);
    const double sf = (double)((uint64_t)1 << (2 * (t->D - cdepth)));
    const double epw = t->pw * sf;
    const double wh3c = epw > 0.0 ? epw * (hc * hc * hc) / 1.5 : 0.0;
    LevelView L;
    L.D = cdepth;
    L.res = 1u << cdepth;
    L.N = nc;
    L.k64 = {stiffness_entry(hc, 0, 0, 0), stiffness_entry(hc, 1, 1, 0), stiffness_entry(hc, 1, 1, 1), epw > 0.0 ? wh3c * 1.005 : 0.0, 0.0};
    t->cb.ensure(nc); t->cx.ensure(nc); t->cr.ensure(2 * nc); t->cdd.ensure(CG_XM * nc); t->cq.ensure(2 * nc);
    launch_coarse_rhs(t->D, cdepth, t->nf.p, epw > 0.0 ? wh3c * 0.5 : 0.0, t->cb.p, &t->floor_list, s);
    launch_fill<double>(t->cx.p, 0.0, nc, s);
    gs_sweeps<double>(t, L, t->cx.p, t->cb.p, gs);
    const bool tru = mode == PRX_RESID_TRUE;
    const double bb = tru ? dev_dot<double>(t, t->cb.p, t->cb.p, nc) : 0.0;
    Cg<double> cg = make_cg<double>(t, L, t->cx.p, t->cb.p, t->cr.p, t->cdd.p, t->cq.p, tru ? 0 : 10, CG_XM, t->cr.p + nc,
                                    t->cq.p + nc);
    CgOut c = run_cg<double>(cg, iters, eps, tru ? PRX_RESID_TRUE : PRX_RESID_REFERENCE, bb);
    rep.coarse_iters = c.iters;
    rep.coarse_rel = c.rel;
    rep.applies += c.applies;
}

// FP64 true relative residual ||b - A x|| / ||b|| of the current solution.
static double true_rel(prx_tree* t) {
    const LevelView L = t->fine_view();
    t->r64.ensure(t->N);
    const double bb = dev_dot<double>(t, t->b.p, t->b.p, t->N);
    int nb = op_pass<double>(t, L, OP_RESID, t->x.p, t->b.p, t->r64.p, nullptr, t->partial.p);
    launch_fin(t->st, t->partial.p, nb, FIN_SLOT, 0, t->stream);
    read_state(t);
    const double rr = t->st_host->slot[0];
    return bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
}

static void fine_solve_fp64(prx_tree* t, const SolveCfg& c, uint64_t cap, prx_solve_report& rep) {
    const LevelView L = t->fine_view();
    const int64_t n = t->N;
    t->r64.ensure(2 * n); t->d64.ensure(CG_XM * n); t->q64.ensure(2 * n);
    {
        Timer tm(t->stream);
        gs_sweeps<double>(t, L, t->x.p, t->b.p, c.gs);
        rep.ms_fine_gs = tm.ms();
    }
    Timer tm(t->stream);
    double bb = 0.0;
    if (c.o.residual_mode == PRX_RESID_TRUE) bb = dev_dot<double>(t, t->b.p, t->b.p, n);
    Cg<double> cg = make_cg<double>(t, L, t->x.p, t->b.p, t->r64.p, t->d64.p, t->q64.p, c.o.refresh_period, CG_XM, t->r64.p + n,
                                    t->q64.p + n);
    CgOut o = run_cg<double>(cg, cap, c.eps, c.o.residual_mode, bb);
    rep.ms_fine_cg = tm.ms();
    rep.fine_iters = o.iters;
    rep.fine_rel = o.rel;
    // delta0: |b - A x|^2 where the CG started (after smoothing); replacements keep it
    if (bb > 0.0) rep.fine_r0_rel = std::sqrt(t->st_host->delta0 / bb);
    rep.applies += o.applies;
    rep.outer_iters = 1;
    rep.status = o.status;
}

// Mixed precision: CG runs on the correction e (FP32 vectors, the FP32
// operator scaled by 1/h, FP64 scalars and reductions) while the FP64 solution
// x stays in HBM untouched.  Whenever the FP32 recursive residual has fallen
// by the factor inner_tol since the last replacement (or claims convergence),
// x += e is folded into the FP64 solution, the FP64 true residual b - A x is
// formed in the same pass and replaces the FP32 residual, and CG continues
// along its current direction (a "reliable update"), so the Krylov space is
// kept across the FP64 corrections.  Convergence is decided on the FP64 true
// residual only.
static void fine_solve_mixed(prx_tree* t, const SolveCfg& c, uint64_t cap, prx_solve_report& rep) {
    const LevelView L = t->fine_view();
    const int64_t n = t->N;
    const hipStream_t s = t->stream;
    t->e32.ensure(n); t->ri32.ensure(2 * n); t->d32.ensure(CG_XM * n); t->q32.ensure(2 * n);
    const double ih = 1.0 / t->h;
    const double theta = c.o.inner_tol > 0.0 && c.o.inner_tol < 1.0 ? c.o.inner_tol : 0.05;
    const double bb = dev_dot<double>(t, t->b.p, t->b.p, n);
    // smoothing on the correction: e = GS^k(0; A/h, r/h), x += e
    {
        Timer tm(s);
        launch_fill<float>(t->e32.p, 0.0f, n, s);
        op_resid_update(t, L, t->e32.p, t->b.p, t->ri32.p, ih);
        if (c.gs > 0) gs_sweeps<float>(t, L, t->e32.p, t->ri32.p, c.gs);
        rep.ms_fine_gs = tm.ms();
    }
    Timer tm(s);
    int nb = op_resid_update(t, L, t->e32.p, t->b.p, t->ri32.p, ih);
    rep.applies += 2;
    launch_fin(t->st, t->partial.p, nb, FIN_SLOT, 0, s);
    read_state(t);
    double rr = t->st_host->slot[0];
    const double r0 = rr;
    const double thr = c.o.residual_mode == PRX_RESID_TRUE ? (bb * c.eps) * c.eps : (r0 * c.eps) * c.eps;
    const double thr32 = thr * ih * ih;  // the same test on the scaled FP32 residual
    uint64_t outer = 0;
    int status = PRX_OK;
    Cg<float> cg = make_cg<float>(t, L, t->e32.p, nullptr, t->ri32.p, t->d32.p, t->q32.p, 0, CG_XM, t->ri32.p + n, t->q32.p + n);
    if (rr > thr && rr > 0.0) {
        launch_fill<float>(t->e32.p, 0.0f, n, s);
        cg.start_from_residual(1.0, cap);
        read_state(t);
        cg.replace_residual(t->st_host->delta, std::max(theta * theta * t->st_host->delta, thr32));
        for (;;) {
            cg.run(cap);
            const CgState h = *t->st_host;
            // fold e into x and replace the residual by the FP64 true one
            nb = op_resid_update(t, L, t->e32.p, t->b.p, cg.rcur(), ih);
            rep.applies++;
            launch_fin(t->st, t->partial.p, nb, FIN_SLOT, 0, s);
            read_state(t);
            rr = t->st_host->slot[0];
            ++outer;
            if (rr <= thr) break;
            if (!h.done || h.status == PRX_ERR_NOT_CONVERGED || cg.it >= cap || outer >= c.o.max_outer) {
                status = PRX_ERR_NOT_CONVERGED;
                break;
            }
            if (h.status == PRX_ERR_BREAKDOWN) {
                status = PRX_ERR_BREAKDOWN;
                break;
            }
            launch_fill<float>(t->e32.p, 0.0f, n, s);
            const double dt = dev_dot<float>(t, cg.rcur(), cg.rcur(), n);
            *t->st_host = h;
            cg.replace_residual(dt, std::max(theta * theta * dt, thr32));
        }
    }
    rep.ms_fine_cg = tm.ms();
    rep.fine_iters = cg.it;
    rep.outer_iters = outer;
    rep.applies += cg.applies;
    const double den = c.o.residual_mode == PRX_RESID_TRUE ? bb : r0;
    rep.fine_rel = den > 0.0 ? std::sqrt(rr / den) : 0.0;
    if (c.o.residual_mode == PRX_RESID_TRUE && bb > 0.0) rep.fine_r0_rel = std::sqrt(r0 / bb);
    rep.status = status;
}

// ------------------------------------------------------------------ multigrid-preconditioned CG
// Preconditioner M ~ (L + wg I)^-1: one V-cycle over the dense grids of
// depths dc..D.  Each level smooths with damped Jacobi (nu sweeps from x = 0
// before the coarse correction, nu after), restricts the residual with P^T,
// prolongates the coarse correction with P (trilinear hats) and solves
// depth dc exactly with a dense inverse.  Level d uses the hat stiffness of
// its own h, which is the Galerkin product P^T L_{d+1} P of nested hat
// spaces up to the truncated boundary rows, and the lumped mass wg * 8^(D-d).
// The per-point screening is left out of M (its entries are ~wh3 = pw h^3/1.5,
// about 1e-5 of the stencil diagonal at pw = 4) and enters through the exact
// operator of the outer CG.  Equal pre- and post-smoothing, R = P^T and an
// exact coarse solve make M symmetric positive definite.
template <typename T>
static DArr<T>& mg_pool(prx_tree* t);
template <>
DArr<double>& mg_pool<double>(prx_tree* t) { return t->mg64; }
template <>
DArr<float>& mg_pool<float>(prx_tree* t) { return t->mg32; }

template <typename T>
struct Mg {
    uint32_t D = 0, dc = 0;
    int nu = 2;
    std::vector<LevelView> lv;  // [d - dc]
    std::vector<std::vector<double>> jw;  // per level: the nu smoothing weights / a_ii
    std::vector<T*> xa, xb, bv, rv;
    int nc = 0;                 // nodes of the coarsest grid
    uint64_t cycles = 0;
};

// Dense inverse of L + wg I on depth dc (nodes in Morton order), entries
// scaled by `scale`; Cholesky factorisation on the host, FP64.
static void mg_dense_inverse(uint32_t dc, double s, double wg, std::vector<double>& inv) {
    const int r = 1 << dc, n = r * r * r;
    std::vector<double> a((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        uint32_t x, y, z;
        demorton3((uint32_t)i, x, y, z);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int jx = (int)x + dx, jy = (int)y + dy, jz = (int)z + dz;
                    if (jx < 0 || jy < 0 || jz < 0 || jx >= r || jy >= r || jz >= r) continue;
                    const int nz = (dx != 0) + (dy != 0) + (dz != 0);
                    const double w = nz == 0 ? 32.0 * s + wg : nz == 2 ? -2.0 * s : nz == 3 ? -s : 0.0;
                    a[(size_t)i * n + morton3((uint32_t)jx, (uint32_t)jy, (uint32_t)jz)] = w;
                }
    }
    // a = C C^T (lower C in place)
    for (int j = 0; j < n; ++j) {
        double d = a[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= a[(size_t)j * n + k] * a[(size_t)j * n + k];
        if (!(d > 0.0)) throw Error(PRX_ERR_BREAKDOWN, "coarsest multigrid operator is not positive definite");
        const double cjj = std::sqrt(d);
        a[(size_t)j * n + j] = cjj;
        for (int i = j + 1; i < n; ++i) {
            double v = a[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) v -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
            a[(size_t)i * n + j] = v / cjj;
        }
    }
    // inverse column by column: C y = e_c, C^T x = y
    inv.assign((size_t)n * n, 0.0);
    std::vector<double> y(n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {
            double v = i == c ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) v -= a[(size_t)i * n + k] * y[k];
            y[i] = v / a[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double v = y[i];
            for (int k = i + 1; k < n; ++k) v -= a[(size_t)k * n + i] * inv[(size_t)k * n + c];
            inv[(size_t)i * n + c] = v / a[(size_t)i * n + i];
        }
    }
}

// Level views, Jacobi weights, vectors (one pool) and the coarsest inverse.
// The FP32 system is the FP64 one scaled by 1/h (as in the mixed solve).
template <typename T>
static Mg<T> mg_build(prx_tree* t, int nu, double omega) {
    Mg<T> m;
    m.D = t->D;
    m.dc = std::min<uint32_t>(t->D, 2);
    m.nu = std::max(1, nu);
    const int nl = (int)(m.D - m.dc) + 1;
    const double ih = 1.0 / t->h;
    const double scale = sizeof(T) == 4 ? ih : 1.0;
    m.lv.resize(nl);
    m.jw.resize(nl);
    // Step weights: omega > 0 is damped Jacobi (the same weight every step);
    // omega <= 0 takes the nu steps at the reciprocals of the roots of the
    // degree-nu Chebyshev polynomial on [lmax/alpha, lmax] (alpha = -omega, or
    // 6), which damps that upper part of the spectrum of D^-1 (L + wg I)
    // uniformly.  lmax = 1.5 is the maximum of the stencil's symbol over
    // D = 32 s (at frequency (pi, 0, 0)); the truncated rows only lower it.
    // The polynomial stays within [-1, 1] on [0, lmax], so the smoother
    // converges and the V-cycle is positive definite.
    std::vector<double> omega_i(m.nu, omega);
    if (!(omega > 0.0)) {
        const double lmax = 1.5, alpha = omega < 0.0 ? -omega : 6.0, lmin = lmax / alpha;
        const double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin);
        for (int i = 0; i < m.nu; ++i) omega_i[i] = 1.0 / (th + de * std::cos(M_PI * (2 * i + 1) / (2.0 * m.nu)));
    }
    m.xa.assign(nl, nullptr); m.xb.assign(nl, nullptr); m.bv.assign(nl, nullptr); m.rv.assign(nl, nullptr);
    int64_t total = 0;
    std::vector<int64_t> off(4 * nl, -1);
    for (int l = 0; l < nl; ++l) {
        const uint32_t d = m.dc + (uint32_t)l;
        const int64_t n = (int64_t)1 << (3 * d);
        const double h = 1.0 / (double)(1u << d);
        const double K0 = stiffness_entry(h, 0, 0, 0), Ke = stiffness_entry(h, 1, 1, 0), Kc = stiffness_entry(h, 1, 1, 1);
        const double wg = t->wg * (double)((uint64_t)1 << (3 * (m.D - d)));
        LevelView& L = m.lv[l];
        L.D = d;
        L.res = 1u << d;
        L.N = n;
        L.k64 = {K0, Ke, Kc, wg, 0.0};
        L.k32 = {(float)(K0 * ih), (float)(Ke * ih), (float)(Kc * ih), (float)(wg * ih), 0.0f};
        L.scr = nullptr;
        L.pts = nullptr;
        m.jw[l].resize(m.nu);
        for (int i = 0; i < m.nu; ++i) m.jw[l][i] = omega_i[i] / ((32.0 * -Kc + wg) * scale);
        // fine: xa = smoothing partner, rv = residual (input and output are the caller's)
        // middle: xa, xb, bv, rv ; coarsest: bv, xa
        const int cnt = l == nl - 1 ? 2 : (l == 0 ? 2 : 4);
        for (int k = 0; k < cnt; ++k) {
            off[4 * l + k] = total;
            total += n;
        }
    }
    DArr<T>& pool = mg_pool<T>(t);
    PRX_HIP(hipStreamSynchronize(t->stream));
    pool.ensure(total);
    for (int l = 0; l < nl; ++l) {
        T* p = pool.p;
        if (l == nl - 1) {
            m.xa[l] = p + off[4 * l];
            m.rv[l] = p + off[4 * l + 1];
        } else if (l == 0) {
            m.bv[l] = p + off[4 * l];
            m.xa[l] = p + off[4 * l + 1];
        } else {
            m.xa[l] = p + off[4 * l];
            m.xb[l] = p + off[4 * l + 1];
            m.bv[l] = p + off[4 * l + 2];
            m.rv[l] = p + off[4 * l + 3];
        }
    }
    m.nc = 1 << (3 * m.dc);
    const double hc = 1.0 / (double)(1u << m.dc);
    std::vector<double> inv;
    mg_dense_inverse(m.dc, -stiffness_entry(hc, 1, 1, 1) * scale, m.lv[0].k64.wg * scale, inv);
    PRX_HIP(hipStreamSynchronize(t->stream));
    t->mg_ainv.ensure((int64_t)inv.size());
    PRX_HIP(hipMemcpyAsync(t->mg_ainv.p, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    return m;
}

// z = M b on the finest level (b and z are N-vectors of the caller).  The
// first two pre-smoothing steps from x = 0 are one pass over b (OP_JACOBI2);
// the last post-smoothing pass of the finest level also leaves the partial
// sums of b.z in t->partial.  Returns their count (0: none were formed).
// done (optional): every kernel of the cycle returns at once when *done is
// set, so the PCG loop can enqueue a cycle before it knows whether it is needed.
template <typename T>
static int mg_vcycle(prx_tree* t, Mg<T>& m, const T* b_fine, T* z_fine, const int32_t* done = nullptr) {
    const hipStream_t s = t->stream;
    const int top = (int)m.lv.size() - 1;
    std::vector<T*> cur(top + 1, nullptr), oth(top + 1, nullptr);
    int nb_rz = 0;
    auto jacobi = [&](int l, const T* b, T*& c, T*& o, double w, bool rz) {
        StencilIO<T> io;
        io.x = c;
        io.b = b;
        io.y = o;
        io.jw = w;
        io.partial = rz ? t->partial.p : nullptr;
        io.done = done;
        int nb = 0;
        launch_stencil<T>(m.lv[l], OP_JACOBI, io, &nb, t->tile, s);
        if (rz) nb_rz = nb;
        std::swap(c, o);
    };
    // writes per level: pre nu-1 (nu >= 2) or 1, post nu; the first goes to
    // the buffer that makes the last one land in the level's output
    const int writes = m.nu >= 2 ? 2 * m.nu - 1 : 2;
    for (int l = top; l > 0; --l) {
        const LevelView& L = m.lv[l];
        const T* b = l == top ? b_fine : m.bv[l];
        T* out = l == top ? z_fine : m.xb[l];
        T* c = (writes & 1) ? out : m.xa[l];
        T* o = (writes & 1) ? m.xa[l] : out;
        if (m.nu >= 2) {
            StencilIO<T> io;
            io.x = b;
            io.y = c;
            io.jw = m.jw[l][0];
            io.jw2 = m.jw[l][1];
            io.done = done;
            int nb = 0;
            launch_stencil<T>(L, OP_JACOBI2, io, &nb, t->tile, s);
            for (int i = 2; i < m.nu; ++i) jacobi(l, b, c, o, m.jw[l][i], false);
        } else {
            launch_mg_scale<T>(c, b, m.jw[l][0], L.N, s, done);
        }
        StencilIO<T> io;
        io.x = c;
        io.b = b;
        io.y = m.rv[l];
        io.done = done;
        int nb = 0;
        launch_stencil<T>(L, OP_RESID, io, &nb, t->tile, s);
        launch_mg_restrict<T>(L.D - 1, m.rv[l], m.bv[l - 1], s, done);
        cur[l] = c;
        oth[l] = o;
    }
    launch_mg_dense<T>(m.nc, t->mg_ainv.p, top == 0 ? b_fine : m.bv[0], top == 0 ? z_fine : m.xa[0], s, done);
    cur[0] = top == 0 ? z_fine : m.xa[0];
    for (int l = 1; l <= top; ++l) {
        const T* b = l == top ? b_fine : m.bv[l];
        launch_mg_prolong_add<T>(m.lv[l].D, cur[l - 1], cur[l], s, done);
        // post-smoothing: the same steps in reverse order (M stays symmetric)
        for (int i = 0; i < m.nu; ++i) jacobi(l, b, cur[l], oth[l], m.jw[l][m.nu - 1 - i], l == top && i == m.nu - 1);
    }
    if (top > 0 && cur[top] != z_fine) throw Error(PRX_ERR_STATE, "multigrid buffer parity");
    if (!done) m.cycles++;  // conditional cycles are counted by the caller once it knows which ran
    return nb_rz;
}

// Preconditioned CG (M = one V-cycle) on level L, the exact operator with
// screening; x is updated in place.  Stops when |r|^2 <= bnorm2 eps^2 (r the
// recursive residual) or after max_iters iterations.  The update
// d = z + beta d is fused into the operator pass as in the plain CG, and x
// is updated lazily in pairs.  The host does not wait for an iteration
// before enqueueing the next: each iteration's state is copied to pinned
// memory behind its last fold, and the host reads iteration k-1's copy only
// after it has enqueued iteration k, so the device never waits on the host.
// The one iteration enqueued past convergence (or the cap) costs only its
// launches: every kernel returns at once when the device flag `done` is set.
template <typename T>
static CgOut run_pcg(prx_tree* t, Mg<T>& m, const LevelView& L, T* x, const T* b, bool x_zero, T* r, T* z, T* dd, T* q,
                     double eps, double bnorm2, uint64_t max_iters) {
    const hipStream_t s = t->stream;
    const int64_t n = L.N;
    CgOut out;
    init_state(t, eps, max_iters, bnorm2);
    int nb = 0;
    if (x_zero) {
        if (r != b) PRX_HIP(hipMemcpyAsync(r, b, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
        launch_dot<T>(r, r, n, t->partial.p, &nb, s);
    } else {
        nb = op_pass<T>(t, L, OP_RESID, x, b, r, nullptr, t->partial.p);
        out.applies++;
    }
    launch_fin(t->st, t->partial.p, nb, FIN_SLOT, 0, s);
    nb = mg_vcycle<T>(t, m, r, z);
    if (nb == 0) launch_dot<T>(r, z, n, t->partial.p, &nb, s);
    launch_fin(t->st, t->partial.p, nb, FIN_PINIT, 0, s);
    launch_fill<T>(dd, T(0), n, s);
    read_state(t);
    CgState* snap = t->state_snapshots();
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (auto& e : ev) PRX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    struct EvGuard {
        hipEvent_t* e;
        ~EvGuard() { for (int i = 0; i < 2; ++i) if (e[i]) (void)hipEventDestroy(e[i]); }
    } evg{ev};
    uint64_t k = 0;
    while (!t->st_host->done) {
        T* dold = dd + (k & 1) * n;
        T* dnew = dd + ((k + 1) & 1) * n;
        StencilIO<T> io;
        io.r = z;
        io.dold = dold;
        io.dnew = dnew;
        io.beta = &t->st->beta;
        io.y = q;
        io.partial = t->partial.p;
        io.done = &t->st->done;
        if (L.scr) {
            T* sb = sbuf_for<T>(t);
            launch_screen_gather<T>(L, nullptr, nullptr, nullptr, sb, s, z, dold, &t->st->beta);
            io.s = sb;
        }
        launch_stencil<T>(L, OP_DIR_APPLY, io, &nb, t->tile, s);
        out.applies++;
        launch_fin(t->st, t->partial.p, nb, FIN_ALPHA, 0, s);
        launch_cg_update<T>(x, r, dold, dnew, q, t->st, n, t->partial.p, &nb, s);
        launch_fin(t->st, t->partial.p, nb, FIN_PRES, (int)((k + 1) & 1), s);
        ++k;
        PRX_HIP(hipMemcpyAsync(snap + (k & 1), t->st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        PRX_HIP(hipEventRecord(ev[k & 1], s));
        if (k >= 2) {  // iteration k-1 finished the solve: iteration k's kernels returned at once
            PRX_HIP(hipEventSynchronize(ev[(k - 1) & 1]));
            if (snap[(k - 1) & 1].done) break;
        }
        nb = mg_vcycle<T>(t, m, r, z, &t->st->done);
        if (nb == 0) launch_dot<T>(r, z, n, t->partial.p, &nb, s);
        launch_fin(t->st, t->partial.p, nb, FIN_PBETA, 0, s);
    }
    read_state(t);
    // the conditional cycles that ran: one after each iteration but the last
    if (t->st_host->it > 1) m.cycles += (uint64_t)(t->st_host->it - 1);
    if (t->st_host->xpend) {
        launch_flush_x<T>(x, dd, dd + n, t->st, n, s);
        launch_clear_xpend(t->st, s);
        t->st_host->xpend = 0;
    }
    const CgState& h = *t->st_host;
    out.iters = (uint64_t)h.it;
    out.status = h.status;
    out.rel = bnorm2 > 0.0 ? std::sqrt(h.slot[1] / bnorm2) : 0.0;
    return out;
}

static int mg_nu(const prx_solve_opts& o) { return o.mg_smooth > 0 ? o.mg_smooth : 2; }
static double mg_omega(const prx_solve_opts& o) { return o.mg_omega < 1.3 ? o.mg_omega : 0.9; }

// FP64 PCG from x = 0 to ||b - A x|| <= eps ||b||, confirmed on the true
// residual (a PCG restart from the current x if the recursive residual drifted).
static void mg_solve_fp64(prx_tree* t, const SolveCfg& c, uint64_t cap, prx_solve_report& rep) {
    const LevelView L = t->fine_view();
    const int64_t n = t->N;
    const hipStream_t s = t->stream;
    t->r64.ensure(n); t->d64.ensure(2 * n); t->q64.ensure(n); t->tmp64.ensure(n);
    Timer ts(s);
    Mg<double> m = mg_build<double>(t, mg_nu(c.o), mg_omega(c.o));
    rep.ms_coarse = ts.ms();
    Timer tm(s);
    launch_fill<double>(t->x.p, 0.0, n, s);
    const double bb = dev_dot<double>(t, t->b.p, t->b.p, n);
    const double thr = (bb * c.eps) * c.eps;
    uint64_t iters = 0;
    int status = PRX_OK;
    bool first = true;
    double true2 = bb;  // ||b - A x||^2 of the last confirmation (x = 0 before the first)
    for (int pass = 0; pass < 4; ++pass) {
        CgOut o = run_pcg<double>(t, m, L, t->x.p, t->b.p, first, t->r64.p, t->tmp64.p, t->d64.p, t->q64.p, c.eps, bb,
                                  cap > iters ? cap - iters : 1);
        first = false;
        iters += o.iters;
        rep.applies += o.applies;
        status = o.status;
        if (o.status != PRX_OK) break;
        int nb = op_pass<double>(t, L, OP_RESID, t->x.p, t->b.p, t->r64.p, nullptr, t->partial.p);
        rep.applies++;
        launch_fin(t->st, t->partial.p, nb, FIN_SLOT, 0, s);
        read_state(t);
        true2 = t->st_host->slot[0];
        if (true2 <= thr) break;
        if (iters >= cap) { status = PRX_ERR_NOT_CONVERGED; break; }
    }
    // four restarts whose recursive residual converged while the true one did not
    if (status == PRX_OK && true2 > thr) status = PRX_ERR_NOT_CONVERGED;
    rep.ms_fine_cg = tm.ms();
    rep.fine_iters = iters;
    rep.outer_iters = 1;
    rep.coarse_iters = m.cycles;
    rep.fine_rel = bb > 0.0 ? std::sqrt(true2 / bb) : 0.0;
    // true2 is ||b - A x||^2 of the returned x unless PCG broke down before its confirmation
    if (status != PRX_ERR_BREAKDOWN) rep.true_rel = rep.fine_rel;
    rep.status = status;
}

// Mixed precision: FP64 refinement x += e around FP32 PCG solves of
// (A/h) e = (b - A x)/h, each to a residual reduction mg_inner_tol (or less
// when less reaches the target); the true residual b - A x is formed in FP64
// (in the same pass that folds e into x) and decides convergence.
static void mg_solve_mixed(prx_tree* t, const SolveCfg& c, uint64_t cap, prx_solve_report& rep) {
    const LevelView L = t->fine_view();
    const int64_t n = t->N;
    const hipStream_t s = t->stream;
    t->e32.ensure(n); t->ri32.ensure(n); t->d32.ensure(2 * n); t->q32.ensure(n); t->r32.ensure(n);
    const double ih = 1.0 / t->h;
    const double tol = c.o.mg_inner_tol > 0.0 && c.o.mg_inner_tol < 1.0 ? c.o.mg_inner_tol : 1e-4;
    Timer ts(s);
    Mg<float> m = mg_build<float>(t, mg_nu(c.o), mg_omega(c.o));
    rep.ms_coarse = ts.ms();
    Timer tm(s);
    launch_fill<double>(t->x.p, 0.0, n, s);
    const double bb = dev_dot<double>(t, t->b.p, t->b.p, n);
    const double thr = (bb * c.eps) * c.eps;
    // x = 0: the first residual is b itself (scaled into the FP32 system)
    launch_cast_scale_d2f(t->ri32.p, t->b.p, ih, n, s);
    double rr = bb;
    int nb = 0;
    uint64_t iters = 0, outer = 0;
    int status = PRX_OK;
    while (rr > thr) {
        if (outer >= c.o.max_outer || iters >= cap) { status = PRX_ERR_NOT_CONVERGED; break; }
        launch_fill<float>(t->e32.p, 0.0f, n, s);
        // reduce by mg_inner_tol, or only by what the target still needs (with margin 2)
        const double need = 0.5 * std::sqrt(thr / rr);
        CgOut o = run_pcg<float>(t, m, L, t->e32.p, t->ri32.p, true, t->ri32.p, t->r32.p, t->d32.p, t->q32.p,
                                 std::max(tol, need), rr * ih * ih, cap - iters);
        iters += o.iters;
        rep.applies += o.applies;
        if (o.status == PRX_ERR_BREAKDOWN) { status = o.status; break; }
        nb = op_resid_update(t, L, t->e32.p, t->b.p, t->ri32.p, ih);
        rep.applies++;
        launch_fin(t->st, t->partial.p, nb, FIN_SLOT, 0, s);
        read_state(t);
        const double rn = t->st_host->slot[0];
        ++outer;
        if (!(rn < rr)) { rr = rn; status = PRX_ERR_NOT_CONVERGED; break; }
        rr = rn;
    }
    rep.ms_fine_cg = tm.ms();
    rep.fine_iters = iters;
    rep.outer_iters = outer;
    rep.coarse_iters = m.cycles;
    rep.fine_rel = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    // rr is ||b - A x||^2 of the returned x, formed in FP64 by the last update pass
    if (status != PRX_ERR_BREAKDOWN) rep.true_rel = rep.fine_rel;
    rep.status = status;
}

static void solve_mg(prx_tree* t, const SolveCfg& c, prx_solve_report& rep) {
    Timer total(t->stream);
    rep.true_rel = -1.0;  // set by the solver when its last pass formed ||b - A x|| of the returned x
    if (c.o.precision == PRX_PREC_MIXED) mg_solve_mixed(t, c, c.cg_iters, rep);
    else mg_solve_fp64(t, c, c.cg_iters, rep);
    rep.ms_total = total.ms();
    if (rep.true_rel < 0.0) rep.true_rel = true_rel(t);
}

static void solve_cascadic(prx_tree* t, const SolveCfg& c, prx_solve_report& rep) {
    std::memset(&rep, 0, sizeof rep);
    t->ensure_state();
    t->precond = c.o.precond;
    t->cg_form = c.o.cg_form;
    if (c.o.precond == PRX_PRECOND_MG) {
        solve_mg(t, c, rep);
        return;
    }
    const hipStream_t s = t->stream;
    Timer total(s);
    if (c.base >= t->D) {
        launch_fill<double>(t->x.p, 0.0, t->N, s);
    } else {
        Timer tm(s);
        const double ceps = c.o.residual_mode == PRX_RESID_TRUE && c.o.coarse_eps > 0.0 ? c.o.coarse_eps : c.eps;
        coarse_solve(t, c.base, c.gs, c.cg_iters, ceps, c.o.residual_mode, rep);
        launch_prolongate(t->D, c.base, t->cx.p, t->x.p, s);
        rep.ms_coarse = tm.ms();
    }
    const uint64_t cap = c.base >= t->D ? c.cg_iters : c.cg_iters / 4;
    if (c.o.precision == PRX_PREC_MIXED) fine_solve_mixed(t, c, cap, rep);
    else fine_solve_fp64(t, c, cap, rep);
    rep.ms_total = total.ms();
    rep.true_rel = true_rel(t);
}

// ------------------------------------------------------------------ splat and assembly
void upload_points(prx_tree* t, const prx_point* pts, size_t n) {
    const auto h0 = std::chrono::steady_clock::now();
    const int64_t P = (int64_t)n;
    if (P >= ((int64_t)1 << 29)) throw Error(PRX_ERR_INVALID, "at most 2^29 points per tree");
    std::vector<double> soa(6 * (size_t)std::max<int64_t>(P, 1));
    for (int64_t i = 0; i < P; ++i) {
        soa[i] = pts[i].px; soa[P + i] = pts[i].py; soa[2 * P + i] = pts[i].pz;
        soa[3 * P + i] = pts[i].nx; soa[4 * P + i] = pts[i].ny; soa[5 * P + i] = pts[i].nz;
    }
    t->P = P;
    t->pos.ensure(3 * std::max<int64_t>(P, 1));
    t->nrm.ensure(3 * std::max<int64_t>(P, 1));
    if (P > 0) {
        PRX_HIP(hipMemcpyAsync(t->pos.p, soa.data(), sizeof(double) * 3 * P, hipMemcpyHostToDevice, t->stream));
        PRX_HIP(hipMemcpyAsync(t->nrm.p, soa.data() + 3 * P, sizeof(double) * 3 * P, hipMemcpyHostToDevice, t->stream));
    }
    PRX_HIP(hipStreamSynchronize(t->stream));
    t->setup_ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
}

// positions and normals already on the device, each [n][3] (x, y, z per
// point): transposed into the tree's SoA arrays on the device
__global__ void k_aos3_to_soa(const double* __restrict__ a, int64_t n, double* __restrict__ soa) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        soa[i] = a[3 * i];
        soa[n + i] = a[3 * i + 1];
        soa[2 * n + i] = a[3 * i + 2];
    }
}
void upload_points_device(prx_tree* t, const double* pos, const double* nrm, size_t n) {
    const auto h0 = std::chrono::steady_clock::now();
    const int64_t P = (int64_t)n;
    if (P >= ((int64_t)1 << 29)) throw Error(PRX_ERR_INVALID, "at most 2^29 points per tree");
    t->P = P;
    t->pos.ensure(3 * std::max<int64_t>(P, 1));
    t->nrm.ensure(3 * std::max<int64_t>(P, 1));
    if (P > 0) {
        const unsigned nb = (unsigned)std::min<int64_t>((P + 255) / 256, 65536);
        hipLaunchKernelGGL(k_aos3_to_soa, dim3(nb), dim3(256), 0, t->stream, pos, P, t->pos.p);
        hipLaunchKernelGGL(k_aos3_to_soa, dim3(nb), dim3(256), 0, t->stream, nrm, P, t->nrm.p);
        PRX_LAUNCH_CHECK();
    }
    PRX_HIP(hipStreamSynchronize(t->stream));
    t->setup_ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
}

void splat(prx_tree* t) {
    const hipStream_t s = t->stream;
    const int64_t P = t->P, M8 = 8 * P;
    free_list(t->floor_list);
    free_list(t->trunc_list);
    free_list(t->eval_list);
    free_cell_lists(t->cells);
    t->cells_ok = false;
    t->perm.ensure(std::max<int64_t>(P, 1));
    t->cell.ensure(3 * std::max<int64_t>(P, 1));
    t->t64.ensure(3 * std::max<int64_t>(P, 1));
    t->t32.ensure(3 * std::max<int64_t>(P, 1));
    t->pts.P = P;
    t->pts.pos = t->pos.p;
    t->pts.nrm = t->nrm.p;
    t->pts.perm = t->perm.p;
    t->pts.cell = t->cell.p;
    t->pts.t64 = t->t64.p;
    t->pts.t32 = t->t32.p;
    t->nf.ensure(3 * t->N);
    t->w.ensure(t->N);
    for (int i = 1; i < 5; ++i) t->setup_ms[i] = 0.0;
    if (P > 0) {
        Timer tsort(s);
        DArr<uint32_t> kf, kt, val, kf2, kt2, val2, val3, ck, cv, ck2, inv;
        kf.ensure(M8); kt.ensure(M8); val.ensure(M8); kf2.ensure(M8); kt2.ensure(M8); val2.ensure(M8); val3.ensure(M8);
        ck.ensure(P); cv.ensure(P); ck2.ensure(P); inv.ensure(P);
        launch_point_keys(t->D, t->pos.p, P, kf.p, kt.p, val.p, ck.p, cv.p, s);
        const int kbits = 3 * (int)t->D + 1;
        size_t need = 0, need2 = 0, need3 = 0;
        PRX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kf.p, kf2.p, val.p, val2.p, (int)M8, 0, kbits, s));
        PRX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need2, ck.p, ck2.p, cv.p, t->perm.p, (int)P, 0, kbits, s));
        PRX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need3, val.p, val2.p, (int)M8, s));
        size_t sb = std::max(need, std::max(need2, need3));
        DArr<uint8_t> scratch;
        scratch.ensure((int64_t)sb);
        // spatial order of the points (screening gathers read neighbouring nodes)
        PRX_HIP(hipcub::DeviceRadixSort::SortPairs(scratch.p, sb, ck.p, ck2.p, cv.p, t->perm.p, (int)P, 0, kbits, s));
        launch_invert_perm(t->perm.p, P, inv.p, s);
        launch_point_sorted_data(t->D, t->pos.p, P, t->perm.p, t->cell.p, t->t64.p, t->t32.p, s);
        t->setup_ms[1] = tsort.ms();
        // floor list (splat), trunc list (screening); radix sort is stable, so
        // entries of one node stay in input point order
        Timer tfloor(s);
        PRX_HIP(hipcub::DeviceRadixSort::SortPairs(scratch.p, sb, kf.p, kf2.p, val.p, val2.p, (int)M8, 0, kbits, s));
        build_node_list(t->floor_list, t->D, kf2.p, val2.p, M8, t->pts, inv.p, scratch.p, sb, s);
        t->setup_ms[2] = tfloor.ms();
        // the screening lists: the reference builds the per-point blocks in
        // assemble_system (src/fem_tree.rs:199-227)
        Timer ttrunc(s);
        PRX_HIP(hipcub::DeviceRadixSort::SortPairs(scratch.p, sb, kt.p, kt2.p, val.p, val3.p, (int)M8, 0, kbits, s));
        build_node_list(t->trunc_list, t->D, kt2.p, val3.p, M8, t->pts, inv.p, scratch.p, sb, s);
        // and the occupied cells with each touched node's (cell, corner) pairs (face-form CG)
        build_cell_lists(t->cells, t->D, t->pts, t->trunc_list, s);
        t->cells_ok = true;
        t->setup_ms[3] = ttrunc.ms();
    }
    Timer tsplat(s);
    launch_splat(t->D, t->floor_list, t->pts, t->nf.p, t->w.p, s);
    t->setup_ms[4] = tsplat.ms();
    t->splatted = true;
    t->qt_ok = false;
    t->assembled = false;
}

void assemble(prx_tree* t, double pw) {
    if (!t->splatted) throw Error(PRX_ERR_STATE, "assemble before splat_normal_field");
    const hipStream_t s = t->stream;
    t->set_coefficients(pw);
    t->b.ensure(t->N);
    t->x.ensure(t->N);
    launch_rhs_laplacian(t->D, t->nf.p, t->b.p, &t->floor_list, s);
    if (pw > 0.0) launch_rhs_screening(t->D, t->trunc_list, t->pts, t->wh3, t->b.p, s);
    launch_fill<double>(t->x.p, 0.0, t->N, s);
    t->ensure_state();
    PRX_HIP(hipStreamSynchronize(s));
    t->assembled = true;
}

prx_tree* create_tree(uint32_t depth, int device, hipStream_t shared) {
    if (depth > PRX_MAX_DEPTH) throw Error(PRX_ERR_INVALID, "depth must be <= 10");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw Error(PRX_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) throw Error(PRX_ERR_INVALID, "device ordinal out of range");
    std::unique_ptr<prx_tree> t(new prx_tree());
    t->device = device;
    PRX_HIP(hipSetDevice(device));
    if (shared) {
        t->stream = shared;
        t->own_stream = false;
    } else {
        PRX_HIP(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    }
    t->D = depth;
    t->res = 1u << depth;
    t->N = (int64_t)1 << (3 * depth);
    t->h = 1.0 / (double)t->res;
    if (const char* e = std::getenv("PRX_TILE")) t->tile = std::atoi(e) == 8 ? 8 : 16;
    return t.release();
}

bool finite_points(const prx_point* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i].px) || !std::isfinite(p[i].py) || !std::isfinite(p[i].pz) || !std::isfinite(p[i].nx) ||
            !std::isfinite(p[i].ny) || !std::isfinite(p[i].nz))
            return false;
    return true;
}

// Per-node list of the evaluate corners of the points (built on first use).
static void build_eval_list(prx_tree* t) {
    if (t->eval_list.T > 0 || t->P == 0) return;
    const hipStream_t s = t->stream;
    const int64_t P = t->P, M8 = 8 * P;
    DArr<uint32_t> k, v, k2, v2, inv;
    k.ensure(M8); v.ensure(M8); k2.ensure(M8); v2.ensure(M8); inv.ensure(P);
    launch_point_keys_eval(t->D, t->pos.p, P, k.p, v.p, s);
    const int kbits = 3 * (int)t->D + 1;
    size_t need = 0, need2 = 0;
    PRX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, k.p, k2.p, v.p, v2.p, (int)M8, 0, kbits, s));
    PRX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need2, k.p, k2.p, (int)M8, s));
    size_t sb = std::max(need, need2);
    DArr<uint8_t> scratch;
    scratch.ensure((int64_t)sb);
    PRX_HIP(hipcub::DeviceRadixSort::SortPairs(scratch.p, sb, k.p, k2.p, v.p, v2.p, (int)M8, 0, kbits, s));
    launch_invert_perm(t->perm.p, P, inv.p, s);
    build_node_list(t->eval_list, t->D, k2.p, v2.p, M8, t->pts, inv.p, scratch.p, sb, s);
    PRX_HIP(hipStreamSynchronize(s));
}

// vjp of the field at the points (see include/prx.h): one adjoint solve with
// the forward operator (A is symmetric), then point-local terms.
static void vjp_points(prx_tree* t, const double* gF_host, double* gpos_host, double* gnrm_host, int precision, double eps,
                       prx_vjp_report* rep, bool on_device = false) {
    if (!t->assembled) throw Error(PRX_ERR_STATE, "vjp before assemble/solve");
    const hipStream_t s = t->stream;
    const int64_t N = t->N, P = t->P;
    Timer tm(s);
    t->ensure_state();
    build_eval_list(t);
    DArr<double> gF;
    gF.ensure(std::max<int64_t>(P, 1));
    if (P > 0)
        PRX_HIP(hipMemcpyAsync(gF.p, gF_host, sizeof(double) * P, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    t->adj_g.ensure(N);
    t->adj_lam.ensure(N);
    launch_fill<double>(t->adj_g.p, 0.0, N, s);
    launch_adj_load(t->eval_list, gF.p, t->adj_g.p, s);
    // solve A lam = g with the fine solver: swap the RHS and solution slots
    t->b.swap(t->adj_g);
    t->x.swap(t->adj_lam);
    prx_solve_report r;
    std::memset(&r, 0, sizeof r);
    try {
        launch_fill<double>(t->x.p, 0.0, N, s);
        SolveCfg c;
        c.base = t->D;
        c.gs = 20;
        c.cg_iters = 4u * 200000u;
        c.eps = eps;
        prx_solve_opts_default(&c.o);
        c.o.precision = precision;
        c.o.residual_mode = PRX_RESID_TRUE;
        c.o.refresh_period = 0;
        c.o.inner_tol = 0.05;
        c.o.max_outer = 200;
        if (t->precond == PRX_PRECOND_MG) {
            c.o.precond = PRX_PRECOND_MG;
            if (precision == PRX_PREC_MIXED) mg_solve_mixed(t, c, 10000, r);
            else mg_solve_fp64(t, c, 10000, r);
        } else if (precision == PRX_PREC_MIXED) {
            fine_solve_mixed(t, c, 200000, r);
        } else {
            fine_solve_fp64(t, c, 200000, r);
        }
    } catch (...) {
        t->x.swap(t->adj_lam);
        t->b.swap(t->adj_g);
        throw;
    }
    t->x.swap(t->adj_lam);
    t->b.swap(t->adj_g);
    // nf-adjoint at the splat nodes, then the per-point terms
    t->adj_gnf.ensure(3 * std::max<int64_t>(t->floor_list.T, 1));
    launch_adj_gnf(t->D, t->floor_list, t->adj_lam.p, t->adj_gnf.p, s);
    t->adj_out.ensure(6 * std::max<int64_t>(P, 1));
    launch_adj_points(t->D, t->pts, t->floor_list, gF.p, t->x.p, t->adj_lam.p, t->nf.p, t->w.p, t->adj_gnf.p,
                      t->pw > 0.0 ? t->wh3 : 0.0, t->adj_out.p, t->adj_out.p + 3 * P, s);
    if (P > 0) {
        const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        PRX_HIP(hipMemcpyAsync(gpos_host, t->adj_out.p, sizeof(double) * 3 * P, k, s));
        PRX_HIP(hipMemcpyAsync(gnrm_host, t->adj_out.p + 3 * P, sizeof(double) * 3 * P, k, s));
    }
    PRX_HIP(hipStreamSynchronize(s));
    if (rep) {
        rep->iters = r.fine_iters;
        rep->rel = r.fine_rel;
        rep->ms = tm.ms();
        rep->status = r.status;
    }
}

int time_operator(prx_tree* t, int precision, int reps, double* ms_gather, double* ms_stencil, int op) {
    try {
        if (!t->assembled) return PRX_ERR_STATE;
        set_device(t);
        t->ensure_state();
        const LevelView L = t->fine_view();
        const int64_t n = t->N;
        const hipStream_t s = t->stream;
        hipEvent_t e0, e1, e2;
        PRX_HIP(hipEventCreate(&e0));
        PRX_HIP(hipEventCreate(&e1));
        PRX_HIP(hipEventCreate(&e2));
        double g = 0.0, k = 0.0;
        // OP_DIR_APPLY: r = xin, d_old = dold, d_new = dnew (beta from the CG
        // state); OP_CG also q = qin and r' to rout (alpha, beta, rpend = 1)
        auto run = [&](auto* xin, auto* yout, auto* sb, auto* dold, auto* dnew, auto* qin, auto* rout) {
            using T = std::remove_pointer_t<decltype(xin)>;
            StencilIO<T> io;
            io.x = xin;
            io.y = yout;
            io.partial = t->partial.p;
            io.s = sb;
            io.r = xin;
            io.dold = dold;
            io.dnew = dnew;
            io.beta = &t->st->beta;
            io.q = qin;
            io.rout = rout;
            io.cst = t->st;
            int nb = 0;
            for (int r = -1; r < reps; ++r) {  // r = -1 warms up
                PRX_HIP(hipEventRecord(e0, s));
                if (op == OP_CG_FACES) {
                    if (L.scr) {
                        io.qt = qt_for<T>(t, L);
                        T* scs = io.qt + L.N;
                        io.sc = scs;
                        launch_cell_screening<T>(L, t->cells, xin, dold, io.qt, t->st, nullptr, scs + std::max<int64_t>(L.scr->T, 1),
                                                 scs, s);
                    }
                    PRX_HIP(hipEventRecord(e1, s));
                    launch_cg_faces<T>(L, io, &nb, s);
                    PRX_HIP(hipEventRecord(e2, s));
                    PRX_HIP(hipEventSynchronize(e2));
                    float a = 0.f, b = 0.f;
                    PRX_HIP(hipEventElapsedTime(&a, e0, e1));
                    PRX_HIP(hipEventElapsedTime(&b, e1, e2));
                    if (r >= 0) { g += a; k += b; }
                    continue;
                }
                if (L.scr) {
                    if (op == OP_DIR_APPLY) launch_screen_gather<T>(L, nullptr, nullptr, nullptr, sb, s, xin, dold, &t->st->beta);
                    else if (op == OP_CG) launch_screen_gather<T>(L, nullptr, nullptr, nullptr, sb, s, xin, dold, nullptr, qin, t->st);
                    else launch_screen_gather<T>(L, xin, nullptr, nullptr, sb, s);
                }
                PRX_HIP(hipEventRecord(e1, s));
                launch_stencil<T>(L, op, io, &nb, t->tile, s);
                PRX_HIP(hipEventRecord(e2, s));
                PRX_HIP(hipEventSynchronize(e2));
                float a = 0.f, b = 0.f;
                PRX_HIP(hipEventElapsedTime(&a, e0, e1));
                PRX_HIP(hipEventElapsedTime(&b, e1, e2));
                if (r >= 0) { g += a; k += b; }
            }
        };
        if (op != OP_APPLY && op != OP_DIR_APPLY && op != OP_CG && op != OP_CG_FACES) return PRX_ERR_INVALID;
        if (op == OP_CG && L.D < 3) return PRX_ERR_INVALID;
        if (op == OP_CG_FACES && L.D < 4) return PRX_ERR_INVALID;
        if (op == OP_CG_FACES && L.scr) {  // defined screening inputs for the timed pass
            if (precision == PRX_PREC_MIXED) qt_for<float>(t, L);
            else qt_for<double>(t, L);
        }
        if (op == OP_DIR_APPLY || op == OP_CG || op == OP_CG_FACES) {  // beta = 0.5, alpha = 0.25, r update pending, not done
            CgState h{};
            h.beta = 0.5;
            h.alpha = 0.25;
            h.rpend = 1;
            *t->st_host = h;
            PRX_HIP(hipMemcpyAsync(t->st, t->st_host, sizeof(CgState), hipMemcpyHostToDevice, s));
        }
        if (precision == PRX_PREC_MIXED) {
            t->d32.ensure(2 * n);
            t->q32.ensure(2 * n);
            t->ri32.ensure(2 * n);
            launch_fill<float>(t->d32.p, 1.0f, 2 * n, s);
            launch_fill<float>(t->ri32.p, 0.5f, n, s);
            launch_fill<float>(t->q32.p, 0.125f, n, s);
            float* sb = L.scr ? sbuf_for<float>(t) : (float*)nullptr;
            if (op == OP_DIR_APPLY) run(t->ri32.p, t->q32.p, sb, t->d32.p, t->d32.p + n, (float*)nullptr, (float*)nullptr);
            else if (op == OP_CG || op == OP_CG_FACES) run(t->ri32.p, t->q32.p + n, sb, t->d32.p, t->d32.p + n, t->q32.p, t->ri32.p + n);
            else run(t->d32.p, t->q32.p, sb, (float*)nullptr, (float*)nullptr, (float*)nullptr, (float*)nullptr);
        } else {
            t->d64.ensure(2 * n);
            t->q64.ensure(2 * n);
            t->r64.ensure(2 * n);
            launch_fill<double>(t->d64.p, 1.0, 2 * n, s);
            launch_fill<double>(t->r64.p, 0.5, n, s);
            launch_fill<double>(t->q64.p, 0.125, n, s);
            double* sb = L.scr ? sbuf_for<double>(t) : (double*)nullptr;
            if (op == OP_DIR_APPLY) run(t->r64.p, t->q64.p, sb, t->d64.p, t->d64.p + n, (double*)nullptr, (double*)nullptr);
            else if (op == OP_CG || op == OP_CG_FACES) run(t->r64.p, t->q64.p + n, sb, t->d64.p, t->d64.p + n, t->q64.p, t->r64.p + n);
            else run(t->d64.p, t->q64.p, sb, (double*)nullptr, (double*)nullptr, (double*)nullptr, (double*)nullptr);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        if (ms_gather) *ms_gather = g / reps;
        if (ms_stencil) *ms_stencil = k / reps;
        return PRX_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    }
}

}  // namespace prx

// ====================================================================== C ABI
extern "C" {

const char* prx_last_error(void) { return prx::g_err.c_str(); }

int prx_device_count(int* out) {
    PRX_API_BEGIN
    PRX_REQUIRE(out, PRX_ERR_INVALID, "null pointer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *out = n;
    PRX_API_END
}

void prx_params_default(prx_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->depth = 8;
    p->samples_per_node = 1.5;
    p->scale = 1.1;
    p->cg_accuracy = 1e-3;
    p->cg_iters = 500;
    p->point_weight = 0.0;
    p->verbose = 0;
    p->precision = PRX_PREC_FP64;
    p->residual_mode = PRX_RESID_REFERENCE;
    p->device = 0;
    p->gs_iters = 20;
    p->base_depth = -1;
    p->refresh_period = 10;
    p->inner_tol = 0.05;
    p->max_outer = 64;
}

void prx_solve_opts_default(prx_solve_opts* o) {
    if (!o) return;
    o->precision = PRX_PREC_FP64;
    o->residual_mode = PRX_RESID_REFERENCE;
    o->refresh_period = 10;
    o->gs_iters = 20;
    o->inner_tol = 0.05;
    o->max_outer = 64;
    o->precond = PRX_PRECOND_NONE;
    o->mg_smooth = 2;
    o->mg_omega = 0.0;
    o->mg_inner_tol = 1e-4;
    o->coarse_eps = 0.0;
    o->cg_form = PRX_CG_FUSED;
    o->batch_launch = PRX_BATCH_LAUNCH_AUTO;
}

int prx_tree_create(uint32_t depth, int device, prx_tree** out) {
    PRX_API_BEGIN
    PRX_REQUIRE(out, PRX_ERR_INVALID, "null pointer");
    *out = nullptr;
    *out = prx::create_tree(depth, device);
    PRX_API_END
}

void prx_tree_destroy(prx_tree* t) {
    if (!t) return;
    (void)hipSetDevice(t->device);
    delete t;
}

int prx_tree_release_workspace(prx_tree* t) {
    PRX_API_BEGIN
    PRX_REQUIRE(t, PRX_ERR_INVALID, "null tree");
    set_device(t);
    PRX_HIP(hipStreamSynchronize(t->stream));
    t->release_workspace();
    PRX_API_END
}

int64_t prx_tree_node_count(const prx_tree* t) { return t ? t->N : -1; }

int prx_tree_export_nodes(prx_tree* t, uint32_t* i2o, int64_t* g2i) {
    PRX_API_BEGIN
    PRX_REQUIRE(t, PRX_ERR_INVALID, "null tree");
    set_device(t);
    if (i2o) {
        DArr<uint32_t> d;
        d.ensure(3 * t->N);
        launch_export_offsets(t->D, d.p, t->stream);
        PRX_HIP(hipMemcpyAsync(i2o, d.p, sizeof(uint32_t) * 3 * t->N, hipMemcpyDeviceToHost, t->stream));
        PRX_HIP(hipStreamSynchronize(t->stream));
    }
    if (g2i) {
        DArr<int64_t> d;
        d.ensure(t->N);
        launch_export_g2i(t->D, d.p, t->stream);
        PRX_HIP(hipMemcpyAsync(g2i, d.p, sizeof(int64_t) * t->N, hipMemcpyDeviceToHost, t->stream));
        PRX_HIP(hipStreamSynchronize(t->stream));
    }
    PRX_API_END
}

int prx_splat_normal_field(prx_tree* t, const prx_point* pts, size_t n) {
    PRX_API_BEGIN
    PRX_REQUIRE(t, PRX_ERR_INVALID, "null tree");
    PRX_REQUIRE(pts || n == 0, PRX_ERR_INVALID, "null points");
    set_device(t);
    prx::upload_points(t, pts, n);
    prx::splat(t);
    PRX_API_END
}

int prx_screening_sizes(prx_tree* t, int64_t* out) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && out, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->splatted, PRX_ERR_STATE, "splat_normal_field not called");
    out[0] = t->trunc_list.T;
    out[1] = t->trunc_list.M;
    out[2] = t->cells_ok ? t->cells.C : 0;
    PRX_API_END
}

int prx_pass_timing(prx_tree* t, int on, float* ms, size_t cap, size_t* n) {
    PRX_API_BEGIN
    PRX_REQUIRE(t, PRX_ERR_INVALID, "null tree");
    set_device(t);
    auto& P = t->ptime;
    if (on) {
        P.on = true;
        P.used = 0;
        P.ms.clear();
    } else {
        PRX_REQUIRE(n, PRX_ERR_INVALID, "null count");
        prx::ptime_read(t);
        P.on = false;
        *n = P.ms.size();
        if (ms) for (size_t i = 0; i < P.ms.size() && i < cap; ++i) ms[i] = P.ms[i];
        P.ms.clear();
    }
    PRX_API_END
}

int prx_splat_times(prx_tree* t, double* ms) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && ms, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->splatted, PRX_ERR_STATE, "splat_normal_field not called");
    for (int i = 0; i < 5; ++i) ms[i] = t->setup_ms[i];
    PRX_API_END
}

int prx_splat_normal_field_device(prx_tree* t, const double* pos_device, const double* nrm_device, size_t n) {
    PRX_API_BEGIN
    PRX_REQUIRE(t, PRX_ERR_INVALID, "null tree");
    PRX_REQUIRE((pos_device && nrm_device) || n == 0, PRX_ERR_INVALID, "null points");
    set_device(t);
    prx::upload_points_device(t, pos_device, nrm_device, n);
    prx::splat(t);
    PRX_API_END
}

int prx_export_normal_field(prx_tree* t, double* nf, double* w) {
    PRX_API_BEGIN
    PRX_REQUIRE(t, PRX_ERR_INVALID, "null tree");
    PRX_REQUIRE(t->splatted, PRX_ERR_STATE, "splat_normal_field not called");
    set_device(t);
    if (nf) {
        std::vector<double> soa(3 * (size_t)t->N);
        PRX_HIP(hipMemcpyAsync(soa.data(), t->nf.p, sizeof(double) * 3 * t->N, hipMemcpyDeviceToHost, t->stream));
        PRX_HIP(hipStreamSynchronize(t->stream));
        for (int64_t i = 0; i < t->N; ++i) {
            nf[3 * i] = soa[i];
            nf[3 * i + 1] = soa[t->N + i];
            nf[3 * i + 2] = soa[2 * t->N + i];
        }
    }
    if (w) {
        PRX_HIP(hipMemcpyAsync(w, t->w.p, sizeof(double) * t->N, hipMemcpyDeviceToHost, t->stream));
        PRX_HIP(hipStreamSynchronize(t->stream));
    }
    PRX_API_END
}

int prx_assemble(prx_tree* t, double point_weight) {
    PRX_API_BEGIN
    PRX_REQUIRE(t, PRX_ERR_INVALID, "null tree");
    set_device(t);
    prx::assemble(t, point_weight);
    PRX_API_END
}

int prx_export_rhs(prx_tree* t, double* rhs) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && rhs, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    set_device(t);
    PRX_HIP(hipMemcpyAsync(rhs, t->b.p, sizeof(double) * t->N, hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    PRX_API_END
}

int prx_csr_nnz(prx_tree* t, int64_t* nnz) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && nnz, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    set_device(t);
    DArr<int64_t> cnt, sum;
    cnt.ensure(t->N);
    sum.ensure(1);
    const NodeList* L = t->pw > 0.0 ? &t->trunc_list : nullptr;
    launch_csr_count(t->D, L, &t->pts, t->pw > 0.0 ? 1 : 0, cnt.p, t->stream);
    size_t need = 0;
    PRX_HIP(hipcub::DeviceReduce::Sum(nullptr, need, cnt.p, sum.p, (int)t->N, t->stream));
    DArr<uint8_t> tmp;
    tmp.ensure((int64_t)need);
    PRX_HIP(hipcub::DeviceReduce::Sum(tmp.p, need, cnt.p, sum.p, (int)t->N, t->stream));
    PRX_HIP(hipMemcpyAsync(nnz, sum.p, sizeof(int64_t), hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    PRX_API_END
}

int prx_export_csr(prx_tree* t, int64_t* off, int64_t* col, double* val) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && off && col && val, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    set_device(t);
    const NodeList* L = t->pw > 0.0 ? &t->trunc_list : nullptr;
    const int diag = t->pw > 0.0 ? 1 : 0;
    DArr<int64_t> cnt;
    cnt.ensure(t->N + 1);
    launch_csr_count(t->D, L, &t->pts, diag, cnt.p, t->stream);
    std::vector<int64_t> h(t->N);
    PRX_HIP(hipMemcpyAsync(h.data(), cnt.p, sizeof(int64_t) * t->N, hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    off[0] = 0;
    for (int64_t i = 0; i < t->N; ++i) off[i + 1] = off[i] + h[i];
    const int64_t nnz = off[t->N];
    DArr<int64_t> doff, dcol;
    DArr<double> dval;
    doff.ensure(t->N + 1);
    dcol.ensure(nnz);
    dval.ensure(nnz);
    PRX_HIP(hipMemcpyAsync(doff.p, off, sizeof(int64_t) * (t->N + 1), hipMemcpyHostToDevice, t->stream));
    launch_csr_fill(t->D, L, &t->pts, diag, t->K0, t->Ke, t->Kc, t->wg, t->wh3, doff.p, dcol.p, dval.p, t->stream);
    PRX_HIP(hipMemcpyAsync(col, dcol.p, sizeof(int64_t) * nnz, hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipMemcpyAsync(val, dval.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    PRX_API_END
}

int prx_apply(prx_tree* t, int precision, const double* x, double* y) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && x && y, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    set_device(t);
    const LevelView L = t->fine_view();
    const int64_t n = t->N;
    const hipStream_t s = t->stream;
    t->r64.ensure(n);
    t->q64.ensure(n);
    PRX_HIP(hipMemcpyAsync(t->r64.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (precision == PRX_PREC_MIXED) {
        t->d32.ensure(n);
        t->q32.ensure(n);
        launch_cast_scale_d2f(t->d32.p, t->r64.p, 1.0, n, s);
        prx::op_pass<float>(t, L, OP_APPLY, t->d32.p, nullptr, t->q32.p, nullptr, t->partial.p);
        std::vector<float> hy(n);
        PRX_HIP(hipMemcpyAsync(hy.data(), t->q32.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; ++i) y[i] = (double)hy[i] * t->h;
    } else {
        prx::op_pass<double>(t, L, OP_APPLY, t->r64.p, nullptr, t->q64.p, nullptr, t->partial.p);
        PRX_HIP(hipMemcpyAsync(y, t->q64.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
    }
    PRX_API_END
}

int prx_mg_vcycle(prx_tree* t, int precision, int smooth, double omega, const double* in, double* out) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && in && out, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    set_device(t);
    t->ensure_state();
    const int64_t n = t->N;
    const hipStream_t s = t->stream;
    prx_solve_opts o;
    prx_solve_opts_default(&o);
    o.mg_smooth = smooth;
    o.mg_omega = omega;
    t->r64.ensure(n);
    t->q64.ensure(n);
    PRX_HIP(hipMemcpyAsync(t->r64.p, in, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (precision == PRX_PREC_MIXED) {
        // the FP32 system is A/h: z = (A/h)^-1 (b/h)
        t->d32.ensure(n);
        t->q32.ensure(n);
        prx::Mg<float> m = prx::mg_build<float>(t, prx::mg_nu(o), prx::mg_omega(o));
        launch_cast_scale_d2f(t->d32.p, t->r64.p, 1.0 / t->h, n, s);
        prx::mg_vcycle<float>(t, m, t->d32.p, t->q32.p);
        std::vector<float> hz(n);
        PRX_HIP(hipMemcpyAsync(hz.data(), t->q32.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; ++i) out[i] = (double)hz[i];
    } else {
        prx::Mg<double> m = prx::mg_build<double>(t, prx::mg_nu(o), prx::mg_omega(o));
        prx::mg_vcycle<double>(t, m, t->r64.p, t->q64.p);
        PRX_HIP(hipMemcpyAsync(out, t->q64.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
    }
    PRX_API_END
}

int prx_coarse_system(prx_tree* t, uint32_t cdepth, double* rhs_rowmajor, double* coef) {
    PRX_API_BEGIN
    PRX_REQUIRE(t, PRX_ERR_INVALID, "null tree");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    PRX_REQUIRE(cdepth <= t->D, PRX_ERR_INVALID, "coarse depth above the tree depth");
    set_device(t);
    const int64_t nc = (int64_t)1 << (3 * cdepth);
    const double hc = 1.0 / (double)(1u << cdepth);
    const double epw = t->pw * (double)((uint64_t)1 << (2 * (t->D - cdepth)));
    const double wh3c = epw > 0.0 ? epw * (hc * hc * hc) / 1.5 : 0.0;
    if (coef) {
        coef[0] = prx::stiffness_entry(hc, 0, 0, 0);
        coef[1] = prx::stiffness_entry(hc, 1, 1, 0);
        coef[2] = prx::stiffness_entry(hc, 1, 1, 1);
        coef[3] = epw > 0.0 ? wh3c * 1.005 : 0.0;
    }
    if (rhs_rowmajor) {
        DArr<double> m, r;
        m.ensure(nc);
        r.ensure(nc);
        launch_coarse_rhs(t->D, cdepth, t->nf.p, epw > 0.0 ? wh3c * 0.5 : 0.0, m.p, &t->floor_list, t->stream);
        launch_morton_to_rowmajor(cdepth, m.p, r.p, t->stream);
        PRX_HIP(hipMemcpyAsync(rhs_rowmajor, r.p, sizeof(double) * nc, hipMemcpyDeviceToHost, t->stream));
        PRX_HIP(hipStreamSynchronize(t->stream));
    }
    PRX_API_END
}

int prx_prolongate(prx_tree* t, uint32_t cdepth, const double* coarse, double* fine) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && coarse && fine, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(cdepth <= t->D, PRX_ERR_INVALID, "coarse depth above the tree depth");
    set_device(t);
    const int64_t nc = (int64_t)1 << (3 * cdepth);
    DArr<double> r, m, f;
    r.ensure(nc);
    m.ensure(nc);
    f.ensure(t->N);
    PRX_HIP(hipMemcpyAsync(r.p, coarse, sizeof(double) * nc, hipMemcpyHostToDevice, t->stream));
    launch_rowmajor_to_morton(cdepth, r.p, m.p, t->stream);
    launch_prolongate(t->D, cdepth, m.p, f.p, t->stream);
    PRX_HIP(hipMemcpyAsync(fine, f.p, sizeof(double) * t->N, hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    PRX_API_END
}

int prx_gs(prx_tree* t, int precision, uint32_t sweeps, const double* b, double* x) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && b && x, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    set_device(t);
    const LevelView L = t->fine_view();
    const int64_t n = t->N;
    const hipStream_t s = t->stream;
    t->r64.ensure(n);
    t->d64.ensure(n);
    PRX_HIP(hipMemcpyAsync(t->r64.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyAsync(t->d64.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (precision == PRX_PREC_MIXED) {
        t->d32.ensure(n);
        t->q32.ensure(n);
        launch_cast_scale_d2f(t->q32.p, t->r64.p, 1.0 / t->h, n, s);
        launch_cast_scale_d2f(t->d32.p, t->d64.p, 1.0, n, s);
        prx::gs_sweeps<float>(t, L, t->d32.p, t->q32.p, (int)sweeps);
        std::vector<float> hx(n);
        PRX_HIP(hipMemcpyAsync(hx.data(), t->d32.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; ++i) x[i] = hx[i];
    } else {
        prx::gs_sweeps<double>(t, L, t->d64.p, t->r64.p, (int)sweeps);
        PRX_HIP(hipMemcpyAsync(x, t->d64.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
    }
    PRX_API_END
}

int prx_cg(prx_tree* t, int precision, const double* b, double* x, uint64_t max_iters, double eps, int residual_mode,
           int refresh_period, uint64_t* iters, double* rel) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && b && x, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    PRX_REQUIRE(precision == PRX_PREC_FP64, PRX_ERR_INVALID, "prx_cg runs the FP64 CG; mixed precision is a solve option");
    set_device(t);
    t->ensure_state();
    const LevelView L = t->fine_view();
    const int64_t n = t->N;
    const hipStream_t s = t->stream;
    DArr<double> db, dx;
    db.ensure(n);
    dx.ensure(n);
    t->r64.ensure(2 * n); t->d64.ensure(CG_XM * n); t->q64.ensure(2 * n);
    PRX_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
    double bb = residual_mode == PRX_RESID_TRUE ? prx::dev_dot<double>(t, db.p, db.p, n) : 0.0;
    prx::Cg<double> cg = prx::make_cg<double>(t, L, dx.p, db.p, t->r64.p, t->d64.p, t->q64.p, refresh_period, prx::CG_XM,
                                              t->r64.p + n, t->q64.p + n);
    prx::CgOut o = prx::run_cg<double>(cg, max_iters, eps, residual_mode, bb);
    PRX_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    PRX_HIP(hipStreamSynchronize(s));
    if (iters) *iters = o.iters;
    if (rel) *rel = o.rel;
    PRX_API_END
}

int prx_solve_cascadic(prx_tree* t, uint32_t base_depth, uint32_t gs, uint64_t cg_iters, double eps, const prx_solve_opts* opts,
                       prx_solve_report* report) {
    PRX_API_BEGIN
    PRX_REQUIRE(t, PRX_ERR_INVALID, "null tree");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    set_device(t);
    prx::SolveCfg c;
    c.base = base_depth;
    c.gs = (int)gs;
    c.cg_iters = cg_iters;
    c.eps = eps;
    prx_solve_opts_default(&c.o);
    if (opts) c.o = *opts;
    PRX_REQUIRE(c.o.cg_form >= PRX_CG_FUSED && c.o.cg_form <= PRX_CG_FUSED_FACES, PRX_ERR_INVALID, "unknown cg_form");
    prx_solve_report rep;
    prx::solve_cascadic(t, c, rep);
    if (report) *report = rep;
    PRX_API_END
}

int prx_get_solution(prx_tree* t, double* x) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && x, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->assembled, PRX_ERR_STATE, "assemble not called");
    set_device(t);
    PRX_HIP(hipMemcpyAsync(x, t->x.p, sizeof(double) * t->N, hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    PRX_API_END
}

int prx_set_solution(prx_tree* t, const double* x) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && x, PRX_ERR_INVALID, "null pointer");
    set_device(t);
    t->x.ensure(t->N);
    PRX_HIP(hipMemcpyAsync(t->x.p, x, sizeof(double) * t->N, hipMemcpyHostToDevice, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    PRX_API_END
}

int prx_evaluate(prx_tree* t, const double* q, size_t m, double* out) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && (q || m == 0) && (out || m == 0), PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->x.p, PRX_ERR_STATE, "no solution");
    set_device(t);
    if (m == 0) return PRX_OK;
    DArr<double> dq, dout;
    dq.ensure(3 * (int64_t)m);
    dout.ensure((int64_t)m);
    PRX_HIP(hipMemcpyAsync(dq.p, q, sizeof(double) * 3 * m, hipMemcpyHostToDevice, t->stream));
    launch_evaluate(t->D, t->x.p, dq.p, (int64_t)m, dout.p, t->stream);
    PRX_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * m, hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    PRX_API_END
}

int prx_evaluate_device(prx_tree* t, const double* q_device, size_t m, double* out_device) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && (q_device || m == 0) && (out_device || m == 0), PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->x.p, PRX_ERR_STATE, "no solution");
    set_device(t);
    if (m == 0) return PRX_OK;
    launch_evaluate(t->D, t->x.p, q_device, (int64_t)m, out_device, t->stream);
    PRX_HIP(hipStreamSynchronize(t->stream));
    PRX_API_END
}

int prx_evaluate_corners(prx_tree* t, double* out) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && out, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->x.p, PRX_ERR_STATE, "no solution");
    set_device(t);
    const int64_t r1 = (int64_t)t->res + 1, n = r1 * r1 * r1;
    DArr<double> d;
    d.ensure(n);
    launch_evaluate_corners(t->D, t->x.p, d.p, t->stream);
    PRX_HIP(hipMemcpyAsync(out, d.p, sizeof(double) * n, hipMemcpyDeviceToHost, t->stream));
    PRX_HIP(hipStreamSynchronize(t->stream));
    PRX_API_END
}

int prx_time_evaluate_corners(prx_tree* t, int reps, int general, double* ms, double* sumsq) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && reps > 0, PRX_ERR_INVALID, "bad argument");
    PRX_REQUIRE(t->x.p, PRX_ERR_STATE, "no solution");
    set_device(t);
    t->ensure_state();
    const int64_t r1 = (int64_t)t->res + 1, n = r1 * r1 * r1;
    DArr<double> d;
    d.ensure(n);
    auto run = [&] {
        if (general) launch_evaluate_corners_general(t->D, t->x.p, d.p, t->stream);
        else launch_evaluate_corners(t->D, t->x.p, d.p, t->stream);
    };
    run();
    hipEvent_t e0, e1;
    PRX_HIP(hipEventCreate(&e0));
    PRX_HIP(hipEventCreate(&e1));
    PRX_HIP(hipEventRecord(e0, t->stream));
    for (int i = 0; i < reps; ++i) run();
    PRX_HIP(hipEventRecord(e1, t->stream));
    PRX_HIP(hipEventSynchronize(e1));
    float f = 0.f;
    PRX_HIP(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = (double)f / reps;
    const double ss = dev_dot<double>(t, d.p, d.p, n);
    if (sumsq) *sumsq = ss;
    PRX_API_END
}

int prx_iso_value(prx_tree* t, double* iso) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && iso, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->x.p && t->splatted, PRX_ERR_STATE, "no solution");
    set_device(t);
    t->ensure_state();
    if (t->P == 0) { *iso = 0.0; return PRX_OK; }
    DArr<double> v;
    v.ensure(t->P);
    launch_evaluate_pointsSoA(t->D, t->x.p, t->pos.p, t->P, v.p, t->stream);
    DArr<double> ones;
    ones.ensure(t->P);
    launch_fill<double>(ones.p, 1.0, t->P, t->stream);
    const double s = prx::dev_dot<double>(t, v.p, ones.p, t->P);
    *iso = s / (double)t->P;
    PRX_API_END
}

int prx_vjp_points(prx_tree* t, const double* gF, double* gpos, double* gnrm, int precision, double eps,
                   prx_vjp_report* report) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && (gF || t->P == 0) && (gpos || t->P == 0) && (gnrm || t->P == 0), PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(eps > 0.0 && eps < 1.0, PRX_ERR_INVALID, "eps must be in (0,1)");
    set_device(t);
    prx::vjp_points(t, gF, gpos, gnrm, precision, eps, report);
    PRX_API_END
}

int prx_vjp_points_device(prx_tree* t, const double* gF, double* gpos, double* gnrm, int precision, double eps,
                          prx_vjp_report* report) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && (gF || t->P == 0) && (gpos || t->P == 0) && (gnrm || t->P == 0), PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(eps > 0.0 && eps < 1.0, PRX_ERR_INVALID, "eps must be in (0,1)");
    set_device(t);
    prx::vjp_points(t, gF, gpos, gnrm, precision, eps, report, true);
    PRX_API_END
}

int prx_csr_spmv(int64_t n, const int64_t* off, const int64_t* col, const double* val, const double* x, double* y) {
    PRX_API_BEGIN
    PRX_REQUIRE(n >= 0 && off && col && val && x && y, PRX_ERR_INVALID, "null pointer");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw prx::Error(PRX_ERR_DEVICE, "no HIP device available");
    const int64_t nnz = off[n];
    hipStream_t s;
    PRX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DArr<int64_t> doff, dcol;
    DArr<double> dval, dx, dy;
    doff.ensure(n + 1); dcol.ensure(nnz); dval.ensure(nnz); dx.ensure(n); dy.ensure(n);
    PRX_HIP(hipMemcpyAsync(doff.p, off, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyAsync(dcol.p, col, sizeof(int64_t) * nnz, hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyAsync(dval.p, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
    launch_csr_spmv(n, doff.p, dcol.p, dval.p, dx.p, dy.p, s);
    PRX_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    PRX_HIP(hipStreamSynchronize(s));
    PRX_HIP(hipStreamDestroy(s));
    PRX_API_END
}

}  // extern "C"
