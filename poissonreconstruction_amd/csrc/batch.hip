// Batched screened-Poisson solves: K point clouds of uneven size solved
// together.  Each cloud is a complete FEMTree (its own depth, points, splat,
// RHS and screening lists, built by the single-tree code), and the solve runs
// over one concatenated node array: instance k owns nodes [off[k], off[k+1])
// (each a multiple of 4096), so every 16^3 brick lies in one instance and
// every batched kernel finds its instance's weights, CG scalars and run flag
// from the brick.  The per-instance algorithm is the cascadic solve of
// FEMTree::solve_cascadic (src/fem_tree.rs:458-487) at base depth D-2: coarse
// FP64 CG after multicolour GS, prolongation (src/fem_tree.rs:310-343), fine
// GS and CG (src/solvers.rs:20-69), with the stopping rules of the single
// tree (reference ||r|| <= eps ||r0|| or the FP64 true residual).
//
// One CG iteration of all instances is the fused CG step of the single tree
// (screening gather, one operator pass, a two-stage fold; x every CG_XM
// iterations) or, with the two-pass form, five launches (gather, direction update +
// operator, alpha, vector update, beta); a chunk of iterations is captured
// once into a hipGraph and replayed until every instance is done, and the GS
// sweeps are captured the same way.  An
// autotuner picks the operator's brick size, the GS block size and the chunk
// length by timing them on the batch itself.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <functional>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "tree.hpp"

namespace prx {

__global__ void k_copy_add_u32(uint32_t* dst, const uint32_t* src, int64_t n, uint32_t add) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + add;
}
__global__ void k_fill_u32(uint32_t* dst, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}
static unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }
static void copy_add_u32(uint32_t* dst, const uint32_t* src, int64_t n, uint32_t add, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_add_u32, dim3(blocks_for(n)), dim3(256), 0, s, dst, src, n, add);
    PRX_LAUNCH_CHECK();
}
static void fill_u32(uint32_t* dst, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n)), dim3(256), 0, s, dst, n, v);
    PRX_LAUNCH_CHECK();
}
template <typename T>
static void dmalloc_n(T** p, int64_t n) {
    PRX_HIP(hipMalloc((void**)p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
}

// One level (fine or coarse) of all instances, concatenated.
struct BLevel {
    int K = 0;
    int64_t N = 0;
    std::vector<uint64_t> off_h;
    std::vector<BCoef> coef_h;
    DArr<uint64_t> off;
    DArr<uint32_t> brick_inst;
    DArr<BCoef> coef;
    DArr<double> x, b, r, dd, q, sbuf;
    // the mixed-precision fine solve: FP32 correction e, its CG vectors and
    // screening values, and x2 (x + e formed by the FP64 residual update)
    DArr<float> e32, r32, dd32, q32, sbuf32;
    DArr<double> x2;
    NodeList scr;  // global-id screening list (fine level)
    bool has_scr = false;
    int64_t P = 0;
    DArr<uint32_t> cell, pinst;
    DArr<uint32_t> pt_off;  // [K+1] first sorted point of each instance (instances without screening have none)
    DArr<double> t64;
    DArr<float> t32;
    // the face-form FP32 pass: the instances' occupied cells and (cell, corner)
    // pairs concatenated (global offsets), and [N] q at the touched nodes,
    // [T] screening sums per touched node, [8C] sums per cell corner
    CellLists cells;
    bool has_cells = false;
    DArr<float> qt32;
    BLevel() { scr.mode = 1; }
    ~BLevel() {
        free_list(scr);
        free_cell_lists(cells);
    }
    BPoints pts() const {
        BPoints p;
        p.P = P;
        p.cell = cell.p;
        p.t64 = t64.p;
        p.t32 = t32.p;
        p.inst = pinst.p;
        return p;
    }
    const NodeList* scr_ptr() const { return has_scr ? &scr : nullptr; }
};

struct GraphEntry {
    const BLevel* L;
    int kind, chunk, refresh, tile, gsb, prec;
    bool fused, faces;
    hipGraphExec_t exec;
};

// The vectors one precision's CG and smoothing run on over a level: FP64 on
// the solution itself, FP32 on the correction e of the mixed solve.
template <typename T>
struct CgVecs {
    T *x, *r, *q, *dd, *sbuf;
};
static CgVecs<double> vecs_of(BLevel& L, double*) { return {L.x.p, L.r.p, L.q.p, L.dd.p, L.sbuf.p}; }
static CgVecs<float> vecs_of(BLevel& L, float*) { return {L.e32.p, L.r32.p, L.q32.p, L.dd32.p, L.sbuf32.p}; }
template <typename T>
static CgVecs<T> vecs(BLevel& L) { return vecs_of(L, (T*)nullptr); }

}  // namespace prx

struct prx_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<prx_tree*> inst;
    prx::BLevel fine, coarse;
    bool built = false;
    prx::DArr<prx::CgState> st;
    prx::CgState* st_host = nullptr;
    prx::DArr<double> partial, fold;
    prx::DArr<int32_t> flags, ones;
    std::vector<int32_t> flags_h;
    int tile = 16, chunk = 16, gs_block = 16;  // gs_block: brick edge of the GS passes
    bool fused = true;                         // the fused CG step (prx_solve_opts.cg_form != PRX_CG_TWO_PASS)
    bool faces = false;                        // FP32 fused step in face form (cg_form PRX_CG_FUSED_FACES)
    bool tuned = false;
    double tune_ms = 0.0;
    uint64_t graph_launches = 0;
    // the persistent CG launches: control block, per-instance state, work lists
    prx::DArr<prx::PCtl> pctl;
    prx::DArr<prx::PInst> pinst;
    prx::DArr<int32_t> plist_i;
    prx::DArr<uint32_t> plist_u;
    prx::DArr<uint8_t> pargs;  // the argument block of a persistent launch
    uint64_t persistent_launches = 0, persistent_rounds = 0;
    int persistent_blocks = 0, persistent_xcd_match = 0;
    double persistent_phase_ms[4] = {0.0, 0.0, 0.0, 0.0};  // of the current solve
    std::vector<prx::GraphEntry> graphs;

    void drop_graphs() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
    ~prx_batch() {
        drop_graphs();
        for (prx_tree* t : inst) delete t;
        if (st_host) (void)hipHostFree(st_host);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace prx {

static BatchTab tab_of(prx_batch* B, const BLevel& L, const int32_t* run, int stride, int when) {
    BatchTab t;
    t.K = L.K;
    t.off = L.off.p;
    t.brick_inst = L.brick_inst.p;
    t.coef = L.coef.p;
    t.st = B->st.p;
    t.run = run;
    t.run_stride = stride;
    t.run_when = when;
    return t;
}
// instances whose CG is still running (done == 0)
static BatchTab tab_running(prx_batch* B, const BLevel& L) {
    return tab_of(B, L, &B->st.p[0].done, (int)(sizeof(CgState) / sizeof(int32_t)), 0);
}
static BatchTab tab_all(prx_batch* B, const BLevel& L) { return tab_of(B, L, B->ones.p, 1, 1); }
static BatchTab tab_flagged(prx_batch* B, const BLevel& L) { return tab_of(B, L, B->flags.p, 1, 1); }

static void read_states(prx_batch* B) {
    PRX_HIP(hipMemcpyAsync(B->st_host, B->st.p, sizeof(CgState) * B->inst.size(), hipMemcpyDeviceToHost, B->stream));
    PRX_HIP(hipStreamSynchronize(B->stream));
}
static void write_states(prx_batch* B) {
    PRX_HIP(hipMemcpyAsync(B->st.p, B->st_host, sizeof(CgState) * B->inst.size(), hipMemcpyHostToDevice, B->stream));
}
static void write_flags(prx_batch* B) {
    PRX_HIP(hipMemcpyAsync(B->flags.p, B->flags_h.data(), sizeof(int32_t) * B->flags_h.size(), hipMemcpyHostToDevice, B->stream));
}

static void setup_level(prx_batch* B, BLevel& L, const std::vector<uint32_t>& depth, const std::vector<BCoef>& coef) {
    const hipStream_t s = B->stream;
    L.K = (int)depth.size();
    L.off_h.assign(L.K + 1, 0);
    for (int k = 0; k < L.K; ++k) L.off_h[k + 1] = L.off_h[k] + ((uint64_t)1 << (3 * depth[k]));
    L.N = (int64_t)L.off_h[L.K];
    if (L.N >= ((int64_t)1 << 32)) throw Error(PRX_ERR_INVALID, "a batch holds at most 2^32 - 1 nodes per level");
    L.coef_h = coef;
    L.off.ensure(L.K + 1);
    L.coef.ensure(L.K);
    PRX_HIP(hipMemcpyAsync(L.off.p, L.off_h.data(), sizeof(uint64_t) * (L.K + 1), hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyAsync(L.coef.p, L.coef_h.data(), sizeof(BCoef) * L.K, hipMemcpyHostToDevice, s));
    std::vector<uint32_t> bi((size_t)(L.N >> 12));
    for (int k = 0; k < L.K; ++k)
        for (uint64_t g = L.off_h[k] >> 12; g < L.off_h[k + 1] >> 12; ++g) bi[g] = (uint32_t)k;
    L.brick_inst.ensure((int64_t)bi.size());
    PRX_HIP(hipMemcpyAsync(L.brick_inst.p, bi.data(), sizeof(uint32_t) * bi.size(), hipMemcpyHostToDevice, s));
    L.x.ensure(L.N);
    L.b.ensure(L.N);
    PRX_HIP(hipStreamSynchronize(s));  // bi goes out of scope
}

// The cascadic CG's vectors of a level (allocated by the batched cascadic
// solve, released by the multigrid one, which solves through the trees)
static void ensure_cg(BLevel& L) {
    L.r.ensure(2 * L.N);       // fused CG: r and q in ping-pong pairs
    L.dd.ensure(CG_XM * L.N);  // directions: a ring of CG_XM (the two-pass form uses two)
    L.q.ensure(2 * L.N);
}
static void release_cg(BLevel& L) {
    L.r.release();
    L.dd.release();
    L.q.release();
}
// the mixed fine solve's vectors (the FP64 CG vectors of the level are not needed then)
static void ensure_cg32(BLevel& L, hipStream_t s) {
    L.e32.ensure(L.N);
    L.r32.ensure(2 * L.N);
    L.dd32.ensure(CG_XM * L.N);
    L.q32.ensure(2 * L.N);
    L.x2.ensure(L.N);
    L.sbuf32.ensure(std::max<int64_t>(L.P, 1));
    if (L.has_cells && !L.qt32.p) {
        // q at the touched nodes (zero elsewhere), the sums per touched node and per cell corner
        const int64_t n = L.N + std::max<int64_t>(L.scr.T, 1) + 8 * std::max<int64_t>(L.cells.C, 1);
        L.qt32.ensure(n);
        launch_fill<float>(L.qt32.p, 0.0f, n, s);
    }
}
static void release_cg32(BLevel& L) {
    L.e32.release();
    L.r32.release();
    L.dd32.release();
    L.q32.release();
    L.x2.release();
    L.sbuf32.release();
    L.qt32.release();
}

// Concatenate the instances' occupied-cell lists (face-form screening) in the
// order of the screening lists: cell starts offset by the instance's first
// sorted point, pair ranges by the pairs before it, pairs by 8 x its first cell.
static void build_cells(prx_batch* B, BLevel& L) {
    const hipStream_t s = B->stream;
    free_cell_lists(L.cells);
    L.has_cells = false;
    int64_t C = 0, MC = 0;
    std::vector<int64_t> mc(L.K, 0);
    for (int k = 0; k < L.K; ++k) {
        prx_tree* t = B->inst[k];
        if (!(t->pw > 0.0 && t->trunc_list.T > 0)) continue;
        if (!t->cells_ok) return;  // no cell lists for this instance: the batch keeps the stored-q pass
        uint32_t m = 0;
        PRX_HIP(hipMemcpy(&m, t->cells.coff + t->trunc_list.T, sizeof(uint32_t), hipMemcpyDeviceToHost));
        mc[k] = m;
        C += t->cells.C;
        MC += m;
    }
    if (C == 0 || C >= ((int64_t)1 << 28) || MC >= ((int64_t)1 << 32)) return;
    CellLists& G = L.cells;
    G.C = C;
    dmalloc_n(&G.cstart, C + 1);
    dmalloc_n(&G.coff, L.scr.T + 1);
    dmalloc_n(&G.cval, MC);
    int64_t Co = 0, MCo = 0, Po = 0, To = 0;
    for (int k = 0; k < L.K; ++k) {
        prx_tree* t = B->inst[k];
        if (!(t->pw > 0.0 && t->trunc_list.T > 0)) continue;
        const CellLists& S = t->cells;
        copy_add_u32(G.cstart + Co, S.cstart, S.C, (uint32_t)Po, s);
        copy_add_u32(G.coff + To, S.coff, t->trunc_list.T, (uint32_t)MCo, s);
        copy_add_u32(G.cval + MCo, S.cval, mc[k], (uint32_t)(8 * Co), s);
        Co += S.C;
        MCo += mc[k];
        Po += t->P;
        To += t->trunc_list.T;
    }
    fill_u32(G.cstart + C, 1, (uint32_t)Po, s);
    fill_u32(G.coff + To, 1, (uint32_t)MCo, s);
    PRX_HIP(hipStreamSynchronize(s));
    L.has_cells = true;
}

// Concatenate the per-instance screening lists and point data (global node
// ids and sorted point positions).
static void build_screening(prx_batch* B, BLevel& L) {
    const hipStream_t s = B->stream;
    free_list(L.scr);
    int64_t T = 0, M = 0, P = 0;
    for (prx_tree* t : B->inst)
        if (t->pw > 0.0 && t->trunc_list.T > 0) { T += t->trunc_list.T; M += t->trunc_list.M; P += t->P; }
    L.has_scr = T > 0;
    L.P = P;
    {
        std::vector<uint32_t> po(L.K + 1, 0);
        for (int k = 0; k < L.K; ++k) {
            prx_tree* t = B->inst[k];
            po[k + 1] = po[k] + (uint32_t)((t->pw > 0.0 && t->trunc_list.T > 0) ? t->P : 0);
        }
        L.pt_off.ensure(L.K + 1);
        PRX_HIP(hipMemcpyAsync(L.pt_off.p, po.data(), sizeof(uint32_t) * (L.K + 1), hipMemcpyHostToDevice, s));
        PRX_HIP(hipStreamSynchronize(s));
    }
    if (!L.has_scr) return;
    if (M >= ((int64_t)1 << 32) || P >= ((int64_t)1 << 32)) throw Error(PRX_ERR_INVALID, "batch screening list too large");
    NodeList& G = L.scr;
    G.T = T;
    G.M = M;
    G.brick_log = 12;
    dmalloc_n(&G.node, T);
    dmalloc_n(&G.off, T + 1);
    dmalloc_n(&G.sp, M);
    dmalloc_n(&G.phi, M);
    dmalloc_n(&G.phi32, M);
    dmalloc_n(&G.diag, T);
    dmalloc_n(&G.brick, (L.N >> 12) + 1);
    dmalloc_n(&G.brick8, (L.N >> 9) + 1);
    L.cell.ensure(3 * P);
    L.t64.ensure(3 * P);
    L.t32.ensure(3 * P);
    L.pinst.ensure(P);
    L.sbuf.ensure(P);
    int64_t To = 0, Mo = 0, Po = 0;
    for (int k = 0; k < L.K; ++k) {
        prx_tree* t = B->inst[k];
        const uint64_t o = L.off_h[k];
        const int64_t nb = (int64_t)(((uint64_t)1 << (3 * t->D)) >> 12), nb8 = nb * 8;
        if (!(t->pw > 0.0 && t->trunc_list.T > 0)) {
            fill_u32(G.brick + (o >> 12), nb, (uint32_t)To, s);
            fill_u32(G.brick8 + (o >> 9), nb8, (uint32_t)To, s);
            continue;
        }
        const NodeList& S = t->trunc_list;
        if (S.brick_log != 12 || !S.brick8) throw Error(PRX_ERR_STATE, "batch instance without 16^3 brick ranges");
        copy_add_u32(G.node + To, S.node, S.T, (uint32_t)o, s);
        copy_add_u32(G.off + To, S.off, S.T, (uint32_t)Mo, s);
        copy_add_u32(G.sp + Mo, S.sp, S.M, (uint32_t)Po, s);
        PRX_HIP(hipMemcpyAsync(G.phi + Mo, S.phi, sizeof(double) * S.M, hipMemcpyDeviceToDevice, s));
        PRX_HIP(hipMemcpyAsync(G.phi32 + Mo, S.phi32, sizeof(float) * S.M, hipMemcpyDeviceToDevice, s));
        PRX_HIP(hipMemcpyAsync(G.diag + To, S.diag, sizeof(double) * S.T, hipMemcpyDeviceToDevice, s));
        copy_add_u32(G.brick + (o >> 12), S.brick, nb, (uint32_t)To, s);
        copy_add_u32(G.brick8 + (o >> 9), S.brick8, nb8, (uint32_t)To, s);
        for (int a = 0; a < 3; ++a) {
            PRX_HIP(hipMemcpyAsync(L.cell.p + a * P + Po, t->cell.p + a * t->P, sizeof(uint32_t) * t->P, hipMemcpyDeviceToDevice, s));
            PRX_HIP(hipMemcpyAsync(L.t64.p + a * P + Po, t->t64.p + a * t->P, sizeof(double) * t->P, hipMemcpyDeviceToDevice, s));
            PRX_HIP(hipMemcpyAsync(L.t32.p + a * P + Po, t->t32.p + a * t->P, sizeof(float) * t->P, hipMemcpyDeviceToDevice, s));
        }
        fill_u32(L.pinst.p + Po, t->P, (uint32_t)k, s);
        To += S.T;
        Mo += S.M;
        Po += t->P;
    }
    fill_u32(G.off + T, 1, (uint32_t)M, s);
    fill_u32(G.brick + (L.N >> 12), 1, (uint32_t)T, s);
    fill_u32(G.brick8 + (L.N >> 9), 1, (uint32_t)T, s);
    PRX_HIP(hipStreamSynchronize(s));
    build_cells(B, L);
}

static void build(prx_batch* B) {
    if (B->built) return;
    B->drop_graphs();
    const int K = (int)B->inst.size();
    std::vector<uint32_t> dF(K), dC(K);
    std::vector<BCoef> cF(K), cC(K);
    for (int k = 0; k < K; ++k) {
        prx_tree* t = B->inst[k];
        dF[k] = t->D;
        dC[k] = t->D - 2;
        BCoef f{};
        f.D = t->D;
        f.res = t->res;
        f.s64 = -t->k64.cc;
        f.wg64 = t->k64.wg;
        f.wh3_64 = t->k64.wh3;
        f.s32 = -t->k32.cc;
        f.wg32 = t->k32.wg;
        f.wh3_32 = t->k32.wh3;
        cF[k] = f;
        // coarse level: src/fem_tree.rs:346-404 (effective pw = pw 4^(D-d), diagonal wh3*1.005)
        const uint32_t cd = t->D - 2;
        const double hc = 1.0 / (double)(1u << cd);
        const double epw = t->pw * (double)((uint64_t)1 << (2 * (t->D - cd)));
        const double wh3c = epw > 0.0 ? epw * (hc * hc * hc) / 1.5 : 0.0;
        BCoef c{};
        c.D = cd;
        c.res = 1u << cd;
        c.s64 = -stiffness_entry(hc, 1, 1, 1);
        c.wg64 = epw > 0.0 ? wh3c * 1.005 : 0.0;
        c.wh3_64 = 0.0;
        cC[k] = c;
    }
    setup_level(B, B->fine, dF, cF);
    setup_level(B, B->coarse, dC, cC);
    build_screening(B, B->fine);
    {
        std::vector<uint32_t> zero(K + 1, 0);
        B->coarse.pt_off.ensure(K + 1);
        PRX_HIP(hipMemcpyAsync(B->coarse.pt_off.p, zero.data(), sizeof(uint32_t) * (K + 1), hipMemcpyHostToDevice, B->stream));
        PRX_HIP(hipStreamSynchronize(B->stream));
    }
    B->pctl.ensure(1);
    B->pinst.ensure(K);
    B->plist_i.ensure(3 * (int64_t)K);
    B->plist_u.ensure(12 * ((int64_t)K + 1) + 3 * (int64_t)K);
    B->st.ensure(K);
    if (B->st_host) (void)hipHostFree(B->st_host);
    B->st_host = nullptr;
    PRX_HIP(hipHostMalloc((void**)&B->st_host, sizeof(CgState) * std::max(K, 1), hipHostMallocDefault));
    B->flags.ensure(K);
    B->ones.ensure(K);
    B->flags_h.assign(K, 0);
    std::vector<int32_t> one(K, 1);
    PRX_HIP(hipMemcpyAsync(B->ones.p, one.data(), sizeof(int32_t) * K, hipMemcpyHostToDevice, B->stream));
    B->partial.ensure(4 * (std::max(B->fine.N, B->coarse.N) >> 9) + 64);  // four per 8^3 brick (fused CG step)
    B->fold.ensure((int64_t)K * FOLD_BLOCKS * 4);
    PRX_HIP(hipStreamSynchronize(B->stream));
    B->built = true;
}

// ------------------------------------------------------------------ batched building blocks
static int ulog_tile(int tile) { return tile == 8 ? 9 : 12; }

// r = b - A x (into y, default L.r) for the instances selected by `tab`;
// partial per brick.
static void resid_pass(prx_batch* B, BLevel& L, const BatchTab& tab, double* y = nullptr) {
    StencilIO<double> io;
    io.x = L.x.p;
    io.b = L.b.p;
    io.y = y ? y : L.r.p;
    io.partial = B->partial.p;
    if (L.has_scr) {
        launch_bscreen_gather<double>(tab, L.pts(), L.x.p, nullptr, nullptr, L.sbuf.p, nullptr, nullptr, B->stream);
        io.s = L.sbuf.p;
    }
    launch_bstencil<double>(tab, L.N, L.scr_ptr(), OP_RESID, io, B->tile, B->stream);
}

// Multicolour GS sweeps of every instance: one brick-tiled pass per colour,
// the same pass (and arithmetic) as the single-tree sweep.
// FP64: on x with right-hand side b; FP32: on the correction e with the
// scaled residual in r slot 0.
template <typename T>
static void gs_enqueue(prx_batch* B, BLevel& L, int sweeps) {
    const BatchTab tab = tab_all(B, L);
    const CgVecs<T> V = vecs<T>(L);
    StencilIO<T> io;
    io.x = V.x;
    io.y = V.x;
    if constexpr (sizeof(T) == 8) io.b = L.b.p;
    else io.b = V.r;
    io.s = V.sbuf;
    for (int k = 0; k < sweeps; ++k)
        for (int c = 0; c < 8; ++c) {
            if (L.has_scr) launch_bscreen_gather<T>(tab, L.pts(), V.x, nullptr, nullptr, V.sbuf, nullptr, nullptr, B->stream);
            io.color = c;
            launch_bstencil<T>(tab, L.N, L.scr_ptr(), OP_GS, io, B->gs_block, B->stream);
        }
}

// The FP32 fused step runs in face form on 16^3 bricks when the level has its
// cell lists (or no screening) and the options ask for it.
template <typename T>
static bool uses_faces(const prx_batch* B, const BLevel& L) {
    return sizeof(T) == 4 && B->faces && B->tile == 16 && (!L.has_scr || (L.has_cells && L.qt32.p));
}

// One fused CG step (pass index k) of every running instance, as
// Cg::enqueue_fused: pass k reads r, q from slot k&1 and the direction from
// ring slot k % CG_XM and writes slot (k+1)&1 and ring slot (k+1) % CG_XM.
// xn: the x terms pending before the pass (the same for every running
// instance: they start or restart together at pass 0).
template <typename T>
static void cg_enqueue_fused(prx_batch* B, BLevel& L, uint64_t k, int refresh, int& xn) {
    const hipStream_t s = B->stream;
    const BatchTab tab = tab_running(B, L);
    const int64_t N = L.N;
    const CgVecs<T> V = vecs<T>(L);
    auto slot = [&](uint64_t j) { return V.dd + (int64_t)(j % CG_XM) * N; };
    T* dold = slot(k);
    T* dnew = slot(k + 1);
    if (xn == CG_XM) {  // the instances also check their own count
        XDirs<T> d{};
        for (int i = 0; i < CG_XM; ++i) d.d[i] = slot(k + 1 + (uint64_t)i);
        launch_bx_terms<T>(tab, V.x, d, V.dd, N, 0, s);
    }
    StencilIO<T> io;
    io.r = V.r + (int64_t)(k & 1) * N;
    io.q = V.q + (int64_t)(k & 1) * N;
    io.rout = V.r + (int64_t)((k + 1) & 1) * N;
    io.dold = dold;
    io.dnew = dnew;
    io.y = V.q + (int64_t)((k + 1) & 1) * N;
    io.partial = B->partial.p;
    if (uses_faces<T>(B, L)) {
        // the face form: io.q / io.y hold the bricks' faces of q, q at the touched nodes lives in qt32
        if (L.has_scr) {
            float* qt = L.qt32.p;
            float* sc = qt + N;
            float* sig = sc + std::max<int64_t>(L.scr.T, 1);
            launch_bcell_screening<float>(tab, L.cells, L.pts(), L.scr.T, L.scr.node, (const float*)io.r, (const float*)dold, qt, sig,
                                          sc, s);
            io.sc = (const T*)sc;
            io.qt = (T*)qt;
        }
        launch_bcg_faces<T>(tab, N, L.scr_ptr(), io, s);
    } else {
        if (L.has_scr) {
            launch_bscreen_gather<T>(tab, L.pts(), nullptr, nullptr, nullptr, V.sbuf, io.r, dold, s, io.q);
            io.s = V.sbuf;
        }
        launch_bstencil<T>(tab, N, L.scr_ptr(), OP_CG, io, B->tile, s);
    }
    const bool fresh = refresh > 0 && (k + 1) % (uint64_t)refresh == 0;
    launch_bfin_cg(tab, B->partial.p, ulog_tile(B->tile), B->fold.p, fresh ? 1 : 0, s);
    xn = xn == CG_XM ? 1 : xn + 1;
    if (fresh) {  // x completed, then r = b - A x replaces the pending update
        if constexpr (sizeof(T) == 8) {
            XDirs<double> d{};
            for (int i = 0; i < xn; ++i) d.d[i] = slot(k + 2 - (uint64_t)xn + (uint64_t)i);
            launch_bx_terms<double>(tab, L.x.p, d, L.dd.p, N, 2, s);
            resid_pass(B, L, tab, io.rout);
            launch_bfin(tab, B->partial.p, ulog_tile(B->tile), FIN_FRESH, 0, s);
            xn = 0;
        } else {
            throw Error(PRX_ERR_INVALID, "the mixed solve replaces its residual by the FP64 one, not by a refresh");
        }
    }
}

// One CG iteration (index k) of every running instance, two-pass form.
static void cg_enqueue(prx_batch* B, BLevel& L, uint64_t k, int refresh) {
    const hipStream_t s = B->stream;
    const BatchTab tab = tab_running(B, L);
    double* dold = L.dd.p + (k & 1) * L.N;
    double* dnew = L.dd.p + ((k + 1) & 1) * L.N;
    StencilIO<double> io;
    io.r = L.r.p;
    io.dold = dold;
    io.dnew = dnew;
    io.y = L.q.p;
    io.partial = B->partial.p;
    if (L.has_scr) {
        launch_bscreen_gather<double>(tab, L.pts(), nullptr, nullptr, nullptr, L.sbuf.p, L.r.p, dold, s);
        io.s = L.sbuf.p;
    }
    launch_bstencil<double>(tab, L.N, L.scr_ptr(), OP_DIR_APPLY, io, B->tile, s);
    launch_bfin(tab, B->partial.p, ulog_tile(B->tile), FIN_ALPHA, 0, s);
    int ulog, fop = FIN_BETA;
    if (refresh > 0 && (k + 1) % (uint64_t)refresh == 0) {
        launch_baxpy_alpha<double>(tab, L.x.p, dold, dnew, L.N, s);
        resid_pass(B, L, tab);
        ulog = ulog_tile(B->tile);
        fop = FIN_BETA_FULL;
    } else {
        ulog = launch_bcg_update<double>(tab, L.x.p, L.r.p, dold, dnew, L.q.p, L.N, B->partial.p, s);
    }
    launch_bfin(tab, B->partial.p, ulog, fop, (int)((k + 1) & 1), s);
}

// kind 0: `chunk` CG iterations starting at an even index; kind 1: GS
// sweeps.  prec 1: the FP32 correction system of the mixed fine solve.
static hipGraphExec_t graph_for(prx_batch* B, BLevel& L, int kind, int n, int refresh, int prec = 0) {
    for (auto& g : B->graphs)
        if (g.L == &L && g.kind == kind && g.chunk == n && g.refresh == refresh && g.tile == B->tile && g.gsb == B->gs_block &&
            g.fused == B->fused && g.faces == B->faces && g.prec == prec)
            return g.exec;
    hipGraph_t graph;
    PRX_HIP(hipStreamSynchronize(B->stream));
    PRX_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    try {
        if (kind == 0 && prec == 1) {
            int xn = CG_XM;
            for (int k = 0; k < n; ++k) cg_enqueue_fused<float>(B, L, (uint64_t)k, 0, xn);
        } else if (kind == 0 && B->fused) {
            // a replay starts with CG_XM terms pending (none at the very first
            // pass, which the kernels check), or none after a refresh that
            // ended the previous replay
            int xn = refresh > 0 ? 0 : CG_XM;
            for (int k = 0; k < n; ++k) cg_enqueue_fused<double>(B, L, (uint64_t)k, refresh, xn);
        } else if (kind == 0) {
            for (int k = 0; k < n; ++k) cg_enqueue(B, L, (uint64_t)k, refresh);
        } else if (prec == 1) {
            gs_enqueue<float>(B, L, n);
        } else {
            gs_enqueue<double>(B, L, n);
        }
    } catch (...) {
        hipGraph_t g2;
        (void)hipStreamEndCapture(B->stream, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        throw;
    }
    PRX_HIP(hipStreamEndCapture(B->stream, &graph));
    hipGraphExec_t exec;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    PRX_HIP(e);
    B->graphs.push_back({&L, kind, n, refresh, B->tile, B->gs_block, prec, B->fused, B->faces, exec});
    return exec;
}

static void launch_graph(prx_batch* B, hipGraphExec_t g) {
    PRX_HIP(hipGraphLaunch(g, B->stream));
    B->graph_launches++;
}

struct BOut {
    std::vector<uint64_t> iters;
    std::vector<double> rel;
    std::vector<int> status;
};

// ------------------------------------------------------------------ persistent launches
// The fields of a persistent CG launch common to both precisions: the level's
// tables, points, work-list storage, control block and scratch.
template <typename T>
static PArgs<T> persist_args(prx_batch* B, BLevel& L) {
    PArgs<T> a;
    a.tab = tab_all(B, L);
    a.N = L.N;
    a.P = L.P;
    a.cell = L.cell.p;
    a.t64 = L.t64.p;
    if constexpr (sizeof(T) == 4) a.tt = (const T*)L.t32.p;
    else a.tt = (const T*)L.t64.p;
    a.pinst = L.pinst.p;
    a.pt_off = L.pt_off.p;
    a.partial = B->partial.p;
    a.partial2 = B->partial.p + (L.N >> 10) + 16;
    a.flag = B->flags.p;
    a.pi = B->pinst.p;
    const int64_t K = L.K, K1 = K + 1;
    PList* ls[3] = {&a.run, &a.repl, &a.xt};
    for (int l = 0; l < 3; ++l) {
        ls[l]->inst = B->plist_i.p + l * K;
        ls[l]->gbrick = B->plist_u.p + 12 * K1 + l * K;
        ls[l]->brick = B->plist_u.p + (4 * l + 0) * K1;
        ls[l]->pts = B->plist_u.p + (4 * l + 1) * K1;
        ls[l]->vblk = B->plist_u.p + (4 * l + 2) * K1;
        ls[l]->kblk = B->plist_u.p + (4 * l + 3) * K1;
    }
    a.ctl = B->pctl.p;
    return a;
}

// One persistent launch over level L with the per-instance state pi; the CG
// states are read back into st_host.
template <typename T>
static void run_persistent(prx_batch* B, BLevel& L, const PArgs<T>& a, std::vector<PInst>& pi) {
    const hipStream_t s = B->stream;
    const int K = L.K;
    std::vector<int32_t> zero(K, 0);
    PCtl c{};
    PRX_HIP(hipMemcpyAsync(B->flags.p, zero.data(), sizeof(int32_t) * K, hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyAsync(B->pinst.p, pi.data(), sizeof(PInst) * K, hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyAsync(B->pctl.p, &c, sizeof(PCtl), hipMemcpyHostToDevice, s));
    int blocks = 0;
    B->pargs.ensure((int64_t)sizeof(PArgs<double>));
    launch_persistent_cg<T>(a, L.scr_ptr(), B->pargs.p, &blocks, s);
    PRX_HIP(hipMemcpyAsync(&c, B->pctl.p, sizeof(PCtl), hipMemcpyDeviceToHost, s));
    PRX_HIP(hipMemcpyAsync(pi.data(), B->pinst.p, sizeof(PInst) * K, hipMemcpyDeviceToHost, s));
    read_states(B);
    B->persistent_launches++;
    B->persistent_rounds += (uint64_t)c.rounds;
    B->persistent_blocks = blocks;
    B->persistent_xcd_match = (int)c.xcd_rr;
    int khz = 0;
    PRX_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, B->device));
    for (int i = 0; i < 4; ++i) B->persistent_phase_ms[i] += khz > 0 ? (double)c.ticks[i] / (double)khz : 0.0;
    if (std::getenv("PRX_PERSIST_TRACE") && khz > 0)
        std::fprintf(stderr, "persistent: %lld rounds (%lld with replacements), phase ms %.1f %.1f %.1f %.1f, mean pass busy per workgroup %.1f ms\n",
                     (long long)c.rounds, (long long)c.repl_rounds, (double)c.ticks[0] / khz, (double)c.ticks[1] / khz,
                     (double)c.ticks[2] / khz, (double)c.ticks[3] / khz, (double)c.pass_busy / khz / blocks);
    if (c.err == 1) throw Error(PRX_ERR_DEVICE, "persistent CG: a grid barrier wait exceeded its bound");
    if (c.err == 2) throw Error(PRX_ERR_DEVICE, "persistent CG: the round cap was reached");
    if (c.err == 3) throw Error(PRX_ERR_DEVICE, "persistent CG: the launch ran past its deadline (60 s unless PRX_PERSIST_DEADLINE_MS sets another)");
    if (c.err == 4)
        throw Error(PRX_ERR_DEVICE, "persistent CG: a brick phase processed " + std::to_string(c.bad_got) + " bricks of " +
                                        std::to_string(c.bad_want) + " (round " + std::to_string(c.bad_round / 2) +
                                        (c.bad_round & 1 ? ", CG pass)" : ", residual pass)"));
}

// CG over every instance of level L from its current x (solve_cg,
// src/solvers.rs:20-69, per instance), iterations graph-replayed in chunks.
static BOut batch_cg(prx_batch* B, BLevel& L, uint64_t max_iters, double eps, int mode, int refresh, bool persistent = false) {
    const hipStream_t s = B->stream;
    const int K = L.K;
    const BatchTab all = tab_all(B, L);
    // ||b||^2 per instance (TRUE mode threshold)
    std::vector<double> bb(K, 0.0);
    if (mode == PRX_RESID_TRUE) {
        const int ul = launch_bdot<double>(all, L.b.p, L.b.p, L.N, B->partial.p, s);
        launch_bfin(all, B->partial.p, ul, FIN_SLOT, 0, s);
        read_states(B);
        for (int k = 0; k < K; ++k) bb[k] = B->st_host[k].slot[0];
    }
    for (int k = 0; k < K; ++k) {
        CgState h{};
        h.eps = eps;
        h.max_iters = (int64_t)max_iters;
        h.bnorm2 = bb[k];
        h.lazy = 1;
        B->st_host[k] = h;
    }
    write_states(B);
    resid_pass(B, L, all);
    launch_bfin(all, B->partial.p, ulog_tile(B->tile), mode == PRX_RESID_TRUE ? FIN_INIT_TRUE : FIN_INIT_REF, 0, s);
    launch_fill<double>(L.dd.p, 0.0, L.N, s);
    int chunk = B->chunk;
    if (refresh > 0) chunk = std::max(refresh, (chunk / refresh) * refresh);
    if (chunk & 1) chunk *= 2;
    if (B->fused) {  // replays keep pass index = ring position: a multiple of CG_XM (and of the refresh)
        const int unit = refresh > 0 ? std::lcm(refresh, CG_XM) : CG_XM;
        chunk = std::max(unit, (chunk + unit / 2) / unit * unit);
    }
    if (persistent) {
        // every iteration, confirmation and restart of every instance in one launch
        if (mode != PRX_RESID_TRUE || refresh > 0 || !B->fused) throw Error(PRX_ERR_INVALID, "persistent CG: true-residual mode, fused step");
        read_states(B);
        PArgs<double> a = persist_args<double>(B, L);
        a.max_iters = (int64_t)max_iters;
        a.max_rounds = (int64_t)max_iters + 16;
        a.x = L.x.p;
        a.r = L.r.p;
        a.q = L.q.p;
        a.ring = L.dd.p;
        a.sbuf = L.sbuf.p;
        a.sbuf64 = L.sbuf.p;
        a.b = L.b.p;
        std::vector<PInst> pi(K);
        for (int k = 0; k < K; ++k) {
            pi[k] = PInst{};
            pi[k].last_rt = INFINITY;
            pi[k].fin = B->st_host[k].done ? 1 : 0;
            pi[k].status = B->st_host[k].status;
        }
        run_persistent<double>(B, L, a, pi);
        // x terms still pending (instances stopped without a confirmation)
        launch_bx_terms<double>(all, L.x.p, XDirs<double>{}, L.dd.p, L.N, 1, s);
        for (int k = 0; k < K; ++k) {
            B->st_host[k].xpend = 0;
            B->st_host[k].xn = 0;
        }
        write_states(B);
        BOut o;
        o.iters.resize(K);
        o.rel.resize(K);
        o.status.resize(K);
        for (int k = 0; k < K; ++k) {
            const CgState& h = B->st_host[k];
            o.iters[k] = (uint64_t)h.it;
            o.status[k] = h.done ? h.status : PRX_ERR_NOT_CONVERGED;
            o.rel[k] = bb[k] > 0.0 ? std::sqrt(h.delta / bb[k]) : 0.0;
        }
        PRX_HIP(hipStreamSynchronize(s));
        return o;
    }
    const hipGraphExec_t g = graph_for(B, L, 0, chunk, refresh);
    std::vector<char> fin(K, 0);
    std::vector<double> last_rt(K, INFINITY);
    read_states(B);
    for (;;) {
        // iterate until every instance is done
        for (;;) {
            bool all_done = true;
            for (int k = 0; k < K; ++k)
                if (!B->st_host[k].done) { all_done = false; break; }
            if (all_done) break;
            launch_graph(B, g);
            read_states(B);
        }
        // complete x: add the pending lazy terms
        if (B->fused) launch_bx_terms<double>(all, L.x.p, XDirs<double>{}, L.dd.p, L.N, 1, s);
        else launch_bflush_x<double>(all, L.x.p, L.dd.p, L.dd.p + L.N, L.N, s);
        for (int k = 0; k < K; ++k) {
            B->st_host[k].xpend = 0;
            B->st_host[k].xn = 0;
        }
        write_states(B);
        if (mode != PRX_RESID_TRUE) break;
        // confirm on the true residual; continue the unconfirmed instances
        bool any = false;
        for (int k = 0; k < K; ++k) {
            const CgState& h = B->st_host[k];
            const bool c = !fin[k] && h.done && h.status == PRX_OK && h.it > 0;
            if (!fin[k] && !c) fin[k] = 1;
            B->flags_h[k] = c ? 1 : 0;
            any = any || c;
        }
        if (!any) break;
        write_flags(B);
        const BatchTab fl = tab_flagged(B, L);
        resid_pass(B, L, fl);
        launch_bfin(fl, B->partial.p, ulog_tile(B->tile), FIN_SLOT, 1, s);
        read_states(B);
        for (int k = 0; k < K; ++k) {
            if (!B->flags_h[k]) continue;
            CgState& h = B->st_host[k];
            const double rt = h.slot[1];
            const bool stalled = rt > 0.25 * last_rt[k];
            if (rt <= h.thr || (uint64_t)h.it >= max_iters || stalled) {
                h.delta = rt;
                if (rt > h.thr) h.status = PRX_ERR_NOT_CONVERGED;
                fin[k] = 1;
                continue;
            }
            last_rt[k] = rt;
            const bool drifted = !(rt <= 4.0 * h.delta);
            h.beta = (!drifted && h.delta_prev > 0.0) ? rt / h.delta_prev : 0.0;
            h.delta = rt;
            h.done = 0;
            h.status = 0;
            const int64_t nk = (int64_t)(L.off_h[k + 1] - L.off_h[k]);
            if (B->fused) {
                // the graph restarts at pass 0, which reads r from slot 0 (the
                // true residual just written there, no update pending) and the
                // direction from ring slot 0: move d_{it-1} there
                h.rpend = 0;
                const int sl = (int)(((h.it - (int64_t)h.ring0) % CG_XM + CG_XM) % CG_XM);
                if (sl != 0)
                    PRX_HIP(hipMemcpyAsync(L.dd.p + L.off_h[k], L.dd.p + (int64_t)sl * L.N + L.off_h[k], sizeof(double) * nk,
                                           hipMemcpyDeviceToDevice, s));
                h.ring0 = (int32_t)(h.it % CG_XM);
            } else if (h.dir_buf == 1) {
                // the graph restarts at an even iteration index and reads the
                // direction from dd[0]; move it there if it is in dd[1]
                PRX_HIP(hipMemcpyAsync(L.dd.p + L.off_h[k], L.dd.p + L.N + L.off_h[k], sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
            }
        }
        write_states(B);
        bool running = false;
        for (int k = 0; k < K; ++k) running = running || !B->st_host[k].done;
        if (!running) break;
    }
    BOut o;
    o.iters.resize(K);
    o.rel.resize(K);
    o.status.resize(K);
    for (int k = 0; k < K; ++k) {
        const CgState& h = B->st_host[k];
        o.iters[k] = (uint64_t)h.it;
        o.status[k] = h.done ? h.status : PRX_ERR_NOT_CONVERGED;
        const double den = mode == PRX_RESID_TRUE ? bb[k] : h.delta0;
        o.rel[k] = den > 0.0 ? std::sqrt(h.delta / den) : 0.0;
    }
    PRX_HIP(hipStreamSynchronize(s));
    return o;
}

// x2 = x + e, r32 = (b - A x2) / h (the instance's 1/h) in r slot 0, partial
// |b - A x2|^2 per brick in FP64, for the instances selected by `tab`: the
// single-tree residual update pass (OP_RESID_UPD) over the batch.
static void resid_update_pass(prx_batch* B, BLevel& F, const BatchTab& tab) {
    StencilIO<double> io;
    io.x64 = F.x.p;
    io.e32 = F.e32.p;
    io.x_out = F.x2.p;
    io.b = F.b.p;
    io.r32 = F.r32.p;
    io.partial = B->partial.p;
    if (F.has_scr) {
        launch_bscreen_gather<double>(tab, F.pts(), nullptr, F.x.p, F.e32.p, F.sbuf.p, nullptr, nullptr, B->stream);
        io.s = F.sbuf.p;
    }
    launch_bstencil<double>(tab, F.N, F.scr_ptr(), OP_RESID_UPD, io, B->tile, B->stream);
}

struct MixedOut {
    BOut cg;
    std::vector<uint64_t> outer;
    std::vector<double> true_rel;
    double ms_gs = 0.0;
};

// The fine level of every instance in mixed precision, as the single-tree
// mixed solve runs it (the fine half of FEMTree::solve_cascadic,
// src/fem_tree.rs:485-486): Gauss-Seidel and CG on the FP32 correction e of
// the system scaled by 1/h, with FP64 scalars and reductions.  Whenever an
// instance's FP32 residual has fallen by theta since its last replacement
// (or its CG stops), one batched FP64 pass folds x += e and forms the true
// residual b - A x, which replaces the FP32 one while CG continues along its
// direction (a reliable update); an instance stops on its FP64 true residual
// ||b - A x|| <= eps ||b||.
static MixedOut batch_fine_mixed(prx_batch* B, BLevel& F, int gs, uint64_t max_iters, double eps, int mode, double theta,
                                 uint64_t max_outer, bool persistent = false) {
    const hipStream_t s = B->stream;
    const int K = F.K;
    const int64_t N = F.N;
    const BatchTab all = tab_all(B, F);
    MixedOut out;
    out.outer.assign(K, 0);
    out.true_rel.assign(K, 0.0);
    // ||b||^2 per instance
    for (int k = 0; k < K; ++k) B->st_host[k] = CgState{};
    write_states(B);
    int ul = launch_bdot<double>(all, F.b.p, F.b.p, N, B->partial.p, s);
    launch_bfin(all, B->partial.p, ul, FIN_SLOT, 0, s);
    read_states(B);
    std::vector<double> bb(K), thr(K), thr32(K), rr(K), r0(K);
    for (int k = 0; k < K; ++k) {
        const double ih = (double)F.coef_h[k].res;
        bb[k] = B->st_host[k].slot[0];
        thr[k] = (bb[k] * eps) * eps;
        thr32[k] = thr[k] * ih * ih;
    }
    // smoothing on the correction: e = GS^gs(0; A/h, (b - A x)/h), then x += e
    {
        Timer tg(s);
        launch_fill<float>(F.e32.p, 0.0f, N, s);
        resid_update_pass(B, F, all);
        launch_bcommit(all, F.x.p, F.x2.p, F.e32.p, N, s);
        if (gs > 0) launch_graph(B, graph_for(B, F, 1, gs, 0, 1));
        out.ms_gs = tg.ms();
    }
    resid_update_pass(B, F, all);
    launch_bfin(all, B->partial.p, ulog_tile(B->tile), FIN_SLOT, 1, s);
    launch_bcommit(all, F.x.p, F.x2.p, F.e32.p, N, s);
    read_states(B);
    for (int k = 0; k < K; ++k) {
        rr[k] = r0[k] = B->st_host[k].slot[1];
        if (mode != PRX_RESID_TRUE) {  // the reference's test: relative to the residual after smoothing
            const double ih = (double)F.coef_h[k].res;
            thr[k] = (rr[k] * eps) * eps;
            thr32[k] = thr[k] * ih * ih;
        }
    }
    // each instance's CG starts from its scaled residual (ring slot 0 = d_-1 = 0)
    for (int k = 0; k < K; ++k) {
        CgState h{};
        h.eps = 1.0;
        h.max_iters = (int64_t)max_iters;
        h.lazy = 1;
        B->st_host[k] = h;
    }
    write_states(B);
    ul = launch_bdot<float>(all, F.r32.p, F.r32.p, N, B->partial.p, s);
    launch_bfin(all, B->partial.p, ul, FIN_INIT_REF, 0, s);
    launch_fill<float>(F.dd32.p, 0.0f, N, s);
    read_states(B);
    std::vector<char> fin(K, 0);
    std::vector<int> status(K, PRX_OK);
    for (int k = 0; k < K; ++k) {
        CgState& h = B->st_host[k];
        if (!(rr[k] > thr[k]) || rr[k] == 0.0) {
            fin[k] = 1;
            h.done = 1;
            continue;
        }
        h.beta = 0.0;
        h.thr = std::max(theta * theta * h.delta, thr32[k]);
        h.done = 0;
        h.status = 0;
        h.rpend = 0;
        h.xn = 0;
    }
    write_states(B);
    if (persistent) {
        // every iteration, replacement and restart of every instance in one launch
        if (B->faces) throw Error(PRX_ERR_INVALID, "persistent CG: the stored-q fused step");
        PArgs<float> a = persist_args<float>(B, F);
        a.max_iters = (int64_t)max_iters;
        a.max_outer = (int64_t)max_outer;
        a.max_rounds = (int64_t)max_iters + 16;
        a.theta = theta;
        a.x = F.e32.p;
        a.r = F.r32.p;
        a.q = F.q32.p;
        a.ring = F.dd32.p;
        a.sbuf = F.sbuf32.p;
        a.x64 = F.x.p;
        a.x2 = F.x2.p;
        a.b = F.b.p;
        a.sbuf64 = F.sbuf.p;
        std::vector<PInst> pi(K);
        for (int k = 0; k < K; ++k) {
            pi[k] = PInst{};
            pi[k].thr = thr[k];
            pi[k].thr32 = thr32[k];
            pi[k].last_rt = INFINITY;
            pi[k].rr = rr[k];
            pi[k].fin = fin[k];
            pi[k].status = PRX_OK;
        }
        run_persistent<float>(B, F, a, pi);
        out.cg.iters.resize(K);
        out.cg.rel.resize(K);
        out.cg.status.resize(K);
        for (int k = 0; k < K; ++k) {
            out.outer[k] = (uint64_t)pi[k].outer;
            out.cg.iters[k] = (uint64_t)B->st_host[k].it;
            out.cg.status[k] = pi[k].fin ? pi[k].status : PRX_ERR_NOT_CONVERGED;
            out.true_rel[k] = bb[k] > 0.0 ? std::sqrt(pi[k].rr / bb[k]) : 0.0;
            const double den = mode == PRX_RESID_TRUE ? bb[k] : r0[k];
            out.cg.rel[k] = den > 0.0 ? std::sqrt(pi[k].rr / den) : 0.0;
        }
        PRX_HIP(hipStreamSynchronize(s));
        return out;
    }
    int chunk = B->chunk;
    chunk = std::max(CG_XM, (chunk + CG_XM / 2) / CG_XM * CG_XM);  // replays keep pass index = ring position
    const hipGraphExec_t g = graph_for(B, F, 0, chunk, 0, 1);
    for (;;) {
        for (;;) {
            bool all_done = true;
            for (int k = 0; k < K; ++k)
                if (!B->st_host[k].done) { all_done = false; break; }
            if (all_done) break;
            launch_graph(B, g);
            read_states(B);
        }
        // complete e: add the pending terms
        launch_bx_terms<float>(all, F.e32.p, XDirs<float>{}, F.dd32.p, N, 1, s);
        bool any = false;
        for (int k = 0; k < K; ++k) {
            B->st_host[k].xpend = 0;
            B->st_host[k].xn = 0;
            B->flags_h[k] = fin[k] ? 0 : 1;
            any = any || !fin[k];
        }
        write_states(B);
        if (!any) break;
        // fold e into x and replace each residual by the FP64 true one
        write_flags(B);
        const BatchTab fl = tab_flagged(B, F);
        resid_update_pass(B, F, fl);
        launch_bfin(fl, B->partial.p, ulog_tile(B->tile), FIN_SLOT, 1, s);
        launch_bcommit(fl, F.x.p, F.x2.p, F.e32.p, N, s);
        ul = launch_bdot<float>(fl, F.r32.p, F.r32.p, N, B->partial.p, s);
        launch_bfin(fl, B->partial.p, ul, FIN_SLOT, 2, s);
        read_states(B);
        for (int k = 0; k < K; ++k) {
            if (!B->flags_h[k]) continue;
            CgState& h = B->st_host[k];
            const double rt = h.slot[1], dt = h.slot[2];
            rr[k] = rt;
            out.outer[k]++;
            if (rt <= thr[k]) { fin[k] = 1; h.done = 1; status[k] = PRX_OK; continue; }
            if (h.status == PRX_ERR_BREAKDOWN) { fin[k] = 1; h.done = 1; status[k] = PRX_ERR_BREAKDOWN; continue; }
            if (h.status == PRX_ERR_NOT_CONVERGED || (uint64_t)h.it >= max_iters || out.outer[k] >= max_outer) {
                fin[k] = 1;
                h.done = 1;
                status[k] = PRX_ERR_NOT_CONVERGED;
                continue;
            }
            const bool drifted = !(dt <= 4.0 * h.delta);
            h.beta = (!drifted && h.delta_prev > 0.0) ? dt / h.delta_prev : 0.0;
            h.delta = dt;
            h.thr = std::max(theta * theta * dt, thr32[k]);
            h.done = 0;
            h.status = 0;
            h.rpend = 0;  // the replacement is the residual of the next pass
            // the graph restarts at pass 0, which reads r from slot 0 (just
            // written) and the direction from ring slot 0: move d_{it-1} there
            const int sl = (int)(((h.it - (int64_t)h.ring0) % CG_XM + CG_XM) % CG_XM);
            const int64_t nk = (int64_t)(F.off_h[k + 1] - F.off_h[k]);
            if (sl != 0)
                PRX_HIP(hipMemcpyAsync(F.dd32.p + F.off_h[k], F.dd32.p + (int64_t)sl * N + F.off_h[k], sizeof(float) * nk,
                                       hipMemcpyDeviceToDevice, s));
            h.ring0 = (int32_t)(h.it % CG_XM);
        }
        write_states(B);
        bool running = false;
        for (int k = 0; k < K; ++k) running = running || !fin[k];
        if (!running) break;
    }
    out.cg.iters.resize(K);
    out.cg.rel.resize(K);
    out.cg.status.resize(K);
    for (int k = 0; k < K; ++k) {
        out.cg.iters[k] = (uint64_t)B->st_host[k].it;
        out.cg.status[k] = status[k];
        out.true_rel[k] = bb[k] > 0.0 ? std::sqrt(rr[k] / bb[k]) : 0.0;
        const double den = mode == PRX_RESID_TRUE ? bb[k] : r0[k];
        out.cg.rel[k] = den > 0.0 ? std::sqrt(rr[k] / den) : 0.0;
    }
    PRX_HIP(hipStreamSynchronize(s));
    return out;
}

// ------------------------------------------------------------------ autotuner
static double time_ms(hipStream_t s, int reps, const std::function<void()>& f) {
    hipEvent_t a, b;
    PRX_HIP(hipEventCreate(&a));
    PRX_HIP(hipEventCreate(&b));
    f();  // warm-up
    PRX_HIP(hipEventRecord(a, s));
    for (int i = 0; i < reps; ++i) f();
    PRX_HIP(hipEventRecord(b, s));
    PRX_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    PRX_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return (double)ms / reps;
}

// Brick size of the operator (16^3 or 8^3 tiles), block size of the GS colour
// kernel and the CG chunk length, measured on this batch's fine level.  The
// chunk trades host synchronisations (one per replay) against skipped
// iterations after the last instance converges: with t_sync the round trip
// and t_skip one iteration of all-done launches, chunk ~ sqrt(2 I t_sync /
// (R t_skip)) for I expected iterations and R confirmation rounds.
static void autotune(prx_batch* B) {
    const auto t0 = std::chrono::steady_clock::now();
    BLevel& L = B->fine;
    const hipStream_t s = B->stream;
    const BatchTab all = tab_all(B, L);
    StencilIO<double> io;
    io.x = L.x.p;
    io.y = L.q.p;
    io.partial = B->partial.p;
    io.s = L.sbuf.p;
    double best = INFINITY;
    for (int tile : {16, 8}) {
        const double ms = time_ms(s, 3, [&] { launch_bstencil<double>(all, L.N, L.scr_ptr(), OP_APPLY, io, tile, s); });
        if (ms < best) { best = ms; B->tile = tile; }
    }
    // GS colour pass (tiled, in place) on a scratch copy of x: its brick edge
    PRX_HIP(hipMemcpyAsync(L.q.p, L.x.p, sizeof(double) * L.N, hipMemcpyDeviceToDevice, s));
    StencilIO<double> gio;
    gio.x = L.q.p;
    gio.y = L.q.p;
    gio.b = L.b.p;
    gio.s = L.sbuf.p;
    best = INFINITY;
    for (int gt : {16, 8}) {
        const double ms = time_ms(s, 3, [&] { launch_bstencil<double>(all, L.N, L.scr_ptr(), OP_GS, gio, gt, s); });
        if (ms < best) { best = ms; B->gs_block = gt; }
    }
    // host round trip and the cost of an iteration whose instances are all done
    const double t_sync = time_ms(s, 5, [&] {
        PRX_HIP(hipMemcpyAsync(B->st_host, B->st.p, sizeof(CgState) * L.K, hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
    });
    for (int k = 0; k < L.K; ++k) {
        B->st_host[k] = CgState{};
        B->st_host[k].done = 1;
    }
    write_states(B);
    B->chunk = 8;
    const hipGraphExec_t g = graph_for(B, L, 0, 8, 0);
    const double t_skip = time_ms(s, 3, [&] { PRX_HIP(hipGraphLaunch(g, s)); }) / 8.0;
    const double I = 500.0, R = 4.0;
    double c = std::sqrt(2.0 * I * t_sync / (R * std::max(t_skip, 1e-4)));
    int ci = (int)std::lround(c / 2.0) * 2;
    B->chunk = std::min(64, std::max(4, ci));
    PRX_HIP(hipStreamSynchronize(s));
    B->tuned = true;
    B->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------ the batched cascadic solve
// Multigrid-preconditioned CG (PRX_PRECOND_MG) on every instance.  Each
// instance is a complete assembled tree, solved by the single-tree path to
// ||b - A x|| <= eps ||b||; the instances are spread over up to 4 streams
// (one host thread each, so one instance's host round trips overlap the
// others' kernels), largest first onto the least-loaded stream.
static void solve_batched_mg(prx_batch* B, uint64_t cg_iters, double eps, const prx_solve_opts& o, prx_batch_report* rep,
                             prx_solve_report* per) {
    const int K = (int)B->inst.size();
    const int S = std::min(K, 4);
    for (prx_tree* t : B->inst)
        if (!t->assembled) throw Error(PRX_ERR_STATE, "batch instance not assembled");
    // the trees carry these solves: the batched levels' CG vectors (and the
    // graphs that hold their addresses) are not needed meanwhile
    B->drop_graphs();
    release_cg(B->fine);
    release_cg(B->coarse);
    release_cg32(B->fine);
    std::vector<int> order(K);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return B->inst[a]->N > B->inst[b]->N; });
    std::vector<std::vector<int>> work(S);
    std::vector<int64_t> load(S, 0);
    for (int k : order) {
        const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        work[s].push_back(k);
        load[s] += B->inst[k]->N;
    }
    std::vector<hipStream_t> streams(S, nullptr);
    for (auto& st : streams) PRX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    std::vector<prx_solve_report> reps(K);
    std::vector<int> rc(K, PRX_OK);
    std::vector<std::string> errs(S);
    PRX_HIP(hipStreamSynchronize(B->stream));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> threads;
    for (int s = 0; s < S; ++s)
        threads.emplace_back([&, s] {
            for (int k : work[s]) {
                prx_tree* t = B->inst[k];
                const hipStream_t keep = t->stream;
                t->stream = streams[s];
                prx_solve_opts oo = o;
                oo.precond = PRX_PRECOND_MG;
                oo.residual_mode = PRX_RESID_TRUE;
                rc[k] = prx_solve_cascadic(t, t->D > 2 ? t->D - 2 : 0, 20, cg_iters, eps, &oo, &reps[k]);
                if (rc[k] != PRX_OK) errs[s] = prx_last_error();
                (void)hipStreamSynchronize(streams[s]);
                t->stream = keep;
                if (rc[k] != PRX_OK) break;
            }
        });
    for (auto& th : threads) th.join();
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (auto& st : streams) (void)hipStreamDestroy(st);
    for (int k = 0; k < K; ++k)
        if (rc[k] != PRX_OK) {
            std::string msg;
            for (const auto& e : errs) if (!e.empty()) { msg = e; break; }
            throw Error(rc[k], "batch instance " + std::to_string(k) + ": " + msg);
        }
    prx_batch_report R{};
    R.instances = (uint32_t)K;
    R.status = PRX_OK;
    R.ms_total = ms;
    R.tile = B->tile;
    R.chunk = B->chunk;
    R.gs_block = B->gs_block;
    for (int k = 0; k < K; ++k) {
        const prx_solve_report& r = reps[k];
        R.total_nodes += (uint64_t)B->inst[k]->N;
        R.total_points += (uint64_t)B->inst[k]->P;
        R.max_true_rel = std::max(R.max_true_rel, r.true_rel);
        R.max_fine_iters = std::max<uint64_t>(R.max_fine_iters, r.fine_iters);
        R.sum_fine_iters += r.fine_iters;
        R.max_coarse_iters = std::max<uint64_t>(R.max_coarse_iters, r.coarse_iters);
        R.ms_coarse += r.ms_coarse;
        R.ms_fine_cg += r.ms_fine_cg;
        if (r.status == PRX_OK) R.converged++;
        else if (R.status == PRX_OK) R.status = r.status;
        if (per) per[k] = r;
    }
    if (rep) *rep = R;
}

static void solve_batched(prx_batch* B, uint32_t gs, uint64_t cg_iters, double eps, const prx_solve_opts& o, int tune,
                          prx_batch_report* rep, prx_solve_report* per) {
    if (B->inst.empty()) throw Error(PRX_ERR_STATE, "empty batch");
    if (o.precond == PRX_PRECOND_MG) {
        solve_batched_mg(B, cg_iters, eps, o, rep, per);
        return;
    }
    const bool mixed = o.precision == PRX_PREC_MIXED;
    if (o.precision != PRX_PREC_FP64 && !mixed) throw Error(PRX_ERR_INVALID, "precision must be PRX_PREC_FP64 or PRX_PREC_MIXED");
    if (mixed && o.cg_form == PRX_CG_TWO_PASS) throw Error(PRX_ERR_INVALID, "the batched mixed solve runs the fused CG step");
    // the instance trees' own solver workspace (left by multigrid or single
    // solves) is not used here; return it before the concatenated levels grow
    PRX_HIP(hipStreamSynchronize(B->stream));
    for (prx_tree* t : B->inst) {
        PRX_HIP(hipStreamSynchronize(t->stream));
        t->release_workspace();
    }
    build(B);
    B->fused = o.cg_form != PRX_CG_TWO_PASS;
    // measured on the batching workload (depth 6-8 instances), the face form is 3% slower than the stored-q
    // step, so batches use it only when asked for explicitly
    B->faces = o.cg_form == PRX_CG_FUSED_FACES;
    const hipStream_t s = B->stream;
    BLevel& F = B->fine;
    BLevel& C = B->coarse;
    const int K = F.K;
    // the coarse level runs the FP64 CG; the fine level the FP64 CG, or the
    // FP32 correction vectors of the mixed solve (the autotuner times FP64
    // passes, so a first mixed solve with tuning holds both for a moment).
    // Captured graphs hold vector addresses: any change drops them.
    auto fine_vectors = [&](bool want64, bool want32) {
        bool changed = false;
        if (want64 != (F.r.p != nullptr)) { want64 ? ensure_cg(F) : release_cg(F); changed = true; }
        if (want32 != (F.e32.p != nullptr)) { want32 ? ensure_cg32(F, B->stream) : release_cg32(F); changed = true; }
        if (changed) B->drop_graphs();
    };
    if (!C.r.p) {
        B->drop_graphs();
        ensure_cg(C);
    }
    fine_vectors(!mixed || (tune && !B->tuned), false);
    // right-hand sides: fine b from each tree, coarse RHS from its normal field
    for (int k = 0; k < K; ++k) {
        prx_tree* t = B->inst[k];
        if (!t->assembled) throw Error(PRX_ERR_STATE, "batch instance not assembled");
        PRX_HIP(hipMemcpyAsync(F.b.p + F.off_h[k], t->b.p, sizeof(double) * t->N, hipMemcpyDeviceToDevice, s));
    }
    if (tune && !B->tuned) {
        launch_fill<double>(F.x.p, 0.0, F.N, s);
        autotune(B);
    }
    if (!tune && !B->tuned) {
        B->tile = 16;
        B->gs_block = 16;
        B->chunk = 16;
    }
    if (mixed) fine_vectors(false, true);
    const int mode = o.residual_mode == PRX_RESID_TRUE ? PRX_RESID_TRUE : PRX_RESID_REFERENCE;
    const int refresh = mode == PRX_RESID_TRUE ? 0 : (o.refresh_period > 0 ? o.refresh_period : 0);
    // one persistent launch per level's CG: true-residual stopping with the
    // fused step (q stored); AUTO keeps the graph replays for now
    const bool pers_ok = mode == PRX_RESID_TRUE && B->fused && !B->faces;
    if (o.batch_launch == PRX_BATCH_LAUNCH_PERSISTENT && !pers_ok)
        throw Error(PRX_ERR_INVALID, "the persistent batched CG needs PRX_RESID_TRUE and the fused CG step with q stored");
    const bool persistent = o.batch_launch == PRX_BATCH_LAUNCH_PERSISTENT;
    for (double& v : B->persistent_phase_ms) v = 0.0;
    Timer total(s);
    double ms_coarse = 0.0, ms_gs = 0.0, ms_cg = 0.0;
    BOut co, fo;
    {
        Timer tm(s);
        for (int k = 0; k < K; ++k) {
            prx_tree* t = B->inst[k];
            const uint32_t cd = t->D - 2;
            const double hc = 1.0 / (double)(1u << cd);
            const double epw = t->pw * (double)((uint64_t)1 << (2 * (t->D - cd)));
            const double wh3c = epw > 0.0 ? epw * (hc * hc * hc) / 1.5 : 0.0;
            launch_coarse_rhs(t->D, cd, t->nf.p, epw > 0.0 ? wh3c * 0.5 : 0.0, C.b.p + C.off_h[k], &t->floor_list, s);
        }
        launch_fill<double>(C.x.p, 0.0, C.N, s);
        if (gs > 0) launch_graph(B, graph_for(B, C, 1, (int)gs, 0));
        const double ceps = mode == PRX_RESID_TRUE && o.coarse_eps > 0.0 ? o.coarse_eps : eps;
        co = batch_cg(B, C, cg_iters, ceps, mode, mode == PRX_RESID_TRUE ? 0 : 10, persistent);
        for (int k = 0; k < K; ++k) {
            prx_tree* t = B->inst[k];
            launch_prolongate(t->D, t->D - 2, C.x.p + C.off_h[k], F.x.p + F.off_h[k], s);
        }
        ms_coarse = tm.ms();
    }
    std::vector<uint64_t> outer(K, 1);
    if (mixed) {
        Timer tm(s);
        const double theta = o.inner_tol > 0.0 && o.inner_tol < 1.0 ? o.inner_tol : 0.05;
        MixedOut mo = batch_fine_mixed(B, F, (int)gs, cg_iters / 4, eps, mode, theta, o.max_outer > 0 ? o.max_outer : 64, persistent);
        ms_gs = mo.ms_gs;
        ms_cg = tm.ms() - mo.ms_gs;
        fo = mo.cg;
        outer = mo.outer;
    } else {
        {
            Timer tm(s);
            if (gs > 0) launch_graph(B, graph_for(B, F, 1, (int)gs, 0));
            ms_gs = tm.ms();
        }
        Timer tm(s);
        fo = batch_cg(B, F, cg_iters / 4, eps, mode, refresh, persistent);
        ms_cg = tm.ms();
    }
    const double ms_total = total.ms();
    // final true residuals and the solutions back into the trees
    const BatchTab all = tab_all(B, F);
    int ul = launch_bdot<double>(all, F.b.p, F.b.p, F.N, B->partial.p, s);
    launch_bfin(all, B->partial.p, ul, FIN_SLOT, 2, s);
    resid_pass(B, F, all, mixed ? F.x2.p : nullptr);
    launch_bfin(all, B->partial.p, ulog_tile(B->tile), FIN_SLOT, 3, s);
    read_states(B);
    for (int k = 0; k < K; ++k) {
        prx_tree* t = B->inst[k];
        PRX_HIP(hipMemcpyAsync(t->x.p, F.x.p + F.off_h[k], sizeof(double) * t->N, hipMemcpyDeviceToDevice, s));
    }
    PRX_HIP(hipStreamSynchronize(s));
    prx_batch_report R{};
    R.instances = (uint32_t)K;
    R.total_nodes = (uint64_t)F.N;
    for (prx_tree* t : B->inst) R.total_points += (uint64_t)t->P;
    R.ms_total = ms_total;
    R.ms_coarse = ms_coarse;
    R.ms_fine_gs = ms_gs;
    R.ms_fine_cg = ms_cg;
    R.ms_tune = B->tune_ms;
    R.tile = B->tile;
    R.chunk = B->chunk;
    R.gs_block = B->gs_block;
    R.graph_launches = B->graph_launches;
    R.persistent_launches = B->persistent_launches;
    R.persistent_rounds = B->persistent_rounds;
    R.persistent_blocks = B->persistent_blocks;
    R.persistent_xcd_match = B->persistent_xcd_match;
    for (int i = 0; i < 4; ++i) R.persistent_phase_ms[i] = B->persistent_phase_ms[i];
    R.status = PRX_OK;
    for (int k = 0; k < K; ++k) {
        const double bb = B->st_host[k].slot[2], rr = B->st_host[k].slot[3];
        const double tr = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
        R.max_true_rel = std::max(R.max_true_rel, tr);
        R.max_fine_iters = std::max<uint64_t>(R.max_fine_iters, fo.iters[k]);
        R.sum_fine_iters += fo.iters[k];
        R.max_coarse_iters = std::max<uint64_t>(R.max_coarse_iters, co.iters[k]);
        if (fo.status[k] == PRX_OK) R.converged++;
        else if (R.status == PRX_OK) R.status = fo.status[k];
        if (per) {
            prx_solve_report& p = per[k];
            std::memset(&p, 0, sizeof p);
            p.coarse_iters = co.iters[k];
            p.coarse_rel = co.rel[k];
            p.fine_iters = fo.iters[k];
            p.fine_rel = fo.rel[k];
            p.outer_iters = outer[k];
            p.true_rel = tr;
            p.status = fo.status[k];
            p.ms_total = ms_total;
        }
    }
    if (rep) *rep = R;
}

}  // namespace prx

// ====================================================================== C ABI
extern "C" {

int prx_batch_create(int device, prx_batch** out) {
    PRX_API_BEGIN
    PRX_REQUIRE(out, PRX_ERR_INVALID, "null pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw prx::Error(PRX_ERR_DEVICE, "no HIP device available");
    PRX_REQUIRE(device >= 0 && device < ndev, PRX_ERR_INVALID, "device ordinal out of range");
    std::unique_ptr<prx_batch> b(new prx_batch());
    b->device = device;
    PRX_HIP(hipSetDevice(device));
    PRX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    *out = b.release();
    PRX_API_END
}

void prx_batch_destroy(prx_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    delete b;
}

int prx_batch_add(prx_batch* b, uint32_t depth, const prx_point* pts, size_t n, double point_weight, uint32_t* index) {
    PRX_API_BEGIN
    PRX_REQUIRE(b && (pts || n == 0), PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(depth >= 6 && depth <= PRX_MAX_DEPTH, PRX_ERR_INVALID, "batched instances need depth 6..10 (base depth D-2 >= 4)");
    PRX_REQUIRE(point_weight >= 0.0 && std::isfinite(point_weight), PRX_ERR_INVALID, "point weight must be finite and >= 0");
    PRX_REQUIRE(prx::finite_points(pts, n), PRX_ERR_NONFINITE, "non-finite point coordinate");
    PRX_HIP(hipSetDevice(b->device));
    std::unique_ptr<prx_tree> t(prx::create_tree(depth, b->device, b->stream));
    prx::upload_points(t.get(), pts, n);
    prx::splat(t.get());
    prx::assemble(t.get(), point_weight);
    b->inst.push_back(t.release());
    b->built = false;
    b->tuned = false;
    if (index) *index = (uint32_t)(b->inst.size() - 1);
    PRX_API_END
}

int prx_batch_size(prx_batch* b, uint32_t* count) {
    PRX_API_BEGIN
    PRX_REQUIRE(b && count, PRX_ERR_INVALID, "null pointer");
    *count = (uint32_t)b->inst.size();
    PRX_API_END
}

int prx_batch_tree(prx_batch* b, uint32_t index, prx_tree** tree) {
    PRX_API_BEGIN
    PRX_REQUIRE(b && tree, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(index < b->inst.size(), PRX_ERR_INVALID, "instance index out of range");
    *tree = b->inst[index];
    PRX_API_END
}

int prx_solve_batched(prx_batch* b, uint32_t gs, uint64_t cg_iters, double eps, const prx_solve_opts* opts, int autotune,
                      prx_batch_report* report, prx_solve_report* per_instance) {
    PRX_API_BEGIN
    PRX_REQUIRE(b, PRX_ERR_INVALID, "null batch");
    PRX_REQUIRE(eps > 0.0 && eps < 1.0, PRX_ERR_INVALID, "eps must be in (0,1)");
    prx_solve_opts o;
    prx_solve_opts_default(&o);
    if (opts) o = *opts;
    PRX_HIP(hipSetDevice(b->device));
    prx::solve_batched(b, gs, cg_iters, eps, o, autotune, report, per_instance);
    PRX_API_END
}

}  // extern "C"
