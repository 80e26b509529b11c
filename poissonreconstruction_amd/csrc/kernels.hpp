// Host-side launchers of the gfx950 kernels, grouped by translation unit.
#pragma once
#include "prx_internal.hpp"

namespace prx {

// ------------------------------------------------------------------ device state of one CG solve
// Scalars stay on the device so an iteration needs no host round-trip; every
// kernel of an iteration returns at once when `done` is set.
// With `lazy` set, the solution update x += alpha d is applied every second
// iteration for two directions at once (both are still in the direction
// ping-pong pair): the vector update then reads and writes x on half the
// iterations.  `xpend` marks a term alpha_pend * d[dir_buf] not yet in x;
// solvers flush it before they read x.  `dir_buf` is the ping-pong index of
// the latest direction (the FIN_BETA caller passes it as the slot).
// Fused CG step (OP_CG + FIN_CG): `rpend` marks the update r -= alpha q of
// the last iteration as still to be formed (by the next pass, from the r and
// q it wrote); `xn` counts the terms alpha_j d_j not yet added to x, with
// coefficients xa[0..xn) (oldest first).  x takes CG_XM of them at once, just
// before the pass that overwrites the oldest direction of the ring.  ring0:
// the iteration count (mod CG_XM) at which the solve last (re)started on ring
// slot 0, so direction j sits in slot (j - ring0 + 1) mod CG_XM.
constexpr int CG_XM = 8;
struct CgState {
    double delta, delta0, dq, alpha, beta, thr, eps, delta_prev;
    double bnorm2;
    double slot[8];
    int64_t it, iters, max_iters;
    int32_t done, status, refresh_due, lazy;
    double alpha_pend;
    int32_t xpend, dir_buf;
    double xa[CG_XM];
    int32_t rpend, xn, ring0;
};

// FIN_BETA_FULL: as FIN_BETA after an iteration that left x fully updated.
// Preconditioned CG (delta = r.z): FIN_PINIT starts from delta = r0.z0,
// FIN_PRES takes |r|^2 after the vector update (lazy-x bookkeeping, the
// iteration count and the stopping test on |r|^2 <= thr), FIN_PBETA takes the
// new r.z and forms beta.
// FIN_CG folds the four interleaved sums of an OP_CG pass (d.q, r.q, q.q,
// r.r): delta = r.r, alpha = delta / d.q, the next |r|^2 = delta - 2 alpha
// r.q + alpha^2 q.q and beta, in one step (slot 1: a refresh of r follows, so
// the stopping test waits for FIN_FRESH).  FIN_FRESH takes |b - A x|^2 of a
// refreshed residual.
enum FinOp {
    FIN_INIT_REF = 0, FIN_INIT_TRUE = 1, FIN_ALPHA = 2, FIN_BETA = 3, FIN_SLOT = 4, FIN_BETA_FULL = 5,
    FIN_PINIT = 6, FIN_PRES = 7, FIN_PBETA = 8, FIN_CG = 9, FIN_FRESH = 10
};

// Points on the device.  Input order: positions/normals (SoA, FP64).  Sorted
// order (by the Morton code of the screening cell): cell corner + weights.
struct DevPoints {
    int64_t P = 0;
    double* pos = nullptr;     // [3P] SoA px|py|pz, input order
    double* nrm = nullptr;     // [3P] SoA
    uint32_t* perm = nullptr;  // [P] sorted -> input
    uint32_t* cell = nullptr;  // [3P] SoA trunc cell as u32 (wrapping `as u32`, src/fem_tree.rs:212)
    double* t64 = nullptr;     // [3P] SoA fractional offsets s - trunc(s)
    float* t32 = nullptr;      // [3P]
};

// One grid level as the operator kernels see it.
struct LevelView {
    uint32_t D = 0, res = 1;
    int64_t N = 1;
    Coef<double> k64{};
    Coef<float> k32{};
    const NodeList* scr = nullptr;  // fine level with pw > 0 only
    const DevPoints* pts = nullptr;
};

// OP_JACOBI: y = x + jw * (b - A x), one damped-Jacobi step (jw = omega / a_ii,
// a_ii constant), with partial sums of b.y when io.partial is set.
// OP_JACOBI2: the first two such steps from x = 0 (weights jw, jw2) in one
// pass over b (the operand io.x = b): y = (jw + jw2) b - jw jw2 A b.
// OP_GS: one colour (io.color = coordinate parities x | y<<1 | z<<2) of the
// multicolour Gauss-Seidel sweep, in place: x_i += (b_i - (A x)_i) / a_ii for
// the nodes of that colour (io.x = io.y = x).
// OP_CG: one fused CG step (tile kernels only): r' = r - alpha q when
// rpend (else r' = r), d' = r' + beta dold, y = A d'; writes r' to rout, d'
// to dnew and y; four partials per brick, interleaved (d'.y, r'.y, y.y, r'.r').
// alpha, beta and rpend come from io.cst.  rout, dnew and y must not alias
// r, dold and q: neighbouring workgroups still read those as halo.
enum StencilOp { OP_APPLY = 0, OP_RESID = 1, OP_RESID_UPD = 2, OP_DIR_APPLY = 3, OP_JACOBI = 4, OP_JACOBI2 = 5, OP_GS = 6, OP_CG = 7,
                 OP_CG_FACES = 8 /* timing selector of launch_cg_faces */ };

// ---- k_grid.hip
void launch_export_offsets(uint32_t D, uint32_t* i2o, hipStream_t s);
void launch_export_g2i(uint32_t D, int64_t* g2i, hipStream_t s);
template <typename T> void launch_fill(T* p, T v, int64_t n, hipStream_t s);
template <typename T> void launch_dot(const T* a, const T* b, int64_t n, double* partial, int* nblk, hipStream_t s);
void launch_fin(CgState* st, const double* partial, int np, int op, int slot, hipStream_t s);
// dprev: the previous direction (lazy x); may be null when the state is not lazy
template <typename T> void launch_cg_update(T* x, T* r, const T* dprev, const T* d, const T* q, const CgState* st, int64_t n, double* partial, int* nblk, hipStream_t s);
template <typename T> void launch_axpy_alpha(T* x, const T* dprev, const T* d, const CgState* st, int64_t n, hipStream_t s);
// x += alpha_pend * d{dir_buf} when a lazy term is pending (the caller then clears xpend)
template <typename T> void launch_flush_x(T* x, const T* d0, const T* d1, const CgState* st, int64_t n, hipStream_t s);
// xpend = xn = 0, after the pending x terms have been added
void launch_clear_xpend(CgState* st, hipStream_t s);
// fused CG: x += sum_j xa[j] * d[j] over the pending terms (d[j] the
// direction of term j, oldest first).  launch_x_terms_full adds them only
// when all CG_XM are pending and the solve runs (before the pass that
// overwrites d[0]); launch_x_flush adds whatever is pending (if_running: only
// while the solve is not done).
template <typename T>
struct XDirs {
    const T* d[CG_XM];
};
template <typename T> void launch_x_terms_full(T* x, const XDirs<T>& d, const CgState* st, int64_t n, hipStream_t s);
template <typename T> void launch_x_flush(T* x, const XDirs<T>& d, const CgState* st, int64_t n, bool if_running, hipStream_t s);

// ---- k_mg.hip: geometric multigrid on the dense Morton grids
// x = w * b
template <typename T> void launch_mg_scale(T* x, const T* b, double w, int64_t n, hipStream_t s, const int32_t* done = nullptr);
// b_c = P^T r_f from depth dc + 1 to dc (trilinear hats, nodes outside the grid absent)
template <typename T> void launch_mg_restrict(uint32_t dc, const T* rf, T* bc, hipStream_t s, const int32_t* done = nullptr);
// x_f += P x_c from depth df - 1 to df
template <typename T> void launch_mg_prolong_add(uint32_t df, const T* xc, T* xf, hipStream_t s, const int32_t* done = nullptr);
// x = Ainv b on a grid of n <= 512 nodes (Ainv row-major, FP64)
template <typename T> void launch_mg_dense(int n, const double* ainv, const T* b, T* x, hipStream_t s, const int32_t* done = nullptr);
template <typename T> void launch_xpby_beta(T* d, const T* r, const CgState* st, int64_t n, hipStream_t s);
template <typename T> void launch_sub(T* r, const T* b, const T* q, int64_t n, double* partial, int* nblk, const CgState* st, hipStream_t s);
void launch_cast_scale_d2f(float* dst, const double* src, double scale, int64_t n, hipStream_t s);
void launch_add_f2d(double* x, const float* e, int64_t n, hipStream_t s);
void launch_coarse_rhs(uint32_t D, uint32_t d, const double* nf, double add, double* rhs_morton, const NodeList* splat, hipStream_t s);
void launch_prolongate(uint32_t D, uint32_t d, const double* coarse_morton, double* fine, hipStream_t s);
void launch_morton_to_rowmajor(uint32_t d, const double* src, double* dst, hipStream_t s);
void launch_rowmajor_to_morton(uint32_t d, const double* src, double* dst, hipStream_t s);
void launch_evaluate(uint32_t D, const double* sol, const double* q, int64_t m, double* out, hipStream_t s);
void launch_evaluate_pointsSoA(uint32_t D, const double* sol, const double* pos, int64_t P, double* out, hipStream_t s);
void launch_evaluate_corners(uint32_t D, const double* sol, double* out, hipStream_t s);
void launch_evaluate_corners_general(uint32_t D, const double* sol, double* out, hipStream_t s);
void launch_csr_spmv(int64_t n, const int64_t* off, const int64_t* col, const double* val, const double* x, double* y, hipStream_t s);

// ---- k_points.hip
// Per point (input order): node keys of the 8 corners for the floor (splat)
// and trunc (screening) cells; invalid corners get key N.
void launch_point_keys(uint32_t D, const double* pos, int64_t P, uint32_t* key_floor, uint32_t* key_trunc, uint32_t* val,
                       uint32_t* cell_key, uint32_t* cell_val, hipStream_t s);
void launch_point_keys_eval(uint32_t D, const double* pos, int64_t P, uint32_t* key, uint32_t* val, hipStream_t s);
void launch_point_sorted_data(uint32_t D, const double* pos, int64_t P, const uint32_t* perm, uint32_t* cell, double* t64,
                              float* t32, hipStream_t s);
// Run-length encode the sorted keys into a NodeList (node/off), fill code,
// phi and sp; allocates the NodeList arrays.
void build_node_list(NodeList& L, uint32_t D, const uint32_t* sorted_key, const uint32_t* sorted_val, int64_t M8,
                     const DevPoints& pts, const uint32_t* inv_perm, void* scratch, size_t scratch_bytes, hipStream_t s);
void launch_invert_perm(const uint32_t* perm, int64_t P, uint32_t* inv, hipStream_t s);
void launch_splat(uint32_t D, const NodeList& L, const DevPoints& pts, double* nf, double* w, hipStream_t s);
void launch_rhs_laplacian(uint32_t D, const double* nf, double* rhs, const NodeList* splat, hipStream_t s);
void launch_rhs_screening(uint32_t D, const NodeList& L, const DevPoints& pts, double wh3, double* rhs, hipStream_t s);
// s_p over the point's corners of x (x64 + e32 when x64 is given; r + beta*d
// when r is given).
// With cst (OP_CG): the operand r' + beta d with r' = r - alpha q when rpend.
template <typename T>
void launch_screen_gather(const LevelView& L, const T* x, const double* x64, const float* e32, T* sbuf, hipStream_t s,
                          const T* r = nullptr, const T* d = nullptr, const double* beta = nullptr, const T* q = nullptr,
                          const CgState* cst = nullptr);
// The occupied cells and each touched node's (cell, corner) pairs (built once
// per splat), and the face-form screening sums from them: sig[8 C] per cell
// corner, then sc[t] per touched node, for the OP_CG operand of r, d, q and
// the CG state; skipped when *done is set.
void build_cell_lists(CellLists& CL, uint32_t D, const DevPoints& pts, const NodeList& S, hipStream_t s);
void free_cell_lists(CellLists& CL);
template <typename T>
void launch_cell_screening(const LevelView& L, const CellLists& CL, const T* r, const T* d, const T* q, const CgState* cst,
                           const int32_t* done, T* sig, T* sc, hipStream_t s);

void launch_csr_count(uint32_t D, const NodeList* L, const DevPoints* pts, int has_diag, int64_t* rowcnt, hipStream_t s);
void launch_csr_fill(uint32_t D, const NodeList* L, const DevPoints* pts, int has_diag, double K0, double Ke, double Kc,
                     double wg, double wh3, const int64_t* off, int64_t* col, double* val, hipStream_t s);

// ---- k_stencil.hip
// One operator pass on level L.  OP_APPLY: y = A x, partial = x.y.
// OP_RESID: y = b - A x, partial = |y|^2.  OP_RESID_UPD (T = double):
// x_out = x64 + e32, r32 = (b - A x_out) * rscale, partial = |b - A x_out|^2.
// x_out must not alias x64: neighbouring workgroups still read x64 as halo.
// OP_DIR_APPLY: dnew = r + beta*dold (beta read on the device), y = A dnew,
// partial = dnew.y; dnew must not alias dold for the same reason.
template <typename T>
struct StencilIO {
    const T* x = nullptr;
    const T* b = nullptr;
    T* y = nullptr;
    const double* x64 = nullptr;
    double* x_out = nullptr;
    const float* e32 = nullptr;
    float* r32 = nullptr;
    double rscale = 1.0;
    const T* r = nullptr;
    const T* dold = nullptr;
    T* dnew = nullptr;
    const double* beta = nullptr;
    const T* s = nullptr;  // per-point screening values wh3*sum(phi*x)
    double* partial = nullptr;
    const int32_t* done = nullptr;
    double jw = 0.0;   // OP_JACOBI weight; OP_JACOBI2: the first step's
    double jw2 = 0.0;  // OP_JACOBI2: the second step's weight
    int color = 0;     // OP_GS colour
    const T* q = nullptr;  // OP_CG: the previous pass's A d (face form: its faces)
    const T* sc = nullptr;       // face form: the screening sums of d' per touched node
    T* qt = nullptr;             // face form: A d at the touched nodes (in: of d_old; out: of d'; node-indexed)
    T* rout = nullptr;     // OP_CG: the updated residual
    const CgState* cst = nullptr;  // OP_CG: alpha, beta, rpend
};
template <typename T>
void launch_stencil(const LevelView& L, int op, const StencilIO<T>& io, int* nblk, int tile, hipStream_t s);
// The fused CG step in face form (16^3 bricks, depth >= 4): as OP_CG, but q
// is not stored: io.q holds the previous pass's faces of A d (FACE_VALS per
// brick), io.y receives this pass's; io.qt (a node-indexed vector) holds
// A d_old at the touched nodes and receives A d' there, where the OP_CG-mode
// gather reads q; io.sc holds the screening sums of d' per touched node
// (launch_cell_screening).  Four partials per brick as OP_CG.
constexpr int FACE_VALS = 1536;
template <typename T>
void launch_cg_faces(const LevelView& L, const StencilIO<T>& io, int* nblk, hipStream_t s);
template <typename T>
void launch_gs_color(const LevelView& L, int color, T* x, const T* b, const T* sbuf, hipStream_t s);

int partial_capacity(int64_t N);

// ------------------------------------------------------------------ batches of grids
// K grids (one per point cloud, depth D_k >= 4) concatenated in one array:
// instance k owns the nodes [off[k], off[k+1]) in its own Morton order, and
// off[k] is a multiple of 4096, so every aligned 16^3 (or 8^3) brick lies in
// one instance (brick_inst, per 16^3 brick).  Node ids in batched node lists
// are global.  Each batched kernel processes instance k only while
// run[k * run_stride] == run_when (e.g. the instance's CG `done` flag is 0).
struct BCoef {
    uint32_t D, res;
    double s64, wg64, wh3_64;  // Laplacian weight scale s (ce = -2s, cc = -s, c0 = 32s), wg, wh3
    float s32, wg32, wh3_32;   // the same scaled by 1/h (FP32 correction system)
    float pad;
};
struct BatchTab {
    int K = 0;
    const uint64_t* off = nullptr;         // [K+1]
    const uint32_t* brick_inst = nullptr;  // [Ntot / 4096]
    const BCoef* coef = nullptr;           // [K]
    CgState* st = nullptr;                 // [K]
    const int32_t* run = nullptr;
    int run_stride = 0, run_when = 0;
};
// Batched point data for the screening gather (sorted order inside each
// instance, instances concatenated).
struct BPoints {
    int64_t P = 0;
    const uint32_t* cell = nullptr;  // [3P] SoA
    const double* t64 = nullptr;     // [3P]
    const float* t32 = nullptr;      // [3P]
    const uint32_t* inst = nullptr;  // [P]
};
// One batched operator pass (tile 16 or 8; beta and the run flag per
// instance from tab; partial per brick).
template <typename T>
void launch_bstencil(const BatchTab& tab, int64_t Ntot, const NodeList* scr, int op, const StencilIO<T>& io, int tile,
                     hipStream_t s);
// With q (fused CG step): the operand as the OP_CG pass forms it.
template <typename T>
void launch_bscreen_gather(const BatchTab& tab, const BPoints& pts, const T* x, const double* x64, const float* e32, T* sbuf,
                           const T* r, const T* d, hipStream_t s, const T* q = nullptr);
// The face form over a batch: the concatenated cell lists (cell indices and
// (cell, corner) pairs with global offsets, instance of a cell = instance of
// its first point) and the batched face-form pass (16^3 bricks).
template <typename T>
void launch_bcell_screening(const BatchTab& tab, const CellLists& CL, const BPoints& pts, int64_t T_touched, const uint32_t* node,
                            const T* r, const T* d, const T* q, T* sig, T* sc, hipStream_t s);
template <typename T>
void launch_bcg_faces(const BatchTab& tab, int64_t Ntot, const NodeList* scr, const StencilIO<T>& io, hipStream_t s);
// Per instance: fold its partials [off[k] >> ulog, off[k+1] >> ulog) in the
// order k_fin uses and apply the same scalar step to st[k].
void launch_bfin(const BatchTab& tab, const double* partial, int ulog, int op, int slot, hipStream_t s);
// FIN_CG per instance, folded in the single-tree order: instances with more
// than 256 bricks first fold their four sums per brick in up to FOLD_BLOCKS
// blocks into fold[(k * FOLD_BLOCKS + b) * 4 + j], then one block per
// instance folds those.
constexpr int FOLD_BLOCKS = 128;
void launch_bfin_cg(const BatchTab& tab, const double* partial, int ulog, double* fold, int slot, hipStream_t s);
// fused CG x terms per instance.  mode 0: all CG_XM terms before a pass
// (directions d, oldest first); mode 2: whatever is pending, while running
// (directions d); mode 1: whatever is pending, each instance finding its
// directions in the ring (slot i at ring + i * Ntot) from its it, ring0, xn.
template <typename T>
void launch_bx_terms(const BatchTab& tab, T* x, const XDirs<T>& d, const T* ring, int64_t Ntot, int mode, hipStream_t s);
// x += alpha_k d ; r -= alpha_k q ; partial per (256 * 16 / sizeof(T)) nodes
template <typename T>
int launch_bcg_update(const BatchTab& tab, T* x, T* r, const T* dprev, const T* d, const T* q, int64_t Ntot, double* partial,
                      hipStream_t s);
template <typename T>
void launch_baxpy_alpha(const BatchTab& tab, T* x, const T* dprev, const T* d, int64_t Ntot, hipStream_t s);
// per instance: x += alpha_pend * d{dir_buf} where a lazy term is pending
template <typename T>
void launch_bflush_x(const BatchTab& tab, T* x, const T* d0, const T* d1, int64_t Ntot, hipStream_t s);
// x = x2, e = 0 on the running instances (after a batched OP_RESID_UPD pass into x2)
void launch_bcommit(const BatchTab& tab, double* x, const double* x2, float* e, int64_t Ntot, hipStream_t s);
// partial = a.b per (256 * 16 / sizeof(T)) nodes; returns log2 of that unit
template <typename T>
int launch_bdot(const BatchTab& tab, const T* a, const T* b, int64_t Ntot, double* partial, hipStream_t s);

// The operator passes' view of a level: grid size, weights (as stored in
// the level, or per instance from a BatchTab) and the screening lists.
template <typename T>
struct LevelArgs {
    uint32_t D, res;
    T ce, cc, wg, c0;  // c0 = -(12ce + 8cc) = 32s, the diagonal of L
    T s;               // weight scale: ce = -2s, cc = -s
    int has_screen;
    int brick_log;
    const uint32_t* brick;
    const uint32_t* node;
    const uint32_t* off;
    const uint32_t* sp;
    const T* phi;
    const double* diag;  // sum phi^2 per touched node
    T wh3;
    // when set, the tile kernels recompute each entry's weight from its point
    // (spc = 8*sorted point + corner, tt = the points' in-cell coordinates)
    const uint32_t* spc;
    const T* tt;
    int64_t P;
};

// ---- persist.hip: the CG of every instance of a batched level in one
// persistent launch (all workgroups co-resident: a cooperative launch).
// Rounds in lockstep: in round k every running instance makes fused CG pass k
// (the ping-pong slot k & 1, ring slot k % CG_XM); instances whose CG stopped
// first get their replacement (mixed: FP64 true residual) or confirmation
// (FP64: the true residual) in the same round, and the device decides, per
// instance, whether it continues.  Phases are separated by a grid barrier;
// the last workgroup to finish a round's fold builds the next round's work
// lists (running instances, instances to replace, instances whose x terms are
// due), so finished instances cost nothing.
constexpr int PBAR_GROUPS = 32;  // first level of the grid barrier: workgroup w arrives at group w % PBAR_GROUPS
struct PCtl {
    uint32_t bar_count, bar_gen;
    int32_t err;                 // 1: a barrier wait exceeded its bound, 2: the round cap was reached, 3: the deadline passed, 4: see done
    int32_t nrun, nrepl, nxt;    // list sizes for the next round
    uint32_t ticket;             // completed fold-phase instances (the last one builds the lists)
    int32_t pad;
    int64_t rounds;
    int64_t repl_rounds;         // rounds with a replacement (confirmation) phase
    uint64_t ticks[4];           // wall-clock ticks per phase kind (replacement, gather, pass, fold), block 0's view
    uint32_t xcd_count[8];       // workgroups per XCD (as its XCC_ID register reads)
    uint32_t xcd_rr;             // workgroups w found on XCD w % 8
    uint64_t pass_busy;          // sum over workgroups and rounds of the time from a pass phase's start to its last brick
    uint32_t group[PBAR_GROUPS][16];  // one 64-byte line per first-level counter
    // brick positions taken so far per (brick phase: replacement, CG pass;
    // round parity; XCD), one 128-byte line each
    uint32_t take[2][2][8][32];
    // bricks processed per (brick phase, round parity), checked against the
    // list after the phase (err 4: a brick phase missed or repeated bricks)
    uint32_t done[2][2][32];
    int64_t bad_round;
    uint32_t bad_got, bad_want;
};
struct PInst {
    double thr, thr32, last_rt, rr;  // FP64 threshold, FP32 one (mixed), last true residual, last |b - A x|^2
    int32_t fin, status, outer, pad;
};
// instances with prefix sums of their 16^3 bricks, screening points, blocks
// of 256 16-byte vectors and 1024-node blocks ([n+1] each)
struct PList {
    int32_t* inst = nullptr;
    uint32_t* gbrick = nullptr;  // [n] the instance's first 16^3 brick in the level
    uint32_t* brick = nullptr;
    uint32_t* pts = nullptr;
    uint32_t* vblk = nullptr;
    uint32_t* kblk = nullptr;
};
template <typename T>
struct PArgs {
    BatchTab tab;                // off, brick_inst, coef, st of the level
    int64_t N = 0;
    int64_t max_iters = 0, max_outer = 0, max_rounds = 0;
    uint64_t deadline = 0;       // wall-clock ticks the launch may run (set by the launcher)
    double theta = 0.0;
    LevelArgs<T> L{};            // screening lists for the tile pass (T weights)
    LevelArgs<double> L64{};     // the same with FP64 weights (the mixed solve's FP64 residual pass)
    int has_scr = 0;
    // sorted points, instance by instance (pt_off[K+1])
    int64_t P = 0;
    const uint32_t* cell = nullptr;
    const T* tt = nullptr;
    const double* t64 = nullptr;
    const uint32_t* pinst = nullptr;
    const uint32_t* pt_off = nullptr;
    // CG vectors of type T: x (FP32 correction e in the mixed solve), r[2N], q[2N], ring[CG_XM N], sbuf[P]
    T* x = nullptr;
    T* r = nullptr;
    T* q = nullptr;
    T* ring = nullptr;
    T* sbuf = nullptr;
    // FP64: the solution x64, x2 = x64 + e, b, per-point screening values
    double* x64 = nullptr;
    double* x2 = nullptr;
    const double* b = nullptr;
    double* sbuf64 = nullptr;
    double* partial = nullptr;   // per 16^3 brick (four in a CG pass)
    double* partial2 = nullptr;  // per 1024-node block
    int32_t* flag = nullptr;     // [K] 1: replace (confirm) in the next round
    PInst* pi = nullptr;
    PList run, repl, xt;
    PCtl* ctl = nullptr;
};
// Launches k_persist for the batched level (T float: the mixed solve's FP32
// correction CG with FP64 residual replacements; T double: the FP64 CG with
// true-residual confirmation).  *blocks receives the grid size used.
// scr: the level's screening list (null: none); dargs: device memory of at
// least sizeof(PArgs<T>) for the argument block.  Returns when the launch ends.
template <typename T>
void launch_persistent_cg(const PArgs<T>& a, const NodeList* scr, void* dargs, int* blocks, hipStream_t s);

// ---- k_adjoint.hip
// gx[node] = sum over the node's evaluate entries of gF[p] * phi (input order)
void launch_adj_load(const NodeList& E, const double* gF, double* gx, hipStream_t s);
// gnf[3t] = h^2 sum_d lam_{j-d} g(d) at the splat-list nodes j = F.node[t]
void launch_adj_gnf(uint32_t D, const NodeList& F, const double* lam, double* gnf, hipStream_t s);
// per point (input order): d L / d position and d L / d normal (AoS, unit cube)
void launch_adj_points(uint32_t D, const DevPoints& pts, const NodeList& F, const double* gF, const double* x,
                       const double* lam, const double* nf, const double* w, const double* gnf, double wh3, double* gpos,
                       double* gnrm, hipStream_t s);

}  // namespace prx
