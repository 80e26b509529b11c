// Host-side state shared by the translation units of libprx: device arrays,
// the FEMTree handle (prx_tree) and the C-ABI error plumbing.
#pragma once
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace prx {

// ------------------------------------------------------------------ device arrays
template <typename T>
struct DArr {
    T* p = nullptr;
    int64_t n = 0;
    DArr() = default;
    DArr(const DArr&) = delete;
    DArr& operator=(const DArr&) = delete;
    ~DArr() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void swap(DArr& o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
    void ensure(int64_t cnt) {
        if (cnt <= n && p) return;
        release();
        PRX_HIP(hipMalloc((void**)&p, (size_t)(cnt > 0 ? cnt : 1) * sizeof(T)));
        n = cnt;
    }
};

inline void free_list(NodeList& L) {
    void* ptrs[] = {L.node, L.off, L.code, L.sp, L.spc, L.phi, L.phi32, L.diag, L.brick, L.brick8};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    int mode = L.mode;
    L = NodeList{};
    L.mode = mode;
}

// Reference stencil constants (src/fem_tree.rs:39-48), same expression order.
inline double mass_1d(int d) { d = d < 0 ? -d : d; return d == 0 ? 2. / 3. : (d == 1 ? 1. / 6. : 0.0); }
inline double stiff_1d(int d) { d = d < 0 ? -d : d; return d == 0 ? 2. : (d == 1 ? -1. : 0.0); }
inline double stiffness_entry(double h, int dx, int dy, int dz) {
    return (stiff_1d(dx) * mass_1d(dy) * mass_1d(dz) + mass_1d(dx) * stiff_1d(dy) * mass_1d(dz) +
            mass_1d(dx) * mass_1d(dy) * stiff_1d(dz)) * h;
}

struct Timer {
    hipEvent_t a{}, b{};
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) {
        PRX_HIP(hipEventCreate(&a));
        PRX_HIP(hipEventCreate(&b));
        PRX_HIP(hipEventRecord(a, s));
    }
    ~Timer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    double ms() {
        PRX_HIP(hipEventRecord(b, s));
        PRX_HIP(hipEventSynchronize(b));
        float t = 0.f;
        PRX_HIP(hipEventElapsedTime(&t, a, b));
        return (double)t;
    }
};

struct CgOut {
    uint64_t iters = 0;
    double rel = 0.0;
    int status = PRX_OK;
    uint64_t applies = 0;
};

}  // namespace prx

using namespace prx;

// ------------------------------------------------------------------ the handle
struct prx_tree {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    uint32_t D = 0, res = 1;
    int64_t N = 1;
    double h = 1.0;
    int tile = 16;

    // points
    int64_t P = 0;
    DArr<double> pos, nrm, t64;
    DArr<uint32_t> perm, cell;
    DArr<float> t32;
    DevPoints pts;
    NodeList floor_list, trunc_list, eval_list;
    bool splatted = false, assembled = false;

    // fields
    DArr<double> nf, w, b, x, x2, adj_g, adj_lam, adj_gnf, adj_out;
    double pw = 0.0;
    Coef<double> k64{};
    Coef<float> k32{};
    double K0 = 0, Ke = 0, Kc = 0, wh3 = 0, wg = 0;

    // solver workspace
    DArr<double> r64, d64, q64, s64, partial, tmp64;
    DArr<float> e32, r32, ri32, d32, q32, s32;
    // face-form CG: q at the touched nodes (node-indexed, zero elsewhere), the
    // screening sums per touched node and per cell corner, and the cell lists
    DArr<float> qt32;
    DArr<double> qt64;
    bool qt_ok = false;
    CellLists cells;
    bool cells_ok = false;
    CgState* st = nullptr;
    CgState* st_host = nullptr;
    CgState* st_snap = nullptr;  // [2] pinned: per-iteration state copies (PCG without host waits)
    CgState* state_snapshots() {
        if (!st_snap) PRX_HIP(hipHostMalloc((void**)&st_snap, 2 * sizeof(CgState), hipHostMallocDefault));
        return st_snap;
    }
    hipEvent_t snap_ev[2] = {nullptr, nullptr};  // behind each snapshot copy
    hipEvent_t* snapshot_events() {
        for (auto& e : snap_ev)
            if (!e) PRX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return snap_ev;
    }

    // coarse level (Morton order inside)
    uint32_t cd = 0;
    DArr<double> cb, cx, cr, cdd, cq, cpart;

    // multigrid preconditioner: level vectors (one pool per precision) and
    // the inverse of the coarsest operator
    DArr<double> mg64, mg_ainv;
    DArr<float> mg32;
    int precond = 0;  // solver of the last solve (the adjoint solve uses the same)
    // last prx_splat_normal_field, ms: host packing + upload, spatial point
    // sort, splat lists, screening lists, splat kernel
    double setup_ms[5] = {0, 0, 0, 0, 0};
    int cg_form = PRX_CG_FUSED;  // CG iteration form of the last solve (prx_solve_opts.cg_form)

    // in-solve timing of the finest level's fused CG pass (prx_pass_timing):
    // a hipEvent pair around each launch, read back after each chunk of
    // iterations (the chunk ends with a host wait anyway)
    struct PassTimer {
        bool on = false;
        std::vector<hipEvent_t> ev;  // 2 per launch in flight
        int used = 0;                // pairs recorded and not yet read
        std::vector<float> ms;       // durations read so far
    } ptime;

    prx_tree() { floor_list.mode = 0; trunc_list.mode = 1; eval_list.mode = 2; }
    ~prx_tree() {
        for (hipEvent_t e : ptime.ev) (void)hipEventDestroy(e);
        for (hipEvent_t e : snap_ev)
            if (e) (void)hipEventDestroy(e);
        free_list(floor_list);
        free_list(trunc_list);
        free_list(eval_list);
        free_cell_lists(cells);
        if (st) (void)hipFree(st);
        if (st_host) (void)hipHostFree(st_host);
        if (st_snap) (void)hipHostFree(st_snap);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }

    // The solver workspace (CG, mixed-precision and multigrid vectors, the
    // screening gathers): every solve allocates what it needs on first use,
    // so this only returns memory, e.g. before a batched solve that needs its
    // own concatenated vectors.  x, b and the point lists stay.
    void release_workspace() {
        r64.release(); d64.release(); q64.release(); s64.release(); tmp64.release();
        e32.release(); r32.release(); ri32.release(); d32.release(); q32.release(); s32.release();
        qt32.release(); qt64.release(); qt_ok = false;
        mg64.release(); mg32.release();
    }

    LevelView fine_view() const {
        LevelView v;
        v.D = D;
        v.res = res;
        v.N = N;
        v.k64 = k64;
        v.k32 = k32;
        v.scr = (pw > 0.0 && trunc_list.T > 0) ? &trunc_list : nullptr;
        v.pts = &pts;
        return v;
    }
    void ensure_state() {
        if (!st) PRX_HIP(hipMalloc((void**)&st, sizeof(CgState)));
        if (!st_host) PRX_HIP(hipHostMalloc((void**)&st_host, sizeof(CgState), hipHostMallocDefault));
        partial.ensure(partial_capacity(N));
    }
    void set_coefficients(double pweight) {
        pw = pweight;
        K0 = stiffness_entry(h, 0, 0, 0);
        Ke = stiffness_entry(h, 1, 1, 0);
        Kc = stiffness_entry(h, 1, 1, 1);
        wh3 = pw > 0.0 ? pw * (h * h * h) / 1.5 : 0.0;
        wg = pw > 0.0 ? wh3 * 0.005 : 0.0;
        k64 = {K0, Ke, Kc, wg, wh3};
        const double ih = 1.0 / h;
        k32 = {(float)(K0 * ih), (float)(Ke * ih), (float)(Kc * ih), (float)(wg * ih), (float)(wh3 * ih)};
    }
};

namespace prx {

extern thread_local std::string g_err;

void set_device(prx_tree* t);
void read_state(prx_tree* t);
// tree set-up as prx_tree_create / prx_splat_normal_field / prx_assemble do it;
// `shared` (optional) is a stream owned by the caller
prx_tree* create_tree(uint32_t depth, int device, hipStream_t shared = nullptr);
void upload_points(prx_tree* t, const prx_point* pts, size_t n);
void splat(prx_tree* t);
void assemble(prx_tree* t, double pw);
bool finite_points(const prx_point* p, size_t n);
// marching cubes of the current solution: vertices [3 nv] (unit cube, or mapped
// by the column-major 4x4 `to_model` when given), triangles [3 nt]
void extract_surface(prx_tree* t, double iso, std::vector<double>& verts, std::vector<uint64_t>& tris,
                     const double* to_model = nullptr);

}  // namespace prx

// ------------------------------------------------------------------ C-ABI error plumbing
#define PRX_API_BEGIN try {
#define PRX_API_END                                   \
    }                                                 \
    catch (const prx::Error& e) {                     \
        prx::g_err = e.what();                        \
        return e.code;                                \
    }                                                 \
    catch (const std::bad_alloc&) {                   \
        prx::g_err = "host allocation failed";        \
        return PRX_ERR_OOM;                           \
    }                                                 \
    catch (const std::exception& e) {                 \
        prx::g_err = e.what();                        \
        return PRX_ERR_DEVICE;                        \
    }                                                 \
    return PRX_OK;

#define PRX_REQUIRE(cond, code, msg) \
    if (!(cond)) throw prx::Error(code, msg)
