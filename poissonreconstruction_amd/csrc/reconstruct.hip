// The pipeline entry point (reconstruct, src/reconstructor.rs:25-70) and the
// generic CSR CG of the reference's solvers module on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "tree.hpp"

// The reference's arithmetic is unfused (Rust does not contract a*b+c); the
// device transform kernels below must round the same way.
#pragma clang fp contract(off)

namespace prx {

template <typename T>
struct DBuf {
    T* p = nullptr;
    explicit DBuf(int64_t n) { PRX_HIP(hipMalloc((void**)&p, (size_t)(n > 0 ? n : 1) * sizeof(T))); }
    ~DBuf() { if (p) (void)hipFree(p); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
};

// f64::min / f64::max ignore a NaN operand (src/geometry.rs:357-362).
__host__ __device__ inline double rs_min(double a, double b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }
__host__ __device__ inline double rs_max(double a, double b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }

// BBox + unit_cube_transform + inverse (src/geometry.rs:349-402, :274-316);
// m is column-major m[col*4+row] like Mat4.
static void transform_from_bbox(const double mn[3], const double mx[3], double scale, double m[16], double mi[16]);
void host_unit_transform(const prx_point* p, size_t n, double scale, double m[16], double mi[16]) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; ++i) {
        const double q[3] = {p[i].px, p[i].py, p[i].pz};
        for (int k = 0; k < 3; ++k) { mn[k] = rs_min(mn[k], q[k]); mx[k] = rs_max(mx[k], q[k]); }
    }
    transform_from_bbox(mn, mx, scale, m, mi);
}
// unit_cube_transform and Mat4::inverse (src/geometry.rs:384-402, :274-316)
static void transform_from_bbox(const double mn[3], const double mx[3], double scale, double m[16], double mi[16]) {
    double c[3], sz[3];
    for (int k = 0; k < 3; ++k) { c[k] = (mn[k] + mx[k]) * 0.5; sz[k] = mx[k] - mn[k]; }
    const double md = rs_max(rs_max(sz[0], sz[1]), sz[2]);
    const double s = md > 0.0 ? scale / md : 1.0;
    std::memset(m, 0, 16 * sizeof(double));
    m[0] = s; m[5] = s; m[10] = s;
    m[12] = -s * c[0] + 0.5; m[13] = -s * c[1] + 0.5; m[14] = -s * c[2] + 0.5; m[15] = 1.0;
    const double a = m[0], b = m[4], cc = m[8], d = m[1], e = m[5], f = m[9], g = m[2], h = m[6], ii = m[10];
    const double det = a * (e * ii - f * h) - b * (d * ii - f * g) + cc * (d * h - e * g);
    std::memset(mi, 0, 16 * sizeof(double));
    if (std::fabs(det) < 1e-30) { mi[0] = mi[5] = mi[10] = mi[15] = 1.0; return; }
    const double id = 1.0 / det;
    const double ai = (e * ii - f * h) * id, bi = (cc * h - b * ii) * id, ci = (b * f - cc * e) * id;
    const double di = (f * g - d * ii) * id, ei = (a * ii - cc * g) * id, fi = (cc * d - a * f) * id;
    const double gi = (d * h - e * g) * id, hi = (b * g - a * h) * id, iii = (a * e - b * d) * id;
    mi[0] = ai; mi[1] = di; mi[2] = gi;
    mi[4] = bi; mi[5] = ei; mi[6] = hi;
    mi[8] = ci; mi[9] = fi; mi[10] = iii;
    mi[12] = -(ai * m[12] + bi * m[13] + ci * m[14]);
    mi[13] = -(di * m[12] + ei * m[13] + fi * m[14]);
    mi[14] = -(gi * m[12] + hi * m[13] + iii * m[14]);
    mi[15] = 1.0;
}

// ------------------------------------------------------------------ the same on the device
// BBox::extend over all points (src/geometry.rs:356-363) as a block reduction;
// min/max ignoring NaN are associative, so the order does not matter.
__global__ __launch_bounds__(256) void k_bbox(const prx_point* __restrict__ p, int64_t n, double* part) {
    __shared__ double sm[6][256];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double q[3] = {p[i].px, p[i].py, p[i].pz};
        for (int k = 0; k < 3; ++k) { mn[k] = rs_min(mn[k], q[k]); mx[k] = rs_max(mx[k], q[k]); }
    }
    for (int k = 0; k < 3; ++k) { sm[k][threadIdx.x] = mn[k]; sm[3 + k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) {
                sm[k][threadIdx.x] = rs_min(sm[k][threadIdx.x], sm[k][threadIdx.x + w]);
                sm[3 + k][threadIdx.x] = rs_max(sm[3 + k][threadIdx.x], sm[3 + k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = sm[k][0];
}
struct Mat16 { double v[16]; };
// transform_point / transform_vector + normalize into SoA position/normal arrays
__global__ void k_transform_points(const prx_point* __restrict__ in, int64_t n, Mat16 M, double* pos, double* nrm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* m = M.v;
    const prx_point p = in[i];
    pos[i] = m[0] * p.px + m[4] * p.py + m[8] * p.pz + m[12];
    pos[n + i] = m[1] * p.px + m[5] * p.py + m[9] * p.pz + m[13];
    pos[2 * n + i] = m[2] * p.px + m[6] * p.py + m[10] * p.pz + m[14];
    double vx = m[0] * p.nx + m[4] * p.ny + m[8] * p.nz;
    double vy = m[1] * p.nx + m[5] * p.ny + m[9] * p.nz;
    double vz = m[2] * p.nx + m[6] * p.ny + m[10] * p.nz;
    const double nn = sqrt(vx * vx + vy * vy + vz * vz);
    if (nn > 0.0) { vx = vx / nn; vy = vy / nn; vz = vz / nn; }
    else { vx = vy = vz = 0.0; }
    nrm[i] = vx;
    nrm[n + i] = vy;
    nrm[2 * n + i] = vz;
}
// the model -> unit map of the points on the device: uploads the AoS points
// once, reduces the bounding box, builds m and its inverse on the host from
// the six numbers, transforms into pos/nrm (device SoA [3n] each)
void device_unit_transform(const prx_point* host_pts, size_t n, double scale, hipStream_t s, double m[16], double mi[16],
                           double* pos, double* nrm) {
    DBuf<prx_point> d((int64_t)n);
    PRX_HIP(hipMemcpyAsync(d.p, host_pts, sizeof(prx_point) * n, hipMemcpyHostToDevice, s));
    const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, ((int64_t)n + 255) / 256));
    DBuf<double> part(6 * nb);
    hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(256), 0, s, d.p, (int64_t)n, part.p);
    PRX_LAUNCH_CHECK();
    std::vector<double> ph(6 * (size_t)nb);
    PRX_HIP(hipMemcpyAsync(ph.data(), part.p, sizeof(double) * ph.size(), hipMemcpyDeviceToHost, s));
    PRX_HIP(hipStreamSynchronize(s));
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 3; ++k) { mn[k] = rs_min(mn[k], ph[6 * b + k]); mx[k] = rs_max(mx[k], ph[6 * b + 3 + k]); }
    transform_from_bbox(mn, mx, scale, m, mi);
    Mat16 M;
    std::memcpy(M.v, m, sizeof M.v);
    hipLaunchKernelGGL(k_transform_points, dim3((unsigned)std::max<int64_t>(1, ((int64_t)n + 255) / 256)), dim3(256), 0, s, d.p,
                       (int64_t)n, M, pos, nrm);
    PRX_LAUNCH_CHECK();
    PRX_HIP(hipStreamSynchronize(s));
}

// transform_point / transform_vector + normalize (src/geometry.rs:256-269,
// :49-56), as applied at src/reconstructor.rs:36-40.
void host_apply_transform(const prx_point* in, size_t n, const double m[16], prx_point* out) {
    for (size_t i = 0; i < n; ++i) {
        const prx_point p = in[i];
        prx_point o;
        o.px = m[0] * p.px + m[4] * p.py + m[8] * p.pz + m[12];
        o.py = m[1] * p.px + m[5] * p.py + m[9] * p.pz + m[13];
        o.pz = m[2] * p.px + m[6] * p.py + m[10] * p.pz + m[14];
        double vx = m[0] * p.nx + m[4] * p.ny + m[8] * p.nz;
        double vy = m[1] * p.nx + m[5] * p.ny + m[9] * p.nz;
        double vz = m[2] * p.nx + m[6] * p.ny + m[10] * p.nz;
        const double nn = std::sqrt(vx * vx + vy * vy + vz * vz);
        if (nn > 0.0) { vx = vx / nn; vy = vy / nn; vz = vz / nn; }
        else { vx = vy = vz = 0.0; }
        o.nx = vx; o.ny = vy; o.nz = vz;
        out[i] = o;
    }
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace prx

extern "C" {

int prx_reconstruct(const prx_point* points, size_t n, const prx_params* params, prx_mesh* out_mesh, prx_stats* out_stats) {
    prx_params p;
    if (params) p = *params;
    else prx_params_default(&p);
    prx_stats st;
    std::memset(&st, 0, sizeof st);
    if (out_mesh) std::memset(out_mesh, 0, sizeof *out_mesh);
    if (n == 0) {  // src/reconstructor.rs:28
        if (out_stats) *out_stats = st;
        return PRX_OK;
    }
    if (!points) return PRX_ERR_INVALID;
    for (size_t i = 0; i < n; ++i) {
        const prx_point& q = points[i];
        if (!std::isfinite(q.px) || !std::isfinite(q.py) || !std::isfinite(q.pz) || !std::isfinite(q.nx) || !std::isfinite(q.ny) ||
            !std::isfinite(q.nz))
            return PRX_ERR_NONFINITE;
    }
    if (n >= ((size_t)1 << 29)) { prx::g_err = "at most 2^29 points"; return PRX_ERR_INVALID; }
    const double t_total = prx::now_s();
    double t0 = prx::now_s();
    double m[16], mi[16];
    prx_tree* t = nullptr;
    int rc = prx_tree_create(p.depth, p.device, &t);
    if (rc != PRX_OK) return rc;
    // bounding box, unit-cube map and the transformed points on the device
    // (src/reconstructor.rs:31-40), written straight into the tree's arrays
    try {
        PRX_HIP(hipSetDevice(t->device));
        t->P = (int64_t)n;
        t->pos.ensure(3 * (int64_t)n);
        t->nrm.ensure(3 * (int64_t)n);
        prx::device_unit_transform(points, n, p.scale, t->stream, m, mi, t->pos.p, t->nrm.p);
    } catch (const prx::Error& e) {
        prx::g_err = e.what();
        prx_tree_destroy(t);
        return e.code;
    }
    st.time_transform = prx::now_s() - t0;

    t0 = prx::now_s();
    try {
        prx::splat(t);
    } catch (const prx::Error& e) {
        prx::g_err = e.what();
        prx_tree_destroy(t);
        return e.code;
    }
    st.fem_nodes = (uint64_t)prx_tree_node_count(t);
    st.time_tree = prx::now_s() - t0;

    t0 = prx::now_s();
    rc = prx_assemble(t, p.point_weight);
    if (rc != PRX_OK) { prx_tree_destroy(t); return rc; }
    const double t_sys = prx::now_s() - t0;
    int64_t nnz = 0;
    rc = prx_csr_nnz(t, &nnz);
    if (rc != PRX_OK) { prx_tree_destroy(t); return rc; }
    st.matrix_nnz = (uint64_t)nnz;
    const uint32_t base = p.base_depth >= 0 ? (uint32_t)p.base_depth : (p.depth > 2 ? p.depth - 2 : 0);
    prx_solve_opts o;
    prx_solve_opts_default(&o);
    o.precision = p.precision;
    o.residual_mode = p.residual_mode;
    o.refresh_period = p.refresh_period;
    o.gs_iters = p.gs_iters;
    o.inner_tol = p.inner_tol > 0.0 ? p.inner_tol : 0.05;
    o.precond = p.precond;
    o.max_outer = p.max_outer > 0 ? p.max_outer : 64;
    prx_solve_report rep;
    rc = prx_solve_cascadic(t, base, (uint32_t)p.gs_iters, p.cg_iters, p.cg_accuracy, &o, &rep);
    if (rc != PRX_OK) { prx_tree_destroy(t); return rc; }
    st.time_system = t_sys;
    st.time_solve = prx::now_s() - t0;

    t0 = prx::now_s();
    double iso = 0.0;
    rc = prx_iso_value(t, &iso);
    if (rc != PRX_OK) { prx_tree_destroy(t); return rc; }
    st.iso_value = iso;
    // extract_surface, vertices mapped back to model space on the device (src/reconstructor.rs:57-64)
    std::vector<double> uv;
    std::vector<uint64_t> tri;
    try {
        PRX_HIP(hipSetDevice(t->device));
        prx::extract_surface(t, iso, uv, tri, mi);
    } catch (const prx::Error& e) {
        prx::g_err = e.what();
        prx_tree_destroy(t);
        return e.code;
    } catch (const std::bad_alloc&) {
        prx::g_err = "host allocation failed";
        prx_tree_destroy(t);
        return PRX_ERR_OOM;
    }
    const size_t nv = uv.size() / 3, nt = tri.size() / 3;
    if (out_mesh) {
        out_mesh->vertices = (double*)std::malloc(sizeof(double) * std::max<size_t>(3 * nv, 1));
        out_mesh->triangles = (uint64_t*)std::malloc(sizeof(uint64_t) * std::max<size_t>(3 * nt, 1));
        if (!out_mesh->vertices || !out_mesh->triangles) {
            prx_mesh_free(out_mesh);
            prx_tree_destroy(t);
            prx::g_err = "host allocation failed";
            return PRX_ERR_OOM;
        }
        std::memcpy(out_mesh->vertices, uv.data(), sizeof(double) * 3 * nv);
        std::memcpy(out_mesh->triangles, tri.data(), sizeof(uint64_t) * 3 * nt);
        out_mesh->n_vertices = nv;
        out_mesh->n_triangles = nt;
    }
    st.mesh_vertices = nv;
    st.mesh_triangles = nt;
    st.time_extract = prx::now_s() - t0;
    prx_tree_destroy(t);

    st.coarse_iters = rep.coarse_iters;
    st.fine_iters = rep.fine_iters;
    st.outer_iters = rep.outer_iters;
    st.final_rel_residual = rep.true_rel;
    st.solve_ms = rep.ms_total;
    st.solve_status = rep.status;
    st.time_total = prx::now_s() - t_total;
    if (out_stats) *out_stats = st;
    return PRX_OK;
}

int prx_unit_transform_device(const prx_point* points, size_t n, double scale, int device, double* m, double* minv,
                              prx_point* out) {
    PRX_API_BEGIN
    PRX_REQUIRE(points || n == 0, PRX_ERR_INVALID, "null pointer");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw prx::Error(PRX_ERR_DEVICE, "no HIP device available");
    PRX_REQUIRE(device >= 0 && device < ndev, PRX_ERR_INVALID, "device ordinal out of range");
    PRX_HIP(hipSetDevice(device));
    hipStream_t s;
    PRX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    double mm[16], mi[16];
    try {
        prx::DBuf<double> pos(3 * (int64_t)n), nrm(3 * (int64_t)n);
        if (n > 0) prx::device_unit_transform(points, n, scale, s, mm, mi, pos.p, nrm.p);
        else prx::host_unit_transform(points, 0, scale, mm, mi);
        if (out && n > 0) {
            std::vector<double> hp(3 * n), hn(3 * n);
            PRX_HIP(hipMemcpyAsync(hp.data(), pos.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
            PRX_HIP(hipMemcpyAsync(hn.data(), nrm.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
            PRX_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < n; ++i)
                out[i] = prx_point{hp[i], hp[n + i], hp[2 * n + i], hn[i], hn[n + i], hn[2 * n + i]};
        }
    } catch (...) {
        (void)hipStreamDestroy(s);
        throw;
    }
    PRX_HIP(hipStreamDestroy(s));
    if (m) std::memcpy(m, mm, sizeof mm);
    if (minv) std::memcpy(minv, mi, sizeof mi);
    PRX_API_END
}

void prx_mesh_free(prx_mesh* mesh) {
    if (!mesh) return;
    std::free(mesh->vertices);
    std::free(mesh->triangles);
    std::memset(mesh, 0, sizeof *mesh);
}

// solve_cg (src/solvers.rs:20-69) on a caller CSR with the device kernels.
int prx_csr_cg(int64_t n, const int64_t* off, const int64_t* col, const double* val, const double* b, double* x,
               uint64_t max_iters, double eps, uint64_t* iters) {
    try {
        if (n < 0 || !off || !col || !val || !b || !x) return PRX_ERR_INVALID;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PRX_ERR_DEVICE;
        using namespace prx;
        const int64_t nnz = off[n];
        hipStream_t s;
        PRX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        DBuf<int64_t> doff(n + 1), dcol(nnz);
        DBuf<double> dval(nnz), db(n), dx(n), dr(n), dd(n), dq(n), part(4096);
        DBuf<CgState> st(1);
        PRX_HIP(hipMemcpyAsync(doff.p, off, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
        PRX_HIP(hipMemcpyAsync(dcol.p, col, sizeof(int64_t) * nnz, hipMemcpyHostToDevice, s));
        PRX_HIP(hipMemcpyAsync(dval.p, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
        PRX_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
        PRX_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
        CgState h{};
        h.eps = eps;
        h.max_iters = (int64_t)max_iters;
        PRX_HIP(hipMemcpyAsync(st.p, &h, sizeof h, hipMemcpyHostToDevice, s));
        int nb = 0;
        launch_csr_spmv(n, doff.p, dcol.p, dval.p, dx.p, dq.p, s);
        launch_sub<double>(dr.p, db.p, dq.p, n, part.p, &nb, nullptr, s);
        launch_fin(st.p, part.p, nb, FIN_INIT_REF, 0, s);
        PRX_HIP(hipMemcpyAsync(dd.p, dr.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        for (uint64_t it = 0; it < max_iters; ++it) {
            launch_csr_spmv(n, doff.p, dcol.p, dval.p, dd.p, dq.p, s);
            launch_dot<double>(dd.p, dq.p, n, part.p, &nb, s);
            launch_fin(st.p, part.p, nb, FIN_ALPHA, 0, s);
            if ((it + 1) % 10 == 0) {
                launch_axpy_alpha<double>(dx.p, nullptr, dd.p, st.p, n, s);
                launch_csr_spmv(n, doff.p, dcol.p, dval.p, dx.p, dq.p, s);
                launch_sub<double>(dr.p, db.p, dq.p, n, part.p, &nb, st.p, s);
            } else {
                launch_cg_update<double>(dx.p, dr.p, nullptr, dd.p, dq.p, st.p, n, part.p, &nb, s);
            }
            launch_fin(st.p, part.p, nb, FIN_BETA, 0, s);
            launch_xpby_beta<double>(dd.p, dr.p, st.p, n, s);
            if ((it & 15) == 15) {
                PRX_HIP(hipMemcpyAsync(&h, st.p, sizeof h, hipMemcpyDeviceToHost, s));
                PRX_HIP(hipStreamSynchronize(s));
                if (h.done) break;
            }
        }
        PRX_HIP(hipMemcpyAsync(&h, st.p, sizeof h, hipMemcpyDeviceToHost, s));
        PRX_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
        PRX_HIP(hipStreamDestroy(s));
        if (iters) *iters = h.done ? (uint64_t)h.iters : max_iters;
        return PRX_OK;
    } catch (const prx::Error& e) {
        return e.code;
    } catch (...) {
        return PRX_ERR_DEVICE;
    }
}

}  // extern "C"
