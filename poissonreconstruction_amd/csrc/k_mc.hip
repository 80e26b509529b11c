// Marching-cubes surface extraction on the device: FEMTree::extract_surface
// (src/fem_tree.rs:576-636) over the complete tree's leaves.
//
// The reference walks the leaves in index (Morton) order, evaluates the field
// at each cell's 8 corners, interpolates a vertex on every cut edge of the
// cell and deduplicates vertices through a hash of the position quantised to
// 1e-9, first insertion keeping its index; triangles come from the case
// table in cell order.  Here a vertex belongs to a cut grid edge: it is
// computed once, with the orientation (a -> b) of the edge in the first cell
// (lowest Morton index) that contains it, which is the cell whose insertion
// the reference keeps, and it is numbered by (first cell, local edge), which
// is the reference's insertion order.  Triangles are written per active cell
// in cell order.  Corner values are the closed-form corner field
// (launch_evaluate_corners, bit-identical to evaluate()).
//
// Arithmetic follows the reference's expressions; contraction to FMA is off.
#pragma clang fp contract(off)

#include <hipcub/hipcub.hpp>

#include <vector>

#include "mc_cases.h"
#include "tree.hpp"

namespace prx {

__constant__ int8_t c_case[256][16];  // triangle edge lists, -1 terminated
__constant__ uint8_t c_corner[8][3];
__constant__ uint8_t c_edge[12][2];

struct McGrid {
    uint32_t D;
    uint64_t res, r1, ne;  // cells per axis, corners per axis, edges per direction (res * r1^2)
    const double* cv;      // corner values, row-major (kz * r1 + ky) * r1 + kx
    double iso, w;
    __host__ __device__ uint64_t corner(uint64_t i, uint64_t j, uint64_t k) const { return (k * r1 + j) * r1 + i; }
};

__device__ inline int cell_case(const McGrid& G, uint32_t i, uint32_t j, uint32_t k, double cv[8]) {
    int mc = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        cv[q] = G.cv[G.corner(i + c_corner[q][0], j + c_corner[q][1], k + c_corner[q][2])];
        if (cv[q] < G.iso) mc |= 1 << q;
    }
    return mc;
}

// cells with at least one triangle
struct CellActive {
    McGrid G;
    __device__ bool operator()(const uint32_t& c) const {
        uint32_t i, j, k;
        demorton3(c, i, j, k);
        double cv[8];
        const int mc = cell_case(G, i, j, k, cv);
        return c_case[mc][0] >= 0;
    }
};

// grid edge `e` of direction `dir`: low corner (i,j,k)
__device__ inline void edge_decode(const McGrid& G, uint64_t e, int& dir, uint64_t& i, uint64_t& j, uint64_t& k) {
    dir = (int)(e / G.ne);
    const uint64_t r = e - (uint64_t)dir * G.ne;
    if (dir == 0) { i = r % G.res; j = (r / G.res) % G.r1; k = r / (G.res * G.r1); }
    else if (dir == 1) { i = r % G.r1; j = (r / G.r1) % G.res; k = r / (G.r1 * G.res); }
    else { i = r % G.r1; j = (r / G.r1) % G.r1; k = r / (G.r1 * G.r1); }
}
__host__ __device__ inline uint64_t edge_encode(const McGrid& G, int dir, uint64_t i, uint64_t j, uint64_t k) {
    if (dir == 0) return (k * G.r1 + j) * G.res + i;
    if (dir == 1) return G.ne + (k * G.res + j) * G.r1 + i;
    return 2 * G.ne + (k * G.r1 + j) * G.r1 + i;
}

// cut edges (endpoints on different sides of the iso value)
struct EdgeCut {
    McGrid G;
    uint64_t base;  // first edge id of this batch
    __device__ bool operator()(const uint32_t& t) const {
        const uint64_t e = base + t;
        int dir;
        uint64_t i, j, k;
        edge_decode(G, e, dir, i, j, k);
        const double a = G.cv[G.corner(i, j, k)];
        const double b = G.cv[G.corner(i + (dir == 0), j + (dir == 1), k + (dir == 2))];
        return (a < G.iso) != (b < G.iso);
    }
};

// First cell (lowest Morton index) containing the cut edge and the edge's
// local index there.  x edge at (i,j,k): cells (i, j-dy, k-dz) with local
// edges 0, 2, 4, 6 for (dy,dz) = (0,0), (1,0), (0,1), (1,1); y edge: cells
// (i-dx, j, k-dz), local 3, 1, 7, 5 for (dx,dz); z edge: cells (i-dx, j-dy, k),
// local 8, 9, 11, 10 for (dx,dy) = (0,0), (1,0), (0,1), (1,1).
__device__ inline void first_cell(const McGrid& G, uint64_t e, uint32_t& cell, int& local) {
    int dir;
    uint64_t i, j, k;
    edge_decode(G, e, dir, i, j, k);
    const int lx[4] = {0, 2, 4, 6}, ly[4] = {3, 1, 7, 5}, lz[4] = {8, 9, 11, 10};
    uint32_t best = 0xffffffffu;
    int bl = -1;
    for (int q = 0; q < 4; ++q) {
        const int u = q & 1, v = q >> 1;
        int64_t ci = (int64_t)i, cj = (int64_t)j, ck = (int64_t)k;
        int le;
        if (dir == 0) { cj -= u; ck -= v; le = lx[q]; }
        else if (dir == 1) { ci -= u; ck -= v; le = ly[q]; }
        else { ci -= u; cj -= v; le = lz[q]; }
        if (ci < 0 || cj < 0 || ck < 0 || ci >= (int64_t)G.res || cj >= (int64_t)G.res || ck >= (int64_t)G.res) continue;
        const uint32_t m = morton3((uint32_t)ci, (uint32_t)cj, (uint32_t)ck);
        if (m < best) { best = m; bl = le; }
    }
    cell = best;
    local = bl;
}

__global__ void k_mc_vkeys(McGrid G, const uint32_t* __restrict__ edges, int64_t A, uint64_t* key, uint32_t* val) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A) return;
    uint32_t cell;
    int le;
    first_cell(G, edges[a], cell, le);
    key[a] = (uint64_t)cell * 12u + (uint64_t)le;
    val[a] = (uint32_t)a;
}

// vertex r: the edge's first cell, its local edge (a -> b), the reference's
// interpolation (src/fem_tree.rs:611-620)
__global__ void k_mc_vpos(McGrid G, const uint64_t* __restrict__ skey, const uint32_t* __restrict__ sval, int64_t A,
                          uint32_t* vid, double* pos) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A) return;
    const uint64_t kk = skey[r];
    const uint32_t cell = (uint32_t)(kk / 12u);
    const int le = (int)(kk % 12u);
    vid[sval[r]] = (uint32_t)r;
    uint32_t ci, cj, ck;
    demorton3(cell, ci, cj, ck);
    const uint64_t off[3] = {ci, cj, ck};
    const int a = c_edge[le][0], b = c_edge[le][1];
    const double va = G.cv[G.corner(ci + c_corner[a][0], cj + c_corner[a][1], ck + c_corner[a][2])];
    const double vb = G.cv[G.corner(ci + c_corner[b][0], cj + c_corner[b][1], ck + c_corner[b][2])];
    const double dv = vb - va;
    const double t = fabs(dv) < 1e-15 ? 0.5 : (G.iso - va) / dv;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double A0 = (double)(off[d] + c_corner[a][d]) * G.w;
        pos[3 * r + d] = A0 + t * (double)((int)c_corner[b][d] - (int)c_corner[a][d]) * G.w;
    }
}

__global__ void k_mc_ntri(McGrid G, const uint32_t* __restrict__ cells, int64_t C, uint32_t* ntri) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint32_t i, j, k;
    demorton3(cells[c], i, j, k);
    double cv[8];
    const int mc = cell_case(G, i, j, k, cv);
    uint32_t n = 0;
    while (n < 5 && c_case[mc][3 * n] >= 0) ++n;
    ntri[c] = n;
}

__device__ inline int64_t find_edge(const uint32_t* edges, int64_t A, uint64_t e) {
    int64_t lo = 0, hi = A;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((uint64_t)edges[mid] < e) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_mc_tris(McGrid G, const uint32_t* __restrict__ cells, int64_t C, const uint32_t* __restrict__ toff,
                          const uint32_t* __restrict__ edges, int64_t A, const uint32_t* __restrict__ vid, uint64_t* tris) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint32_t i, j, k;
    demorton3(cells[c], i, j, k);
    double cv[8];
    const int mc = cell_case(G, i, j, k, cv);
    uint64_t o = toff[c];
    for (int q = 0; q < 15 && c_case[mc][q] >= 0; ++q) {
        const int le = c_case[mc][q], a = c_edge[le][0], b = c_edge[le][1];
        int dir = 0;
        if (c_corner[a][1] != c_corner[b][1]) dir = 1;
        if (c_corner[a][2] != c_corner[b][2]) dir = 2;
        const uint64_t e = edge_encode(G, dir, i + min(c_corner[a][0], c_corner[b][0]), j + min(c_corner[a][1], c_corner[b][1]),
                                       k + min(c_corner[a][2], c_corner[b][2]));
        const int64_t s = find_edge(edges, A, e);
        tris[3 * o + (q % 3)] = (s < A && (uint64_t)edges[s] == e) ? (uint64_t)vid[s] : ~0ull;
        if (q % 3 == 2) ++o;
    }
}

__global__ void k_add_u32(uint32_t* p, int64_t n, uint32_t add) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] += add;
}
__global__ void k_set_u32_at(uint32_t* p, uint32_t v) { *p = v; }

static void upload_tables(hipStream_t s) {
    int8_t cs[256][16];
    for (int c = 0; c < 256; ++c) {
        int n = 0;
        for (const char* p = kMcCase[c]; *p; ++p) cs[c][n++] = (int8_t)(*p <= '9' ? *p - '0' : *p - 'a' + 10);
        while (n < 16) cs[c][n++] = -1;
    }
    PRX_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_case), cs, sizeof cs, 0, hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_corner), kMcCorner, sizeof kMcCorner, 0, hipMemcpyHostToDevice, s));
    PRX_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_edge), kMcEdge, sizeof kMcEdge, 0, hipMemcpyHostToDevice, s));
}

static unsigned nb(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

// Extract the iso surface of the tree's current solution: vertices [3 nv]
// (unit cube) and triangles [3 nt] into host vectors.
// transform_point (src/geometry.rs:256-261) of the vertices, unfused like the reference
struct M16 { double v[16]; };
__global__ void k_mc_to_model(double* pos, int64_t A, M16 M) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A) return;
    const double* m = M.v;
    const double x = pos[3 * r], y = pos[3 * r + 1], z = pos[3 * r + 2];
    pos[3 * r] = m[0] * x + m[4] * y + m[8] * z + m[12];
    pos[3 * r + 1] = m[1] * x + m[5] * y + m[9] * z + m[13];
    pos[3 * r + 2] = m[2] * x + m[6] * y + m[10] * z + m[14];
}

void extract_surface(prx_tree* t, double iso, std::vector<double>& verts, std::vector<uint64_t>& tris, const double* to_model) {
    const hipStream_t s = t->stream;
    upload_tables(s);
    McGrid G;
    G.D = t->D;
    G.res = t->res;
    G.r1 = G.res + 1;
    G.ne = G.res * G.r1 * G.r1;
    G.iso = iso;
    G.w = 1.0 / (double)G.res;
    DArr<double> cv;
    cv.ensure((int64_t)(G.r1 * G.r1 * G.r1));
    launch_evaluate_corners(t->D, t->x.p, cv.p, s);
    G.cv = cv.p;
    DArr<int64_t> nsel;
    nsel.ensure(4);
    std::vector<int64_t> nsel_h(4, 0);
    // active cells, in Morton order
    const int64_t N = t->N;
    DArr<uint32_t> cells;
    cells.ensure(N);
    size_t tb = 0;
    {
        CellActive pred{G};
        hipcub::CountingInputIterator<uint32_t> it(0);
        PRX_HIP(hipcub::DeviceSelect::If(nullptr, tb, it, cells.p, nsel.p, (int)N, pred, s));
        DArr<uint8_t> tmp;
        tmp.ensure((int64_t)tb);
        PRX_HIP(hipcub::DeviceSelect::If(tmp.p, tb, it, cells.p, nsel.p, (int)N, pred, s));
    }
    // cut edges in edge-id order, one direction at a time (each < 2^31 ids)
    const int64_t ne = (int64_t)G.ne;
    DArr<uint32_t> edges;
    int64_t A = 0;
    {
        // count first (edges of a closed surface ~ a few per active cell), then select into place
        std::vector<int64_t> cnt(3, 0);
        DArr<uint32_t> scratch_out;
        for (int dir = 0; dir < 3; ++dir) {
            EdgeCut pred{G, (uint64_t)dir * G.ne};
            hipcub::CountingInputIterator<uint32_t> it(0);
            size_t b2 = 0;
            scratch_out.ensure(std::max<int64_t>(ne, 1));
            PRX_HIP(hipcub::DeviceSelect::If(nullptr, b2, it, scratch_out.p, nsel.p + 1 + dir, (int)ne, pred, s));
            DArr<uint8_t> tmp;
            tmp.ensure((int64_t)b2);
            PRX_HIP(hipcub::DeviceSelect::If(tmp.p, b2, it, scratch_out.p, nsel.p + 1 + dir, (int)ne, pred, s));
            PRX_HIP(hipMemcpyAsync(&cnt[dir], nsel.p + 1 + dir, sizeof(int64_t), hipMemcpyDeviceToHost, s));
            PRX_HIP(hipStreamSynchronize(s));
            // append (ids of this direction offset by dir * ne)
            DArr<uint32_t> grown;
            grown.ensure(std::max<int64_t>(A + cnt[dir], 1));
            if (A > 0) PRX_HIP(hipMemcpyAsync(grown.p, edges.p, sizeof(uint32_t) * A, hipMemcpyDeviceToDevice, s));
            if (cnt[dir] > 0) {
                PRX_HIP(hipMemcpyAsync(grown.p + A, scratch_out.p, sizeof(uint32_t) * cnt[dir], hipMemcpyDeviceToDevice, s));
                if (dir > 0) {
                    const uint32_t add = (uint32_t)((uint64_t)dir * G.ne);
                    hipLaunchKernelGGL(k_add_u32, dim3(nb(cnt[dir])), dim3(256), 0, s, grown.p + A, cnt[dir], add);
                    PRX_LAUNCH_CHECK();
                }
            }
            PRX_HIP(hipStreamSynchronize(s));
            edges.swap(grown);
            A += cnt[dir];
        }
    }
    PRX_HIP(hipMemcpyAsync(nsel_h.data(), nsel.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PRX_HIP(hipStreamSynchronize(s));
    const int64_t C = nsel_h[0];
    // vertices: key (first cell, local edge) -> rank
    DArr<uint64_t> key, skey;
    DArr<uint32_t> val, sval, vid;
    DArr<double> pos;
    key.ensure(std::max<int64_t>(A, 1));
    skey.ensure(std::max<int64_t>(A, 1));
    val.ensure(std::max<int64_t>(A, 1));
    sval.ensure(std::max<int64_t>(A, 1));
    vid.ensure(std::max<int64_t>(A, 1));
    pos.ensure(std::max<int64_t>(3 * A, 1));
    if (A > 0) {
        hipLaunchKernelGGL(k_mc_vkeys, dim3(nb(A)), dim3(256), 0, s, G, edges.p, A, key.p, val.p);
        PRX_LAUNCH_CHECK();
        const int kbits = 3 * (int)t->D + 4;  // cell * 12 + 11 < 2^(3D+4)
        size_t b3 = 0;
        PRX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b3, key.p, skey.p, val.p, sval.p, (int)A, 0, kbits, s));
        DArr<uint8_t> tmp;
        tmp.ensure((int64_t)b3);
        PRX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, b3, key.p, skey.p, val.p, sval.p, (int)A, 0, kbits, s));
        hipLaunchKernelGGL(k_mc_vpos, dim3(nb(A)), dim3(256), 0, s, G, skey.p, sval.p, A, vid.p, pos.p);
        PRX_LAUNCH_CHECK();
        if (to_model) {
            M16 M;
            std::memcpy(M.v, to_model, sizeof M.v);
            hipLaunchKernelGGL(k_mc_to_model, dim3(nb(A)), dim3(256), 0, s, pos.p, A, M);
            PRX_LAUNCH_CHECK();
        }
    }
    // triangles: per active cell in cell order
    DArr<uint32_t> ntri, toff;
    ntri.ensure(std::max<int64_t>(C + 1, 1));
    toff.ensure(std::max<int64_t>(C + 1, 1));
    int64_t T = 0;
    if (C > 0) {
        hipLaunchKernelGGL(k_mc_ntri, dim3(nb(C)), dim3(256), 0, s, G, cells.p, C, ntri.p);
        PRX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_set_u32_at, dim3(1), dim3(1), 0, s, ntri.p + C, 0u);
        PRX_LAUNCH_CHECK();
        size_t b4 = 0;
        PRX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b4, ntri.p, toff.p, (int)(C + 1), s));
        DArr<uint8_t> tmp;
        tmp.ensure((int64_t)b4);
        PRX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, b4, ntri.p, toff.p, (int)(C + 1), s));
        uint32_t th = 0;
        PRX_HIP(hipMemcpyAsync(&th, toff.p + C, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
        T = th;
    }
    DArr<uint64_t> dtris;
    dtris.ensure(std::max<int64_t>(3 * T, 1));
    if (T > 0) {
        hipLaunchKernelGGL(k_mc_tris, dim3(nb(C)), dim3(256), 0, s, G, cells.p, C, toff.p, edges.p, A, vid.p, dtris.p);
        PRX_LAUNCH_CHECK();
    }
    verts.resize((size_t)(3 * A));
    tris.resize((size_t)(3 * T));
    if (A > 0) PRX_HIP(hipMemcpyAsync(verts.data(), pos.p, sizeof(double) * 3 * A, hipMemcpyDeviceToHost, s));
    if (T > 0) PRX_HIP(hipMemcpyAsync(tris.data(), dtris.p, sizeof(uint64_t) * 3 * T, hipMemcpyDeviceToHost, s));
    PRX_HIP(hipStreamSynchronize(s));
}

}  // namespace prx

extern "C" {

int prx_extract_surface(prx_tree* t, double iso, prx_mesh* out) {
    PRX_API_BEGIN
    PRX_REQUIRE(t && out, PRX_ERR_INVALID, "null pointer");
    PRX_REQUIRE(t->x.p, PRX_ERR_STATE, "no solution");
    std::memset(out, 0, sizeof *out);
    PRX_HIP(hipSetDevice(t->device));
    std::vector<double> v;
    std::vector<uint64_t> f;
    prx::extract_surface(t, iso, v, f);
    out->n_vertices = v.size() / 3;
    out->n_triangles = f.size() / 3;
    out->vertices = (double*)std::malloc(sizeof(double) * std::max<size_t>(v.size(), 1));
    out->triangles = (uint64_t*)std::malloc(sizeof(uint64_t) * std::max<size_t>(f.size(), 1));
    if (!out->vertices || !out->triangles) {
        std::free(out->vertices);
        std::free(out->triangles);
        std::memset(out, 0, sizeof *out);
        throw std::bad_alloc();
    }
    std::memcpy(out->vertices, v.data(), sizeof(double) * v.size());
    std::memcpy(out->triangles, f.data(), sizeof(uint64_t) * f.size());
    PRX_API_END
}

}  // extern "C"
