// Point-driven kernels: per-node point lists, normal splatting, RHS assembly,
// the per-point screening gather of the operator, and the verification-mode
// CSR export.
//
// Splat, RHS and CSR values reproduce the reference bit for bit: every node
// sums its contributions sequentially in input point order, with the
// reference's expression order, and contraction to FMA is off in this file.
#pragma clang fp contract(off)

#include <hipcub/hipcub.hpp>

#include "cg_dev.hpp"

namespace prx {

static inline int nblk(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    return (int)(b < 1 ? 1 : b);
}

// Floor cell for the splat (src/fem_tree.rs:138-143): 64-bit cell, corner
// offsets compared as usize.  Trunc cell for screening (src/fem_tree.rs:207-213):
// `as isize` then `(c + d) as u32`.
__device__ inline int64_t cell_floor(double p, double h) { return sat_i64(floor(p / h)); }
__device__ inline int64_t cell_trunc(double p, double h) { return sat_i64(p / h); }

__global__ void k_point_keys(uint32_t D, const double* __restrict__ pos, int64_t P, uint32_t* key_f, uint32_t* key_t,
                             uint32_t* val, uint32_t* cell_key, uint32_t* cell_val) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const uint32_t res = 1u << D, N = 1u << (3 * D);
    const double h = 1.0 / (double)res;
    const double q[3] = {pos[p], pos[P + p], pos[2 * P + p]};
    int64_t cf[3], ct[3];
    for (int a = 0; a < 3; ++a) { cf[a] = cell_floor(q[a], h); ct[a] = cell_trunc(q[a], h); }
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        uint64_t fx = (uint64_t)wadd(cf[0], dx), fy = (uint64_t)wadd(cf[1], dy), fz = (uint64_t)wadd(cf[2], dz);
        key_f[8 * p + k] = (fx < res && fy < res && fz < res) ? morton3((uint32_t)fx, (uint32_t)fy, (uint32_t)fz) : N;
        uint32_t tx = (uint32_t)(uint64_t)wadd(ct[0], dx), ty = (uint32_t)(uint64_t)wadd(ct[1], dy), tz = (uint32_t)(uint64_t)wadd(ct[2], dz);
        key_t[8 * p + k] = (tx < res && ty < res && tz < res) ? morton3(tx, ty, tz) : N;
        val[8 * p + k] = (uint32_t)(8 * p + k);
    }
    uint32_t c[3];
    for (int a = 0; a < 3; ++a) c[a] = (uint32_t)(ct[a] < 0 ? 0 : (ct[a] >= (int64_t)res ? res - 1 : ct[a]));
    cell_key[p] = morton3(c[0], c[1], c[2]);
    cell_val[p] = (uint32_t)p;
}
void launch_point_keys(uint32_t D, const double* pos, int64_t P, uint32_t* key_f, uint32_t* key_t, uint32_t* val,
                       uint32_t* cell_key, uint32_t* cell_val, hipStream_t s) {
    if (P <= 0) return;
    hipLaunchKernelGGL(k_point_keys, dim3(nblk(P)), dim3(256), 0, s, D, pos, P, key_f, key_t, val, cell_key, cell_val);
    PRX_LAUNCH_CHECK();
}

__global__ void k_point_sorted(uint32_t D, const double* __restrict__ pos, int64_t P, const uint32_t* __restrict__ perm,
                               uint32_t* cell, double* t64, float* t32) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const double h = 1.0 / (double)(1u << D);
    const uint32_t p = perm[i];
    for (int a = 0; a < 3; ++a) {
        double q = pos[a * P + p];
        int64_t c = cell_trunc(q, h);
        double t = q / h - (double)c;
        cell[a * P + i] = (uint32_t)(uint64_t)c;
        t64[a * P + i] = t;
        t32[a * P + i] = (float)t;
    }
}
void launch_point_sorted_data(uint32_t D, const double* pos, int64_t P, const uint32_t* perm, uint32_t* cell, double* t64,
                              float* t32, hipStream_t s) {
    if (P <= 0) return;
    hipLaunchKernelGGL(k_point_sorted, dim3(nblk(P)), dim3(256), 0, s, D, pos, P, perm, cell, t64, t32);
    PRX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ node lists
__global__ void k_count_valid(const uint32_t* key, int64_t M8, uint32_t N, uint32_t* out_M) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M8) return;
    if (key[j] < N && (j + 1 == M8 || key[j + 1] >= N)) *out_M = (uint32_t)(j + 1);
}
__global__ void k_heads(const uint32_t* key, int64_t M, uint32_t* head) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    head[j] = (j == 0 || key[j] != key[j - 1]) ? 1u : 0u;
}
__global__ void k_scatter_heads(const uint32_t* key, const uint32_t* head, const uint32_t* pos, int64_t M, uint32_t* node,
                                uint32_t* off) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    if (head[j]) {
        node[pos[j]] = key[j];
        off[pos[j]] = (uint32_t)j;
    }
}
__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

// evaluate's corner offsets (src/fem_tree.rs:645-649): (clamp(f-1), clamp(f)), f = floor(s)
__device__ inline int64_t eval_off(double s, int64_t res, int which) {
    const int64_t f = sat_i64(floor(s));
    const int64_t o = which ? f : wadd(f, -1);
    return o < 0 ? 0 : (o > res - 1 ? res - 1 : o);
}
__device__ inline double bspline1_d(double t) {
    if (t < 0.0 || t > 2.0) return 0.0;
    if (t <= 1.0) return t;
    return 2.0 - t;
}

// weight of corner k of point p: (bx*by)*bz as in src/fem_tree.rs:155-158 /
// :215-218 for the splat (mode 0) and screening (mode 1) cells, and the
// evaluate weights N1(s-o) (src/fem_tree.rs:498-507) for mode 2
__device__ inline double corner_weight(uint32_t D, const double* pos, int64_t P, uint32_t p, int k, int mode) {
    const double h = 1.0 / (double)(1u << D);
    double b[3];
    for (int a = 0; a < 3; ++a) {
        double q = pos[a * P + p];
        if (mode == 2) {
            const double s = q / h;
            b[a] = bspline1_d(s - (double)eval_off(s, (int64_t)1 << D, (k >> a) & 1));
        } else {
            int64_t c = mode == 0 ? cell_floor(q, h) : cell_trunc(q, h);
            double t = q / h - (double)c;
            b[a] = ((k >> a) & 1) ? t : 1.0 - t;
        }
    }
    return b[0] * b[1] * b[2];
}

// node keys of the 8 evaluate corners of each point (always in range)
__global__ void k_point_keys_eval(uint32_t D, const double* __restrict__ pos, int64_t P, uint32_t* key, uint32_t* val) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int64_t res = (int64_t)1 << D;
    const double h = 1.0 / (double)res;
    int64_t o[3][2];
    for (int a = 0; a < 3; ++a) {
        const double s = pos[a * P + p] / h;
        o[a][0] = eval_off(s, res, 0);
        o[a][1] = eval_off(s, res, 1);
    }
    for (int k = 0; k < 8; ++k) {
        key[8 * p + k] = morton3((uint32_t)o[0][k & 1], (uint32_t)o[1][(k >> 1) & 1], (uint32_t)o[2][k >> 2]);
        val[8 * p + k] = (uint32_t)(8 * p + k);
    }
}

void launch_point_keys_eval(uint32_t D, const double* pos, int64_t P, uint32_t* key, uint32_t* val, hipStream_t s) {
    if (P <= 0) return;
    hipLaunchKernelGGL(k_point_keys_eval, dim3(nblk(P)), dim3(256), 0, s, D, pos, P, key, val);
    PRX_LAUNCH_CHECK();
}

__global__ void k_list_entries(uint32_t D, int mode, const uint32_t* code, int64_t M, const double* __restrict__ pos, int64_t P,
                               const uint32_t* __restrict__ inv, uint32_t* sp, uint32_t* spc, double* phi, float* phi32) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const uint32_t c = code[j], p = c >> 3;
    const int k = c & 7;
    double w = corner_weight(D, pos, P, p, k, mode);
    phi[j] = w;
    phi32[j] = (float)w;
    sp[j] = inv[p];
    spc[j] = (inv[p] << 3) | (uint32_t)k;
}
__global__ void k_list_diag(const uint32_t* off, int64_t T, const double* phi, double* diag) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    double s = 0.0;
    for (uint32_t j = off[t]; j < off[t + 1]; ++j) s += phi[j] * phi[j];
    diag[t] = s;
}
// brick[b] = first touched node with index >= b << log
__global__ void k_brick_ranges(const uint32_t* node, int64_t T, int64_t nb, int lg, uint32_t* brick) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nb) return;
    const uint64_t v = (uint64_t)b << lg;
    int64_t lo = 0, hi = T;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((uint64_t)node[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    brick[b] = (uint32_t)lo;
}
__global__ void k_invert_perm(const uint32_t* perm, int64_t P, uint32_t* inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    inv[perm[i]] = (uint32_t)i;
}

void launch_invert_perm(const uint32_t* perm, int64_t P, uint32_t* inv, hipStream_t s) {
    if (P <= 0) return;
    hipLaunchKernelGGL(k_invert_perm, dim3(nblk(P)), dim3(256), 0, s, perm, P, inv);
    PRX_LAUNCH_CHECK();
}

template <typename T>
static void dmalloc(T** p, int64_t n) {
    PRX_HIP(hipMalloc((void**)p, (size_t)(n > 0 ? n : 1) * sizeof(T)));
}

void build_node_list(NodeList& L, uint32_t D, const uint32_t* key, const uint32_t* val, int64_t M8, const DevPoints& pts,
                     const uint32_t* inv, void* scratch, size_t scratch_bytes, hipStream_t s) {
    const uint32_t N = 1u << (3 * D);
    uint32_t* dM;
    PRX_HIP(hipMallocAsync((void**)&dM, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, dM, 0u);
    if (M8 > 0) hipLaunchKernelGGL(k_count_valid, dim3(nblk(M8)), dim3(256), 0, s, key, M8, N, dM);
    PRX_LAUNCH_CHECK();
    uint32_t hM = 0;
    PRX_HIP(hipMemcpyAsync(&hM, dM, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PRX_HIP(hipStreamSynchronize(s));
    const int64_t M = hM;
    L.M = M;
    uint32_t *head, *pos;
    PRX_HIP(hipMallocAsync((void**)&head, sizeof(uint32_t) * (M + 1), s));
    PRX_HIP(hipMallocAsync((void**)&pos, sizeof(uint32_t) * (M + 1), s));
    int64_t T = 0;
    if (M > 0) {
        hipLaunchKernelGGL(k_heads, dim3(nblk(M)), dim3(256), 0, s, key, M, head);
        size_t need = 0;
        PRX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, head, pos, (int)M, s));
        if (need > scratch_bytes) throw Error(PRX_ERR_DEVICE, "scan scratch too small");
        PRX_HIP(hipcub::DeviceScan::ExclusiveSum(scratch, need, head, pos, (int)M, s));
        uint32_t lastpos = 0, lasthead = 0;
        PRX_HIP(hipMemcpyAsync(&lastpos, pos + M - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        PRX_HIP(hipMemcpyAsync(&lasthead, head + M - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        PRX_HIP(hipStreamSynchronize(s));
        T = (int64_t)lastpos + lasthead;
    }
    L.T = T;
    dmalloc(&L.node, T);
    dmalloc(&L.off, T + 1);
    dmalloc(&L.code, M);
    dmalloc(&L.sp, M);
    dmalloc(&L.spc, M);
    dmalloc(&L.phi, M);
    dmalloc(&L.phi32, M);
    dmalloc(&L.diag, T);
    if (M > 0) {
        hipLaunchKernelGGL(k_scatter_heads, dim3(nblk(M)), dim3(256), 0, s, key, head, pos, M, L.node, L.off);
        PRX_HIP(hipMemcpyAsync(L.code, val, sizeof(uint32_t) * M, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_list_entries, dim3(nblk(M)), dim3(256), 0, s, D, L.mode, L.code, M, pts.pos, pts.P, inv, L.sp,
                           L.spc, L.phi, L.phi32);
    }
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, L.off + T, (uint32_t)M);
    if (T > 0) hipLaunchKernelGGL(k_list_diag, dim3(nblk(T)), dim3(256), 0, s, L.off, T, L.phi, L.diag);
    L.brick_log = D >= 4 ? 12 : 3 * (int)D;
    const int64_t nb = (int64_t)N >> L.brick_log;
    dmalloc(&L.brick, nb + 1);
    hipLaunchKernelGGL(k_brick_ranges, dim3(nblk(nb + 1)), dim3(256), 0, s, L.node, T, nb, L.brick_log, L.brick);
    if (D >= 3) {
        const int64_t nb8 = (int64_t)N >> 9;
        dmalloc(&L.brick8, nb8 + 1);
        hipLaunchKernelGGL(k_brick_ranges, dim3(nblk(nb8 + 1)), dim3(256), 0, s, L.node, T, nb8, 9, L.brick8);
    }
    PRX_LAUNCH_CHECK();
    PRX_HIP(hipFreeAsync(head, s));
    PRX_HIP(hipFreeAsync(pos, s));
    PRX_HIP(hipFreeAsync(dM, s));
}

// ------------------------------------------------------------------ splat
// splat_normal_field (src/fem_tree.rs:130-178): per touched node, sum
// n*w and w over its points in input order, then nf = -nf * (1/w).
__global__ void k_splat(int64_t T, const uint32_t* __restrict__ node, const uint32_t* __restrict__ off,
                        const uint32_t* __restrict__ code, const double* __restrict__ phi, const double* __restrict__ nrm,
                        int64_t P, int64_t N, double* nf, double* w) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    double sx = 0.0, sy = 0.0, sz = 0.0, sw = 0.0;
    for (uint32_t j = off[t]; j < off[t + 1]; ++j) {
        const uint32_t p = code[j] >> 3;
        const double ww = phi[j];
        sx += nrm[p] * ww;
        sy += nrm[P + p] * ww;
        sz += nrm[2 * P + p] * ww;
        sw += ww;
    }
    if (sw > 0.0) {
        double inv = 1.0 / sw;
        sx = -sx * inv;
        sy = -sy * inv;
        sz = -sz * inv;
    }
    const uint32_t i = node[t];
    nf[i] = sx;
    nf[N + i] = sy;
    nf[2 * N + i] = sz;
    w[i] = sw;
}
void launch_splat(uint32_t D, const NodeList& L, const DevPoints& pts, double* nf, double* w, hipStream_t s) {
    const int64_t N = (int64_t)1 << (3 * D);
    launch_fill<double>(nf, 0.0, 3 * N, s);
    launch_fill<double>(w, 0.0, N, s);
    if (L.T <= 0) return;
    hipLaunchKernelGGL(k_splat, dim3(nblk(L.T)), dim3(256), 0, s, L.T, L.node, L.off, L.code, L.phi, pts.nrm, pts.P, N, nf, w);
    PRX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ RHS
__device__ inline double mass_1d(int d) { return d == 0 ? 2. / 3. : 1. / 6.; }
__device__ inline double deriv_mass_1d(int d) { return d == 0 ? 0. : (d == 1 ? -0.5 : 0.5); }

// assemble_rhs Laplacian part (src/fem_tree.rs:258-274): neighbours in
// dx, dy, dz loop order, h*h*(v.x G M M + v.y M G M + v.z M M G).
__global__ void k_rhs_laplacian(uint32_t D, const double* __restrict__ nf, double* rhs) {
    const uint32_t res = 1u << D;
    const uint64_t N = 1ull << (3 * D);
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double h = 1.0 / (double)res;
    uint32_t ix, iy, iz;
    demorton3((uint32_t)i, ix, iy, iz);
    double sum = 0.0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                int64_t jx = (int64_t)ix + dx, jy = (int64_t)iy + dy, jz = (int64_t)iz + dz;
                if (jx < 0 || jx >= res || jy < 0 || jy >= res || jz < 0 || jz >= res) continue;
                const uint32_t j = morton3((uint32_t)jx, (uint32_t)jy, (uint32_t)jz);
                const double vx = nf[j], vy = nf[N + j], vz = nf[2 * N + j];
                sum += h * h * (vx * deriv_mass_1d(dx) * mass_1d(dy) * mass_1d(dz) + vy * mass_1d(dx) * deriv_mass_1d(dy) * mass_1d(dz) +
                                vz * mass_1d(dx) * mass_1d(dy) * deriv_mass_1d(dz));
            }
    rhs[i] = sum;
}
// The same sums, bit for bit, from operands staged in LDS.  One workgroup
// per aligned 8^3 brick (a contiguous Morton range) loads the brick's three
// normal-field components and the one-node halo (zeros outside the grid).
//
// Exact restatement.  A term of the reference's sum is h*h*(a + b + c) with
// a = ((vx*G(dx))*M(dy))*M(dz), b = ((vy*M(dx))*G(dy))*M(dz),
// c = ((vz*M(dx))*M(dy))*G(dz).  G(d) is 0 or -+1/2 and h*h = 2^-2D, and
// scaling by a power of two commutes with rounding (no subnormal operands:
// the field components are averages of unit normals), so
//   a = -+1/2 fl(fl(vx*M(dy))*M(dz)),  b = -+1/2 fl(fl(vy*M(dx))*M(dz)),
//   c = -+1/2 fl(fl(vz*M(dx))*M(dy)),
// a part with G(0) = 0 is an exact zero that leaves the sum a + b + c
// unchanged, and so are the terms of out-of-grid neighbours (zero halo; the
// reference skips them; a running sum is never -0, so adding a zero changes
// nothing).  The node's sum is then 2^(-2D-1) times the sequential sum, in
// the same dx, dy, dz order, of the terms' nonzero parts: 2 multiplies per
// part instead of 4, no bounds tests, and, since the neighbour order is a
// sweep along z, each thread forms two z-adjacent nodes of one (x, y) column
// from the 4 planes they share (21 LDS reads per node instead of 81).
//
// The normal field is nonzero only at the splat list's nodes (k_splat writes
// nothing else), so a brick whose 27 neighbour bricks (itself included) hold
// no splat node reads only zeros: every part is +-0, the running sum stays
// +0.0 (the reference's sum starts at +0.0, and +0 + -0 = +0), and the brick
// is written as +0.0 without reading the field.  act8[b .. b+1] is the splat
// list's node range of 8^3 brick b (null: no skipping).  On the 1M-point
// torus 13.7% of the depth-9 bricks have such a neighbour (6.8% at depth 10).
__global__ __launch_bounds__(256) void k_rhs_laplacian_tile(uint32_t D, const double* __restrict__ nf, double* rhs,
                                                            const uint32_t* __restrict__ act8) {
    constexpr int E = 10, E2 = E * E, E3 = E * E * E;
    __shared__ double v[3][E3];
    const uint32_t res = 1u << D;
    const uint64_t N = 1ull << (3 * D);
    const int tid = threadIdx.x;
    const uint32_t bb = blockIdx.x;
    const uint64_t base = (uint64_t)bb << 9;
    uint32_t bx, by, bz;
    demorton3(bb, bx, by, bz);
    const int ox = (int)bx * 8, oy = (int)by * 8, oz = (int)bz * 8;
    typedef double d2 __attribute__((ext_vector_type(2)));
    // interior: thread t loads local Morton nodes 2t and 2t+1 (an x pair)
    const uint32_t m0 = 2u * (uint32_t)tid;
    if (act8) {
        int near = 0;
        if (tid < 27) {
            const int cx = (int)bx + tid % 3 - 1, cy = (int)by + (tid / 3) % 3 - 1, cz = (int)bz + tid / 9 - 1;
            const int nb = (int)(res >> 3);
            if (cx >= 0 && cy >= 0 && cz >= 0 && cx < nb && cy < nb && cz < nb) {
                const uint32_t b = morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
                near = act8[b + 1] > act8[b] ? 1 : 0;
            }
        }
        if (!__syncthreads_or(near)) {
            *(d2*)(rhs + base + m0) = d2{0.0, 0.0};
            return;
        }
    }
    {
        uint32_t lx, ly, lz;
        demorton3(m0, lx, ly, lz);
        const int li = ((int)(lz + 1) * E + (int)ly + 1) * E + (int)lx + 1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const d2 q = *(const d2*)(nf + (uint64_t)c * N + base + m0);
            v[c][li] = q[0];
            v[c][li + 1] = q[1];
        }
    }
    // halo shell (10^3 - 8^3 = 488 nodes), zero outside the grid: every
    // thread issues all of its loads before it writes any of them
    constexpr int NH = (E3 + 255) / 256;
    double hv[NH][3];
    int hl[NH];
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        const int i = tid + 256 * k;
        const int hx = i % E - 1, hy = (i / E) % E - 1, hz = i / E2 - 1;
        const bool shell = i < E3 && !(hx >= 0 && hx < 8 && hy >= 0 && hy < 8 && hz >= 0 && hz < 8);
        const int gx = ox + hx, gy = oy + hy, gz = oz + hz;
        const bool in = shell && gx >= 0 && gy >= 0 && gz >= 0 && gx < (int)res && gy < (int)res && gz < (int)res;
        const uint32_t j = in ? morton3((uint32_t)gx, (uint32_t)gy, (uint32_t)gz) : 0u;
        hl[k] = shell ? i : -1;
#pragma unroll
        for (int c = 0; c < 3; ++c) hv[k][c] = in ? nf[(uint64_t)c * N + j] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NH; ++k)
        if (hl[k] >= 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c][hl[k]] = hv[k][c];
        }
    __syncthreads();
    // thread -> column (cx, cy), nodes z0 and z0 + 1
    const int cx = tid & 7, cy = (tid >> 3) & 7, z0 = (tid >> 6) * 2;
    constexpr double M0 = 2. / 3., M1 = 1. / 6.;  // mass_1d(0), mass_1d(+-1)
    double S[2] = {0.0, 0.0};
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) {
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            // planes z0-1 .. z0+2 of the neighbour column (tile z index z0 .. z0+3)
            const int col = (cy + 1 + dy) * E + cx + 1 + dx;
            double X[4], Y[4], Z[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const int k = (z0 + p) * E2 + col;
                X[p] = dx != 0 ? v[0][k] : 0.0;
                Y[p] = dy != 0 ? v[1][k] : 0.0;
                Z[p] = v[2][k];
            }
            const double mx = dx != 0 ? M1 : M0, my = dy != 0 ? M1 : M0;
#pragma unroll
            for (int o = 0; o < 2; ++o) {
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) {
                    const int p = o + 1 + dz;
                    const double mz = dz != 0 ? M1 : M0;
                    // parts in a, b, c order; G(-1) = +1/2, G(+1) = -1/2
                    double t = 0.0;
                    bool have = false;
                    if (dx != 0) {
                        const double a = (X[p] * my) * mz;
                        t = dx < 0 ? a : -a;
                        have = true;
                    }
                    if (dy != 0) {
                        const double b = (Y[p] * mx) * mz;
                        const double bs = dy < 0 ? b : -b;
                        t = have ? t + bs : bs;
                        have = true;
                    }
                    if (dz != 0) {
                        const double c = (Z[p] * mx) * my;
                        const double cs = dz < 0 ? c : -c;
                        t = have ? t + cs : cs;
                        have = true;
                    }
                    if (have) S[o] = S[o] + t;
                }
            }
        }
    }
    const double scale = ldexp(1.0, -2 * (int)D - 1);
    __syncthreads();  // the tile is read; stage the outputs in v[0] for a coalesced store
    const uint32_t mc = spread3((uint32_t)cx) | (spread3((uint32_t)cy) << 1);
    v[0][mc | (spread3((uint32_t)z0) << 2)] = S[0] * scale;
    v[0][mc | (spread3((uint32_t)z0 + 1u) << 2)] = S[1] * scale;
    __syncthreads();
    d2 out;
    out[0] = v[0][m0];
    out[1] = v[0][m0 + 1];
    *(d2*)(rhs + base + m0) = out;
}

void launch_rhs_laplacian(uint32_t D, const double* nf, double* rhs, const NodeList* splat, hipStream_t s) {
    const int64_t N = (int64_t)1 << (3 * D);
    const uint32_t* act8 = splat && splat->brick8 ? splat->brick8 : nullptr;
    if (D >= 3) hipLaunchKernelGGL(k_rhs_laplacian_tile, dim3((unsigned)(N >> 9)), dim3(256), 0, s, D, nf, rhs, act8);
    else hipLaunchKernelGGL(k_rhs_laplacian, dim3(nblk(N)), dim3(256), 0, s, D, nf, rhs);
    PRX_LAUNCH_CHECK();
}

// Screening RHS (src/fem_tree.rs:277-297): rhs += wh3*0.5*bx*by*bz per point,
// in input order.
__global__ void k_rhs_screening(uint32_t D, int64_t T, const uint32_t* __restrict__ node, const uint32_t* __restrict__ off,
                                const uint32_t* __restrict__ code, const double* __restrict__ pos, int64_t P, double wh3,
                                double* rhs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const double h = 1.0 / (double)(1u << D);
    const double target = 0.5;
    const uint32_t i = node[t];
    double acc = rhs[i];
    for (uint32_t j = off[t]; j < off[t + 1]; ++j) {
        const uint32_t c = code[j], p = c >> 3;
        const int k = c & 7;
        double b[3];
        for (int a = 0; a < 3; ++a) {
            double q = pos[a * P + p];
            int64_t cc = cell_trunc(q, h);
            double tt = q / h - (double)cc;
            b[a] = ((k >> a) & 1) ? tt : 1. - tt;
        }
        acc += wh3 * target * b[0] * b[1] * b[2];
    }
    rhs[i] = acc;
}
void launch_rhs_screening(uint32_t D, const NodeList& L, const DevPoints& pts, double wh3, double* rhs, hipStream_t s) {
    if (L.T <= 0) return;
    hipLaunchKernelGGL(k_rhs_screening, dim3(nblk(L.T)), dim3(256), 0, s, D, L.T, L.node, L.off, L.code, pts.pos, pts.P, wh3, rhs);
    PRX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ screening gather
// s_p = wh3 * sum_k phi_k x(node_k) over the point's in-range trunc corners,
// per point in sorted order.  The operator then adds sum phi * s_p at every
// touched node: together they apply sum_p wh3 phi_p phi_p^T
// (src/fem_tree.rs:206-226).
template <typename T, int MODE>  // MODE 0: x ; MODE 1: x64 + e32 ; MODE 2: r + beta*d ; MODE 3: the OP_CG operand
__global__ void k_screen_gather(uint32_t D, int64_t P, const uint32_t* __restrict__ cell, const T* __restrict__ tt,
                                const T* __restrict__ x, const double* __restrict__ x64, const float* __restrict__ e32, T wh3,
                                T* s, const T* __restrict__ r, const T* __restrict__ dd, const double* beta,
                                const T* __restrict__ qq, const CgState* cst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const double bt = MODE == 2 ? *beta : MODE == 3 ? cst->beta : 0.0;
    const double al = MODE == 3 ? cst->alpha : 0.0;
    const int rp = MODE == 3 ? cst->rpend : 0;
    const uint32_t res = 1u << D;
    const uint32_t c[3] = {cell[i], cell[P + i], cell[2 * P + i]};
    const T t[3] = {tt[i], tt[P + i], tt[2 * P + i]};
    // all corners' loads first (a corner outside the grid reads node 0 with
    // weight 0, which adds an exact zero), then the sum in corner order
    uint32_t m[8];
    T w[8], v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t ox = c[0] + (k & 1), oy = c[1] + ((k >> 1) & 1), oz = c[2] + (k >> 2);
        const bool in = ox < res && oy < res && oz < res;
        const T bx = (k & 1) ? t[0] : T(1) - t[0];
        const T by = ((k >> 1) & 1) ? t[1] : T(1) - t[1];
        const T bz = (k >> 2) ? t[2] : T(1) - t[2];
        w[k] = in ? bx * by * bz : T(0);
        m[k] = in ? morton3(ox, oy, oz) : 0u;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (MODE == 0) v[k] = x[m[k]];
        else if (MODE == 1) v[k] = (T)(x64[m[k]] + (double)e32[m[k]]);
        else if (MODE == 2) v[k] = dir_form(bt, dd[m[k]], r[m[k]]);
        else v[k] = cg_form(bt, al, rp, dd[m[k]], r[m[k]], qq[m[k]]);  // q is read always and used when rp
    }
    T acc = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += w[k] * v[k];
    s[i] = wh3 * acc;
}
template <typename T>
void launch_screen_gather(const LevelView& L, const T* x, const double* x64, const float* e32, T* sbuf, hipStream_t s,
                          const T* r, const T* d, const double* beta, const T* q, const CgState* cst) {
    const DevPoints& pts = *L.pts;
    if (pts.P <= 0) return;
    const T* tt;
    T wh3;
    if constexpr (sizeof(T) == 8) { tt = (const T*)pts.t64; wh3 = (T)L.k64.wh3; }
    else { tt = (const T*)pts.t32; wh3 = (T)L.k32.wh3; }
    if (cst)
        hipLaunchKernelGGL((k_screen_gather<T, 3>), dim3(nblk(pts.P)), dim3(256), 0, s, L.D, pts.P, pts.cell, tt, x, x64, e32, wh3, sbuf,
                           r, d, beta, q, cst);
    else if (r)
        hipLaunchKernelGGL((k_screen_gather<T, 2>), dim3(nblk(pts.P)), dim3(256), 0, s, L.D, pts.P, pts.cell, tt, x, x64, e32, wh3, sbuf,
                           r, d, beta, q, cst);
    else if (x64)
        hipLaunchKernelGGL((k_screen_gather<T, 1>), dim3(nblk(pts.P)), dim3(256), 0, s, L.D, pts.P, pts.cell, tt, x, x64, e32, wh3, sbuf,
                           r, d, beta, q, cst);
    else
        hipLaunchKernelGGL((k_screen_gather<T, 0>), dim3(nblk(pts.P)), dim3(256), 0, s, L.D, pts.P, pts.cell, tt, x, x64, e32, wh3, sbuf,
                           r, d, beta, q, cst);
    PRX_LAUNCH_CHECK();
}
template void launch_screen_gather<double>(const LevelView&, const double*, const double*, const float*, double*, hipStream_t,
                                           const double*, const double*, const double*, const double*, const CgState*);

// ------------------------------------------------------------------ occupied cells (face-form screening)
// A cell starts where the sorted points' trunc cell changes.
__global__ void k_cell_heads(int64_t P, const uint32_t* __restrict__ cell, uint32_t* head) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    head[i] = (i == 0 || cell[i] != cell[i - 1] || cell[P + i] != cell[P + i - 1] || cell[2 * P + i] != cell[2 * P + i - 1]) ? 1u : 0u;
}
__global__ void k_cell_starts(int64_t P, const uint32_t* __restrict__ head, const uint32_t* __restrict__ pos, uint32_t* cstart) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    if (head[i]) cstart[pos[i]] = (uint32_t)i;
}
// key = corner node of cell c (N outside the grid), val = 8 c + corner
__global__ void k_cell_corner_keys(uint32_t D, int64_t C, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cell,
                                   int64_t P, uint32_t* key, uint32_t* val) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t res = 1u << D, N = 1u << (3 * D), p = cstart[c];
    const uint32_t cx = cell[p], cy = cell[P + p], cz = cell[2 * P + p];
    for (int k = 0; k < 8; ++k) {
        const uint32_t ox = cx + (k & 1), oy = cy + ((k >> 1) & 1), oz = cz + (k >> 2);
        key[8 * c + k] = (ox < res && oy < res && oz < res) ? morton3(ox, oy, oz) : N;
        val[8 * c + k] = (uint32_t)(8 * c + k);
    }
}
// coff[t] = the first sorted pair of node[t]; coff[T] = the number of pairs inside the grid
__global__ void k_cell_offsets(int64_t T, const uint32_t* __restrict__ node, const uint32_t* __restrict__ skey, int64_t M,
                               uint32_t N, uint32_t* coff) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > T) return;
    const uint32_t v = t < T ? node[t] : N;
    int64_t lo = 0, hi = M;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (skey[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    coff[t] = (uint32_t)lo;
}
void build_cell_lists(CellLists& CL, uint32_t D, const DevPoints& pts, const NodeList& S, hipStream_t s) {
    const int64_t P = pts.P;
    const uint32_t N = 1u << (3 * D);
    CL.C = 0;
    PRX_HIP(hipMalloc((void**)&CL.coff, sizeof(uint32_t) * (size_t)(S.T + 1)));
    if (P <= 0) {
        PRX_HIP(hipMalloc((void**)&CL.cstart, sizeof(uint32_t)));
        PRX_HIP(hipMalloc((void**)&CL.cval, sizeof(uint32_t)));
        PRX_HIP(hipMemsetAsync(CL.coff, 0, sizeof(uint32_t) * (size_t)(S.T + 1), s));
        return;
    }
    uint32_t *head, *pos;
    PRX_HIP(hipMallocAsync((void**)&head, sizeof(uint32_t) * P, s));
    PRX_HIP(hipMallocAsync((void**)&pos, sizeof(uint32_t) * P, s));
    hipLaunchKernelGGL(k_cell_heads, dim3(nblk(P)), dim3(256), 0, s, P, pts.cell, head);
    size_t need = 0;
    PRX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, head, pos, (int)P, s));
    void* scratch;
    PRX_HIP(hipMallocAsync(&scratch, need, s));
    PRX_HIP(hipcub::DeviceScan::ExclusiveSum(scratch, need, head, pos, (int)P, s));
    uint32_t lastpos = 0, lasthead = 0;
    PRX_HIP(hipMemcpyAsync(&lastpos, pos + P - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PRX_HIP(hipMemcpyAsync(&lasthead, head + P - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PRX_HIP(hipStreamSynchronize(s));
    PRX_HIP(hipFreeAsync(scratch, s));
    const int64_t C = (int64_t)lastpos + lasthead;
    CL.C = C;
    PRX_HIP(hipMalloc((void**)&CL.cstart, sizeof(uint32_t) * (size_t)(C + 1)));
    hipLaunchKernelGGL(k_cell_starts, dim3(nblk(P)), dim3(256), 0, s, P, head, pos, CL.cstart);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, CL.cstart + C, (uint32_t)P);
    uint32_t *key, *val, *skey;
    PRX_HIP(hipMallocAsync((void**)&key, sizeof(uint32_t) * 8 * C, s));
    PRX_HIP(hipMallocAsync((void**)&val, sizeof(uint32_t) * 8 * C, s));
    PRX_HIP(hipMallocAsync((void**)&skey, sizeof(uint32_t) * 8 * C, s));
    PRX_HIP(hipMalloc((void**)&CL.cval, sizeof(uint32_t) * 8 * (size_t)C));
    hipLaunchKernelGGL(k_cell_corner_keys, dim3(nblk(C)), dim3(256), 0, s, D, C, CL.cstart, pts.cell, P, key, val);
    const int kbits = 3 * (int)D + 1;
    need = 0;
    PRX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, key, skey, val, CL.cval, (int)(8 * C), 0, kbits, s));
    PRX_HIP(hipMallocAsync(&scratch, need, s));
    // stable: the pairs of one node stay in cell order
    PRX_HIP(hipcub::DeviceRadixSort::SortPairs(scratch, need, key, skey, val, CL.cval, (int)(8 * C), 0, kbits, s));
    hipLaunchKernelGGL(k_cell_offsets, dim3(nblk(S.T + 1)), dim3(256), 0, s, S.T, S.node, skey, 8 * C, N, CL.coff);
    PRX_LAUNCH_CHECK();
    PRX_HIP(hipFreeAsync(scratch, s));
    PRX_HIP(hipFreeAsync(key, s));
    PRX_HIP(hipFreeAsync(val, s));
    PRX_HIP(hipFreeAsync(skey, s));
    PRX_HIP(hipFreeAsync(head, s));
    PRX_HIP(hipFreeAsync(pos, s));
    PRX_HIP(hipStreamSynchronize(s));
}
void free_cell_lists(CellLists& CL) {
    if (CL.cstart) (void)hipFree(CL.cstart);
    if (CL.coff) (void)hipFree(CL.coff);
    if (CL.cval) (void)hipFree(CL.cval);
    CL = CellLists{};
}

// The screening term of the face-form CG pass, cell by cell.  Each occupied
// cell loads its 8 corners' operand once ((r - alpha q) + beta d_old, as the
// OP_CG-mode gather forms it; 0 outside the grid, where the weight is 0 too),
// forms s_p = wh3 * sum_k phi_k x_k for each of its points (the gather's
// expression) and sums sig[8c + k] = sum_p phi_k(p) s_p.  Skipped when *done.
template <typename T>
__global__ void k_cell_gather(uint32_t D, int64_t C, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cell,
                              const T* __restrict__ tt, int64_t P, T wh3, const T* __restrict__ r, const T* __restrict__ dd,
                              const T* __restrict__ qq, const CgState* cst, const int32_t* done, T* sig) {
    if (done && *done) return;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double bt = cst->beta, al = cst->alpha;
    const int rp = cst->rpend;
    const uint32_t res = 1u << D, p0 = cstart[c], p1 = cstart[c + 1];
    const uint32_t cx = cell[p0], cy = cell[P + p0], cz = cell[2 * P + p0];
    bool in[8];
    uint32_t m[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t ox = cx + (k & 1), oy = cy + ((k >> 1) & 1), oz = cz + (k >> 2);
        in[k] = ox < res && oy < res && oz < res;
        m[k] = in[k] ? morton3(ox, oy, oz) : 0u;
    }
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = in[k] ? cg_form(bt, al, rp, dd[m[k]], r[m[k]], qq[m[k]]) : T(0);
    T sg[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) sg[k] = T(0);
    for (uint32_t p = p0; p < p1; ++p) {
        const T t0 = tt[p], t1 = tt[P + p], t2 = tt[2 * P + p];
        T w[8], acc = T(0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const T bx = (k & 1) ? t0 : T(1) - t0;
            const T by = ((k >> 1) & 1) ? t1 : T(1) - t1;
            const T bz = (k >> 2) ? t2 : T(1) - t2;
            w[k] = in[k] ? bx * by * bz : T(0);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += w[k] * v[k];
        const T sp = wh3 * acc;
#pragma unroll
        for (int k = 0; k < 8; ++k) sg[k] += w[k] * sp;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) sig[8 * c + k] = sg[k];
}
// sc[t] = sum of sig over the (cell, corner) pairs of touched node t, in cell order
template <typename T>
__global__ void k_cell_sums(int64_t Tn, const uint32_t* __restrict__ coff, const uint32_t* __restrict__ cval,
                            const T* __restrict__ sig, T* sc, const int32_t* done) {
    if (done && *done) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Tn) return;
    T acc = T(0);
    for (uint32_t j = coff[t]; j < coff[t + 1]; ++j) acc += sig[cval[j]];
    sc[t] = acc;
}
template <typename T>
void launch_cell_screening(const LevelView& L, const CellLists& CL, const T* r, const T* d, const T* q, const CgState* cst,
                           const int32_t* done, T* sig, T* sc, hipStream_t s) {
    const DevPoints& pts = *L.pts;
    const NodeList& S = *L.scr;
    if (CL.C <= 0 || S.T <= 0) return;
    const T* tt;
    T wh3;
    if constexpr (sizeof(T) == 8) { tt = (const T*)pts.t64; wh3 = (T)L.k64.wh3; }
    else { tt = (const T*)pts.t32; wh3 = (T)L.k32.wh3; }
    hipLaunchKernelGGL((k_cell_gather<T>), dim3(nblk(CL.C)), dim3(256), 0, s, L.D, CL.C, CL.cstart, pts.cell, tt, pts.P, wh3, r, d,
                       q, cst, done, sig);
    hipLaunchKernelGGL((k_cell_sums<T>), dim3(nblk(S.T)), dim3(256), 0, s, S.T, CL.coff, CL.cval, sig, sc, done);
    PRX_LAUNCH_CHECK();
}
template void launch_cell_screening<double>(const LevelView&, const CellLists&, const double*, const double*, const double*,
                                            const CgState*, const int32_t*, double*, double*, hipStream_t);

// The same over a batch: cell c belongs to the instance of its first point;
// the instance's resolution, first node, wh3 and CG state apply, and cells of
// instances that do not run are skipped.
template <typename T>
__global__ void k_bcell_gather(BatchTab tab, int64_t C, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cell,
                               const T* __restrict__ tt, const uint32_t* __restrict__ pinst, int64_t P, const T* __restrict__ r,
                               const T* __restrict__ dd, const T* __restrict__ qq, T* sig) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t p0 = cstart[c], p1 = cstart[c + 1], k = pinst[p0];
    if (tab.run[(int64_t)k * tab.run_stride] != tab.run_when) return;
    const BCoef cf = tab.coef[k];
    const CgState& st = tab.st[k];
    const double bt = st.beta, al = st.alpha;
    const int rp = st.rpend;
    const uint64_t ib = tab.off[k];
    T wh3;
    if constexpr (sizeof(T) == 8) wh3 = (T)cf.wh3_64;
    else wh3 = (T)cf.wh3_32;
    const uint32_t res = cf.res;
    const uint32_t cx = cell[p0], cy = cell[P + p0], cz = cell[2 * P + p0];
    bool in[8];
    uint64_t m[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t ox = cx + (j & 1), oy = cy + ((j >> 1) & 1), oz = cz + (j >> 2);
        in[j] = ox < res && oy < res && oz < res;
        m[j] = in[j] ? ib + morton3(ox, oy, oz) : 0u;
    }
    T v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = in[j] ? cg_form(bt, al, rp, dd[m[j]], r[m[j]], qq[m[j]]) : T(0);
    T sg[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) sg[j] = T(0);
    for (uint32_t p = p0; p < p1; ++p) {
        const T t0 = tt[p], t1 = tt[P + p], t2 = tt[2 * P + p];
        T w[8], acc = T(0);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const T bx = (j & 1) ? t0 : T(1) - t0;
            const T by = ((j >> 1) & 1) ? t1 : T(1) - t1;
            const T bz = (j >> 2) ? t2 : T(1) - t2;
            w[j] = in[j] ? bx * by * bz : T(0);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) acc += w[j] * v[j];
        const T sp = wh3 * acc;
#pragma unroll
        for (int j = 0; j < 8; ++j) sg[j] += w[j] * sp;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) sig[8 * c + j] = sg[j];
}
// k_cell_sums over a batch: touched nodes of instances that do not run are skipped
template <typename T>
__global__ void k_bcell_sums(BatchTab tab, int64_t Tn, const uint32_t* __restrict__ node, const uint32_t* __restrict__ coff,
                             const uint32_t* __restrict__ cval, const T* __restrict__ sig, T* sc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Tn) return;
    const uint32_t k = tab.brick_inst[node[t] >> 12];
    if (tab.run[(int64_t)k * tab.run_stride] != tab.run_when) return;
    T acc = T(0);
    for (uint32_t j = coff[t]; j < coff[t + 1]; ++j) acc += sig[cval[j]];
    sc[t] = acc;
}
template <typename T>
void launch_bcell_screening(const BatchTab& tab, const CellLists& CL, const BPoints& pts, int64_t T_touched, const uint32_t* node,
                            const T* r, const T* d, const T* q, T* sig, T* sc, hipStream_t s) {
    if (CL.C <= 0 || T_touched <= 0) return;
    const T* tt;
    if constexpr (sizeof(T) == 8) tt = (const T*)pts.t64;
    else tt = (const T*)pts.t32;
    hipLaunchKernelGGL((k_bcell_gather<T>), dim3(nblk(CL.C)), dim3(256), 0, s, tab, CL.C, CL.cstart, pts.cell, tt, pts.inst, pts.P,
                       r, d, q, sig);
    hipLaunchKernelGGL((k_bcell_sums<T>), dim3(nblk(T_touched)), dim3(256), 0, s, tab, T_touched, node, CL.coff, CL.cval, sig, sc);
    PRX_LAUNCH_CHECK();
}
template void launch_bcell_screening<double>(const BatchTab&, const CellLists&, const BPoints&, int64_t, const uint32_t*,
                                             const double*, const double*, const double*, double*, double*, hipStream_t);
template void launch_bcell_screening<float>(const BatchTab&, const CellLists&, const BPoints&, int64_t, const uint32_t*,
                                            const float*, const float*, const float*, float*, float*, hipStream_t);
template void launch_cell_screening<float>(const LevelView&, const CellLists&, const float*, const float*, const float*,
                                           const CgState*, const int32_t*, float*, float*, hipStream_t);


template void launch_screen_gather<float>(const LevelView&, const float*, const double*, const float*, float*, hipStream_t,
                                          const float*, const float*, const double*, const float*, const CgState*);

template <typename T, int MODE>
__global__ void k_bscreen_gather(BatchTab tab, int64_t P, const uint32_t* __restrict__ cell, const T* __restrict__ tt,
                                 const uint32_t* __restrict__ pinst, const T* __restrict__ x, const double* __restrict__ x64,
                                 const float* __restrict__ e32, T* s, const T* __restrict__ r, const T* __restrict__ dd,
                                 const T* __restrict__ qq) {
    bscreen_gather_point<T, MODE>(tab, P, cell, tt, pinst, x, x64, e32, s, r, dd, qq, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}
template <typename T>
void launch_bscreen_gather(const BatchTab& tab, const BPoints& pts, const T* x, const double* x64, const float* e32, T* sbuf,
                           const T* r, const T* d, hipStream_t s, const T* q) {
    if (pts.P <= 0) return;
    const T* tt;
    if constexpr (sizeof(T) == 8) tt = (const T*)pts.t64;
    else tt = (const T*)pts.t32;
    if (r && q)
        hipLaunchKernelGGL((k_bscreen_gather<T, 3>), dim3(nblk(pts.P)), dim3(256), 0, s, tab, pts.P, pts.cell, tt, pts.inst, x, x64,
                           e32, sbuf, r, d, q);
    else if (r)
        hipLaunchKernelGGL((k_bscreen_gather<T, 2>), dim3(nblk(pts.P)), dim3(256), 0, s, tab, pts.P, pts.cell, tt, pts.inst, x, x64,
                           e32, sbuf, r, d, q);
    else if (x64)
        hipLaunchKernelGGL((k_bscreen_gather<T, 1>), dim3(nblk(pts.P)), dim3(256), 0, s, tab, pts.P, pts.cell, tt, pts.inst, x, x64,
                           e32, sbuf, r, d, q);
    else
        hipLaunchKernelGGL((k_bscreen_gather<T, 0>), dim3(nblk(pts.P)), dim3(256), 0, s, tab, pts.P, pts.cell, tt, pts.inst, x, x64,
                           e32, sbuf, r, d, q);
    PRX_LAUNCH_CHECK();
}
template void launch_bscreen_gather<double>(const BatchTab&, const BPoints&, const double*, const double*, const float*, double*,
                                            const double*, const double*, hipStream_t, const double*);
template void launch_bscreen_gather<float>(const BatchTab&, const BPoints&, const float*, const double*, const float*, float*,
                                           const float*, const float*, hipStream_t, const float*);

// ------------------------------------------------------------------ verification CSR
// The CSR of assemble_system (src/fem_tree.rs:180-249) as the reference
// stores it: columns = Laplacian neighbours with K != 0, the diagonal when
// pw > 0, and every node that shares a point with the row node (the 64
// screening pairs are pushed without a zero filter); sorted by column.
__device__ inline int64_t find_touched(const NodeList& L, uint32_t i) {
    if (L.T == 0) return -1;
    const uint32_t b = i >> L.brick_log;
    int64_t lo = L.brick[b], hi = L.brick[b + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        uint32_t v = L.node[mid];
        if (v == i) return mid;
        if (v < i) lo = mid + 1;
        else hi = mid;
    }
    return -1;
}

__device__ inline uint32_t row_mask(uint32_t D, uint32_t i, const NodeList& L, const DevPoints& pts, int has_diag) {
    const int64_t res = 1 << D;
    uint32_t ix, iy, iz;
    demorton3(i, ix, iy, iz);
    uint32_t mask = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int nz = (dx != 0) + (dy != 0) + (dz != 0);
                int64_t jx = (int64_t)ix + dx, jy = (int64_t)iy + dy, jz = (int64_t)iz + dz;
                if (jx < 0 || jx >= res || jy < 0 || jy >= res || jz < 0 || jz >= res) continue;
                int bit = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                if (nz != 1) mask |= 1u << bit;  // K != 0 for centre, edges, corners
            }
    if (has_diag) mask |= 1u << 13;
    const int64_t t = find_touched(L, i);
    if (t >= 0) {
        for (uint32_t j = L.off[t]; j < L.off[t + 1]; ++j) {
            const uint32_t p = L.code[j] >> 3, k = L.code[j] & 7;
            const int ex = k & 1, ey = (k >> 1) & 1, ez = k >> 2;
            // the point's cell corner c = i - (ex,ey,ez); partners c + (fx,fy,fz)
            for (int f = 0; f < 8; ++f) {
                const int fx = f & 1, fy = (f >> 1) & 1, fz = f >> 2;
                // the partner exists iff its trunc corner is in range
                const double h = 1.0 / (double)res;
                int64_t c[3];
                for (int a = 0; a < 3; ++a) c[a] = sat_i64(pts.pos[a * pts.P + p] / h);
                uint32_t ox = (uint32_t)(uint64_t)wadd(c[0], fx), oy = (uint32_t)(uint64_t)wadd(c[1], fy),
                         oz = (uint32_t)(uint64_t)wadd(c[2], fz);
                if (ox >= (uint32_t)res || oy >= (uint32_t)res || oz >= (uint32_t)res) continue;
                int bx = fx - ex, by = fy - ey, bz = fz - ez;
                mask |= 1u << ((bz + 1) * 9 + (by + 1) * 3 + (bx + 1));
            }
        }
    }
    return mask;
}

__global__ void k_csr_count(uint32_t D, NodeList L, DevPoints pts, int has_diag, int64_t* cnt) {
    const uint64_t N = 1ull << (3 * D);
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    cnt[i] = __popc(row_mask(D, (uint32_t)i, L, pts, has_diag));
}

__global__ void k_csr_fill(uint32_t D, NodeList L, DevPoints pts, int has_diag, double K0, double Ke, double Kc,
                           double wg, double wh3, const int64_t* off, int64_t* col, double* val) {
    const uint64_t N = 1ull << (3 * D);
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t res = 1 << D;
    const double h = 1.0 / (double)res;
    uint32_t mask = row_mask(D, (uint32_t)i, L, pts, has_diag);
    uint32_t ix, iy, iz;
    demorton3((uint32_t)i, ix, iy, iz);
    uint32_t cols[27];
    int bits[27];
    int n = 0;
    for (int b = 0; b < 27; ++b) {
        if (!(mask & (1u << b))) continue;
        int dx = b % 3 - 1, dy = (b / 3) % 3 - 1, dz = b / 9 - 1;
        uint32_t j = morton3(ix + dx, iy + dy, iz + dz);
        int k = n++;
        while (k > 0 && cols[k - 1] > j) { cols[k] = cols[k - 1]; bits[k] = bits[k - 1]; --k; }
        cols[k] = j;
        bits[k] = b;
    }
    const int64_t t = find_touched(L, (uint32_t)i);
    const int64_t base = off[i];
    for (int e = 0; e < n; ++e) {
        const int b = bits[e];
        const int dx = b % 3 - 1, dy = (b / 3) % 3 - 1, dz = b / 9 - 1;
        const int nz = (dx != 0) + (dy != 0) + (dz != 0);
        double v = 0.0;
        if (nz != 1) v += (nz == 0 ? K0 : (nz == 2 ? Ke : Kc));
        if (nz == 0 && has_diag) v += wg;
        if (t >= 0) {
            for (uint32_t j = L.off[t]; j < L.off[t + 1]; ++j) {
                const uint32_t p = L.code[j] >> 3, k = L.code[j] & 7;
                const int ex = k & 1, ey = (k >> 1) & 1, ez = k >> 2;
                const int fx = ex + dx, fy = ey + dy, fz = ez + dz;
                if (fx < 0 || fx > 1 || fy < 0 || fy > 1 || fz < 0 || fz > 1) continue;
                int64_t c[3];
                double tt[3];
                for (int a = 0; a < 3; ++a) {
                    double q = pts.pos[a * pts.P + p];
                    c[a] = sat_i64(q / h);
                    tt[a] = q / h - (double)c[a];
                }
                uint32_t ox = (uint32_t)(uint64_t)wadd(c[0], fx), oy = (uint32_t)(uint64_t)wadd(c[1], fy),
                         oz = (uint32_t)(uint64_t)wadd(c[2], fz);
                if (ox >= (uint32_t)res || oy >= (uint32_t)res || oz >= (uint32_t)res) continue;
                const double bxj = fx ? tt[0] : 1. - tt[0], byj = fy ? tt[1] : 1. - tt[1], bzj = fz ? tt[2] : 1. - tt[2];
                const double phij = bxj * byj * bzj;
                v += wh3 * L.phi[j] * phij;
            }
        }
        col[base + e] = cols[e];
        val[base + e] = v;
    }
}

void launch_csr_count(uint32_t D, const NodeList* L, const DevPoints* pts, int has_diag, int64_t* rowcnt, hipStream_t s) {
    const int64_t N = (int64_t)1 << (3 * D);
    NodeList Lv = L ? *L : NodeList{};
    DevPoints Pv = pts ? *pts : DevPoints{};
    hipLaunchKernelGGL(k_csr_count, dim3(nblk(N)), dim3(256), 0, s, D, Lv, Pv, has_diag, rowcnt);
    PRX_LAUNCH_CHECK();
}
void launch_csr_fill(uint32_t D, const NodeList* L, const DevPoints* pts, int has_diag, double K0, double Ke, double Kc,
                     double wg, double wh3, const int64_t* off, int64_t* col, double* val, hipStream_t s) {
    const int64_t N = (int64_t)1 << (3 * D);
    NodeList Lv = L ? *L : NodeList{};
    DevPoints Pv = pts ? *pts : DevPoints{};
    hipLaunchKernelGGL(k_csr_fill, dim3(nblk(N)), dim3(256), 0, s, D, Lv, Pv, has_diag, K0, Ke, Kc, wg, wh3, off, col, val);
    PRX_LAUNCH_CHECK();
}

}  // namespace prx
