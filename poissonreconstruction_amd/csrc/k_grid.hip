// Grid-level kernels: node numbering export, vector updates and reductions of
// CG, coarse RHS and prolongation, field evaluation, generic CSR SpMV.
//
// Kernels that must reproduce the reference bit for bit (tree export, coarse
// RHS, prolongation, evaluate, CSR SpMV) evaluate every expression in the
// reference's order; contraction to FMA is off in this file so a*b+c rounds
// twice, as Rust does.
#pragma clang fp contract(off)

#include <algorithm>

#include "cg_dev.hpp"

namespace prx {

static inline int nblocks_for(int64_t n, int per_thread = 1, int threads = 256, int cap = 1 << 30) {
    int64_t b = (n + (int64_t)threads * per_thread - 1) / ((int64_t)threads * per_thread);
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

// Fixed grid of the reduction kernels: the partial count depends on n only,
// so every reduction is summed in the same order on every run.
static inline int red_blocks(int64_t n) { return nblocks_for(n, 8, 256, 2048); }

int partial_capacity(int64_t N) {
    // four per brick (8^3 or 16^3; fused CG step) and the first-stage fold
    // behind them, one per 512 nodes of k_cg_update
    int64_t b = 4 * ((N + 511) / 512) + 4 * 128;
    int64_t r = red_blocks(N);
    int64_t nv = (N + 255) / 256;   // naive stencil: one partial per block
    int64_t m = b > r ? b : r;
    if (nv > m) m = nv;
    return (int)(m < 2048 ? 2048 : m);
}

// ------------------------------------------------------------------ tree export
__global__ void k_export_offsets(uint32_t N, uint32_t* i2o) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint32_t x, y, z;
    demorton3(i, x, y, z);
    i2o[3 * (uint64_t)i] = x;
    i2o[3 * (uint64_t)i + 1] = y;
    i2o[3 * (uint64_t)i + 2] = z;
}
void launch_export_offsets(uint32_t D, uint32_t* i2o, hipStream_t s) {
    uint32_t N = 1u << (3 * D);
    hipLaunchKernelGGL(k_export_offsets, dim3(nblocks_for(N)), dim3(256), 0, s, N, i2o);
    PRX_LAUNCH_CHECK();
}

__global__ void k_export_g2i(uint32_t D, int64_t* g2i) {
    const uint32_t res = 1u << D;
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (uint64_t)res * res * res) return;
    uint32_t x = r % res, y = (r / res) % res, z = r / ((uint64_t)res * res);
    g2i[r] = (int64_t)morton3(x, y, z);
}
void launch_export_g2i(uint32_t D, int64_t* g2i, hipStream_t s) {
    int64_t n = (int64_t)1 << (3 * D);
    hipLaunchKernelGGL(k_export_g2i, dim3(nblocks_for(n)), dim3(256), 0, s, D, g2i);
    PRX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ fills
template <typename T>
__global__ void k_fill(T* p, T v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
template <typename T>
void launch_fill(T* p, T v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(nblocks_for(n, 4, 256, 65536)), dim3(256), 0, s, p, v, n);
    PRX_LAUNCH_CHECK();
}
template void launch_fill<double>(double*, double, int64_t, hipStream_t);
template void launch_fill<float>(float*, float, int64_t, hipStream_t);
template void launch_fill<uint32_t>(uint32_t*, uint32_t, int64_t, hipStream_t);
template void launch_fill<int64_t>(int64_t*, int64_t, int64_t, hipStream_t);

// ------------------------------------------------------------------ reductions
template <typename T>
__global__ __launch_bounds__(256) void k_dot(const T* a, const T* b, int64_t n, double* partial) {
    __shared__ double ws[4];
    using VT = typename V16<T>::type;
    constexpr int VEC = V16<T>::n;
    double acc = 0.0;
    const int64_t nv = n / VEC, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        const VT av = ((const VT*)a)[i], bv = ((const VT*)b)[i];
        const T *as = (const T*)&av, *bs = (const T*)&bv;
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc += (double)as[e] * (double)bs[e];
    }
    for (int64_t i = nv * VEC + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        acc += (double)a[i] * (double)b[i];
    double s = block_sum_256(acc, ws);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
template <typename T>
void launch_dot(const T* a, const T* b, int64_t n, double* partial, int* nblk, hipStream_t s) {
    int nb = red_blocks(n);
    hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(256), 0, s, a, b, n, partial);
    PRX_LAUNCH_CHECK();
    *nblk = nb;
}
template void launch_dot<double>(const double*, const double*, int64_t, double*, int*, hipStream_t);
template void launch_dot<float>(const float*, const float*, int64_t, double*, int*, hipStream_t);

// One CG vector update at a node, evaluated in FP64 and rounded once to the
// vector type (for the FP32 correction system this keeps r consistent with
// x to one FP32 rounding per step; FP32 arithmetic here costs iterations).
template <typename T>
__device__ __forceinline__ void cg_upd(T& x, T& r, T d, T q, double a, double& acc) {
    x = (T)((double)x + a * (double)d);
    r = r_form(a, q, r);
    acc += (double)r * (double)r;
}
template <typename T>
__device__ __forceinline__ void cg_upd_r(T& r, T q, double a, double& acc) {
    r = r_form(a, q, r);
    acc += (double)r * (double)r;
}
// the two pending terms of a lazy pair: x + alpha_pend * dprev + alpha * d
template <typename T>
__device__ __forceinline__ T x_pair(T x, T dp, T d, double ap, double a) {
    return (T)(((double)x + ap * (double)dp) + a * (double)d);
}

// One block folds the partials in a fixed order and advances the CG scalars
// exactly as solve_cg does (src/solvers.rs:36-64).
__global__ __launch_bounds__(1024) void k_fin(CgState* st, const double* partial, int np, int op, int slot) {
    __shared__ double ws[16];
    if (op != FIN_SLOT && op != FIN_INIT_REF && op != FIN_INIT_TRUE && op != FIN_PINIT && st->done) return;
    const double sum = fin_sum(partial, np, ws);
    if (threadIdx.x != 0) return;
    fin_step(st, sum, op, slot);
}
__global__ __launch_bounds__(256) void k_fold4(const CgState* st, const double* __restrict__ partial, int np, double* __restrict__ out) {
    __shared__ double ws[16];
    if (st->done) return;
    fold4_block(partial, np, blockIdx.x, out, ws);
}
// FIN_CG: the first stage writes behind the partials (partial + 4 np)
__global__ __launch_bounds__(1024) void k_fin_cg(CgState* st, const double* partial, int np, int slot) {
    __shared__ double ws[16];
    if (st->done) return;
    double v[4];
    fold4_final(partial, np, partial + 4 * (int64_t)np, ws, v);
    if (threadIdx.x == 0) fin_cg(st, v[0], v[1], v[2], v[3], slot);
}
void launch_fin(CgState* st, const double* partial, int np, int op, int slot, hipStream_t s) {
    if (op == FIN_CG) {
        const int g = fold_blocks(np);
        if (g) {
            hipLaunchKernelGGL(k_fold4, dim3(g), dim3(256), 0, s, st, partial, np, const_cast<double*>(partial) + 4 * (int64_t)np);
            PRX_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_fin_cg, dim3(1), dim3(1024), 0, s, st, partial, np, slot);
        PRX_LAUNCH_CHECK();
        return;
    }
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(1024), 0, s, st, partial, np, op, slot);
    PRX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ batched CG pieces
__global__ __launch_bounds__(1024) void k_bfin(BatchTab tab, const double* partial, int ulog, int op, int slot) {
    __shared__ double ws[16];
    const uint32_t k = blockIdx.x;
    if (!b_runs(tab, k)) return;
    const int64_t p0 = (int64_t)(tab.off[k] >> ulog), p1 = (int64_t)(tab.off[k + 1] >> ulog);
    const double sum = fin_sum(partial + p0, p1 - p0, ws);
    if (threadIdx.x != 0) return;
    fin_step(tab.st + k, sum, op, slot);
}
void launch_bfin(const BatchTab& tab, const double* partial, int ulog, int op, int slot, hipStream_t s) {
    hipLaunchKernelGGL(k_bfin, dim3(tab.K), dim3(1024), 0, s, tab, partial, ulog, op, slot);
    PRX_LAUNCH_CHECK();
}

// FIN_CG per instance, in the single-tree order (fold4_block, fold4_final)
// over the instance's bricks [off[k] >> ulog, off[k+1] >> ulog)
__global__ __launch_bounds__(256) void k_bfold4(BatchTab tab, const double* __restrict__ partial, int ulog, double* __restrict__ fold) {
    __shared__ double ws[16];
    const uint32_t k = blockIdx.y;
    if (!b_runs(tab, k)) return;
    const int64_t p0 = (int64_t)(tab.off[k] >> ulog), np = (int64_t)(tab.off[k + 1] >> ulog) - p0;
    if ((int)blockIdx.x >= fold_blocks((int)np)) return;
    fold4_block(partial + 4 * p0, (int)np, blockIdx.x, fold + (int64_t)k * FOLD_BLOCKS * 4, ws);
}
__global__ __launch_bounds__(1024) void k_bfin_cg(BatchTab tab, const double* partial, int ulog, const double* fold, int slot) {
    __shared__ double ws[16];
    const uint32_t k = blockIdx.x;
    if (!b_runs(tab, k)) return;
    const int64_t p0 = (int64_t)(tab.off[k] >> ulog), np = (int64_t)(tab.off[k + 1] >> ulog) - p0;
    double v[4];
    fold4_final(partial + 4 * p0, (int)np, fold + (int64_t)k * FOLD_BLOCKS * 4, ws, v);
    if (threadIdx.x == 0) fin_cg(tab.st + k, v[0], v[1], v[2], v[3], slot);
}
void launch_bfin_cg(const BatchTab& tab, const double* partial, int ulog, double* fold, int slot, hipStream_t s) {
    hipLaunchKernelGGL(k_bfold4, dim3(FOLD_BLOCKS, tab.K), dim3(256), 0, s, tab, partial, ulog, fold);
    PRX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bfin_cg, dim3(tab.K), dim3(1024), 0, s, tab, partial, ulog, fold, slot);
    PRX_LAUNCH_CHECK();
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_bx_terms(BatchTab tab, T* __restrict__ x, XDirs<T> dd, const T* __restrict__ ring,
                                                  int64_t Ntot) {
    bx_terms_block<T, MODE>(tab, x, dd, ring, Ntot, blockIdx.x);
}
template <typename T>
void launch_bx_terms(const BatchTab& tab, T* x, const XDirs<T>& d, const T* ring, int64_t Ntot, int mode, hipStream_t s) {
    constexpr int VEC = 16 / (int)sizeof(T);
    const unsigned nb = (unsigned)(Ntot / (256 * VEC));  // instances are multiples of 4096 nodes
    if (mode == 0) hipLaunchKernelGGL((k_bx_terms<T, 0>), dim3(nb), dim3(256), 0, s, tab, x, d, ring, Ntot);
    else if (mode == 1) hipLaunchKernelGGL((k_bx_terms<T, 1>), dim3(nb), dim3(256), 0, s, tab, x, d, ring, Ntot);
    else hipLaunchKernelGGL((k_bx_terms<T, 2>), dim3(nb), dim3(256), 0, s, tab, x, d, ring, Ntot);
    PRX_LAUNCH_CHECK();
}
template void launch_bx_terms<double>(const BatchTab&, double*, const XDirs<double>&, const double*, int64_t, int, hipStream_t);
template void launch_bx_terms<float>(const BatchTab&, float*, const XDirs<float>&, const float*, int64_t, int, hipStream_t);

__global__ __launch_bounds__(256) void k_bcommit(BatchTab tab, double* __restrict__ x, const double* __restrict__ x2, float* __restrict__ e) {
    bcommit_block(tab, x, x2, e, blockIdx.x);
}
void launch_bcommit(const BatchTab& tab, double* x, const double* x2, float* e, int64_t Ntot, hipStream_t s) {
    hipLaunchKernelGGL(k_bcommit, dim3((unsigned)(Ntot / 1024)), dim3(256), 0, s, tab, x, x2, e);
    PRX_LAUNCH_CHECK();
}

// r -= alpha_k q ; partial = r.r per block ; x as in k_cg_update (eager, or
// lazy in pairs).  A block covers 1024 vectors (4096 float or 2048 double
// nodes) inside one 4096-node brick, so it belongs to one instance.
template <typename T>
__global__ __launch_bounds__(256) void k_bcg_update(BatchTab tab, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ dp,
                                                    const T* __restrict__ d, const T* __restrict__ q, double* partial) {
    __shared__ double ws[4];
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T);
    const uint32_t k = tab.brick_inst[((uint64_t)blockIdx.x * 1024 * VEC) >> 12];
    if (!b_runs(tab, k)) return;
    const CgState* st = tab.st + k;
    const double a = st->alpha, ap = st->alpha_pend;
    const int mode = !st->lazy ? 0 : (st->xpend ? 2 : 1);
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t j = ((int64_t)blockIdx.x * 1024 + u * 256 + threadIdx.x) * VEC;
        NV rv = ld_nt(r + j);
        const NV qv = ld_nt(q + j);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            T re = rv[e];
            cg_upd_r<T>(re, qv[e], a, acc);
            rv[e] = re;
        }
        st_nt(r + j, rv);
        if (mode != 1) {
            NV xv = ld_nt(x + j);
            const NV dv = ld_nt(d + j);
            if (mode == 0) {
#pragma unroll
                for (int e = 0; e < VEC; ++e) xv[e] = (T)((double)xv[e] + a * (double)dv[e]);
            } else {
                const NV pv = ld_nt(dp + j);
#pragma unroll
                for (int e = 0; e < VEC; ++e) xv[e] = x_pair<T>(xv[e], pv[e], dv[e], ap, a);
            }
            st_nt(x + j, xv);
        }
    }
    const double sm = block_sum_256(acc, ws);
    if (threadIdx.x == 0) partial[blockIdx.x] = sm;
}
template <typename T>
int launch_bcg_update(const BatchTab& tab, T* x, T* r, const T* dp, const T* d, const T* q, int64_t Ntot, double* partial,
                      hipStream_t s) {
    constexpr int VEC = 16 / (int)sizeof(T);
    const int64_t nb = Ntot / (1024 * VEC);
    hipLaunchKernelGGL(k_bcg_update<T>, dim3((unsigned)nb), dim3(256), 0, s, tab, x, r, dp, d, q, partial);
    PRX_LAUNCH_CHECK();
    return VEC == 2 ? 11 : 12;
}
template int launch_bcg_update<double>(const BatchTab&, double*, double*, const double*, const double*, const double*, int64_t, double*,
                                       hipStream_t);
template int launch_bcg_update<float>(const BatchTab&, float*, float*, const float*, const float*, const float*, int64_t, double*,
                                      hipStream_t);

// refresh iteration: x brought fully up to date (the pending lazy term too)
template <typename T>
__global__ __launch_bounds__(256) void k_baxpy_alpha(BatchTab tab, T* __restrict__ x, const T* __restrict__ dp, const T* __restrict__ d) {
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T);
    const int64_t j = ((int64_t)blockIdx.x * 256 + threadIdx.x) * VEC;
    const uint32_t k = tab.brick_inst[(uint64_t)blockIdx.x * 256 * VEC >> 12];
    if (!b_runs(tab, k)) return;
    const CgState* st = tab.st + k;
    const double a = st->alpha, ap = st->alpha_pend;
    const bool pair = st->lazy && st->xpend;
    const NV xv = ld_v(x + j), dv = ld_v(d + j);
    NV xo;
    if (pair) {
        const NV pv = ld_v(dp + j);
#pragma unroll
        for (int e = 0; e < VEC; ++e) xo[e] = x_pair<T>(xv[e], pv[e], dv[e], ap, a);
    } else {
#pragma unroll
        for (int e = 0; e < VEC; ++e) xo[e] = (T)((double)xv[e] + a * (double)dv[e]);
    }
    *(NV*)(x + j) = xo;
}
template <typename T>
void launch_baxpy_alpha(const BatchTab& tab, T* x, const T* dp, const T* d, int64_t Ntot, hipStream_t s) {
    constexpr int VEC = 16 / (int)sizeof(T);
    hipLaunchKernelGGL(k_baxpy_alpha<T>, dim3((unsigned)(Ntot / (256 * VEC))), dim3(256), 0, s, tab, x, dp, d);
    PRX_LAUNCH_CHECK();
}
template void launch_baxpy_alpha<double>(const BatchTab&, double*, const double*, const double*, int64_t, hipStream_t);
template void launch_baxpy_alpha<float>(const BatchTab&, float*, const float*, const float*, int64_t, hipStream_t);

// per instance: x += alpha_pend * d{dir_buf} where a lazy term is pending
template <typename T>
__global__ __launch_bounds__(256) void k_bflush_x(BatchTab tab, T* __restrict__ x, const T* __restrict__ d0, const T* __restrict__ d1) {
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T);
    const int64_t j = ((int64_t)blockIdx.x * 256 + threadIdx.x) * VEC;
    const uint32_t k = tab.brick_inst[(uint64_t)blockIdx.x * 256 * VEC >> 12];
    if (!b_runs(tab, k)) return;
    const CgState* st = tab.st + k;
    if (!(st->lazy && st->xpend)) return;
    const double ap = st->alpha_pend;
    const NV xv = ld_v(x + j), dv = ld_v((st->dir_buf ? d1 : d0) + j);
    NV xo;
#pragma unroll
    for (int e = 0; e < VEC; ++e) xo[e] = (T)((double)xv[e] + ap * (double)dv[e]);
    *(NV*)(x + j) = xo;
}
template <typename T>
void launch_bflush_x(const BatchTab& tab, T* x, const T* d0, const T* d1, int64_t Ntot, hipStream_t s) {
    constexpr int VEC = 16 / (int)sizeof(T);
    hipLaunchKernelGGL(k_bflush_x<T>, dim3((unsigned)(Ntot / (256 * VEC))), dim3(256), 0, s, tab, x, d0, d1);
    PRX_LAUNCH_CHECK();
}
template void launch_bflush_x<double>(const BatchTab&, double*, const double*, const double*, int64_t, hipStream_t);
template void launch_bflush_x<float>(const BatchTab&, float*, const float*, const float*, int64_t, hipStream_t);

template <typename T>
__global__ __launch_bounds__(256) void k_bdot(BatchTab tab, const T* __restrict__ a, const T* __restrict__ b, double* partial) {
    __shared__ double ws[4];
    bdot_block<T>(tab, a, b, partial, blockIdx.x, ws);
}
template <typename T>
int launch_bdot(const BatchTab& tab, const T* a, const T* b, int64_t Ntot, double* partial, hipStream_t s) {
    constexpr int VEC = 16 / (int)sizeof(T);
    hipLaunchKernelGGL(k_bdot<T>, dim3((unsigned)(Ntot / (256 * VEC))), dim3(256), 0, s, tab, a, b, partial);
    PRX_LAUNCH_CHECK();
    return VEC == 2 ? 9 : 10;
}
template int launch_bdot<double>(const BatchTab&, const double*, const double*, int64_t, double*, hipStream_t);
template int launch_bdot<float>(const BatchTab&, const float*, const float*, int64_t, double*, hipStream_t);

// x += alpha d ; r -= alpha q ; partial = r.r   (src/solvers.rs:47, :53, :57)
// Four 16-byte vectors per array per thread (1024 vectors per block, each
// quarter coalesced), streamed with non-temporal loads and stores; one
// partial per block.
template <typename T>
__global__ __launch_bounds__(256) void k_cg_update(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ dp,
                                                   const T* __restrict__ d, const T* __restrict__ q, const CgState* st, int64_t n,
                                                   double* partial) {
    __shared__ double ws[4];
    if (st->done) return;
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T);
    const double a = st->alpha, ap = st->alpha_pend;
    // 0: x += a d every iteration; 1: lazy, x untouched; 2: lazy, both pending terms
    const int mode = !st->lazy ? 0 : (st->xpend ? 2 : 1);
    double acc = 0.0;
    const int64_t nv = n / VEC;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t i = (int64_t)blockIdx.x * 1024 + u * 256 + threadIdx.x;
        if (i < nv) {
            const int64_t j = i * VEC;
            NV rv = ld_nt(r + j);
            const NV qv = ld_nt(q + j);
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                T re = rv[e];
                cg_upd_r<T>(re, qv[e], a, acc);
                rv[e] = re;
            }
            st_nt(r + j, rv);
            if (mode != 1) {
                NV xv = ld_nt(x + j);
                const NV dv = ld_nt(d + j);
                if (mode == 0) {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) xv[e] = (T)((double)xv[e] + a * (double)dv[e]);
                } else {
                    const NV pv = ld_nt(dp + j);
#pragma unroll
                    for (int e = 0; e < VEC; ++e) xv[e] = x_pair<T>(xv[e], pv[e], dv[e], ap, a);
                }
                st_nt(x + j, xv);
            }
        }
    }
    if (blockIdx.x == gridDim.x - 1) {  // tail (n not a multiple of the vector width)
        const int64_t j = nv * VEC + threadIdx.x;
        if (j < n) {
            T rs = r[j];
            cg_upd_r<T>(rs, q[j], a, acc);
            r[j] = rs;
            if (mode == 0) x[j] = (T)((double)x[j] + a * (double)d[j]);
            else if (mode == 2) x[j] = x_pair<T>(x[j], dp[j], d[j], ap, a);
        }
    }
    double sm = block_sum_256(acc, ws);
    if (threadIdx.x == 0) partial[blockIdx.x] = sm;
}
template <typename T>
void launch_cg_update(T* x, T* r, const T* dp, const T* d, const T* q, const CgState* st, int64_t n, double* partial, int* nblk,
                      hipStream_t s) {
    constexpr int VEC = 16 / (int)sizeof(T);
    const int64_t nv = n / VEC;
    const int nb = (int)std::max<int64_t>(1, (nv + 1023) / 1024);
    hipLaunchKernelGGL(k_cg_update<T>, dim3(nb), dim3(256), 0, s, x, r, dp, d, q, st, n, partial);
    PRX_LAUNCH_CHECK();
    *nblk = nb;
}
template void launch_cg_update<double>(double*, double*, const double*, const double*, const double*, const CgState*, int64_t, double*,
                                       int*, hipStream_t);
template void launch_cg_update<float>(float*, float*, const float*, const float*, const float*, const CgState*, int64_t, double*, int*,
                                      hipStream_t);

// refresh iteration: x brought fully up to date (the pending lazy term too)
template <typename T>
__global__ void k_axpy_alpha(T* __restrict__ x, const T* __restrict__ dp, const T* __restrict__ d, const CgState* st, int64_t n) {
    if (st->done) return;
    const double a = st->alpha, ap = st->alpha_pend;
    const bool pair = st->lazy && st->xpend;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = pair ? x_pair<T>(x[i], dp[i], d[i], ap, a) : (T)((double)x[i] + a * (double)d[i]);
}
template <typename T>
void launch_axpy_alpha(T* x, const T* dp, const T* d, const CgState* st, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy_alpha<T>, dim3(nblocks_for(n, 4, 256, 65536)), dim3(256), 0, s, x, dp, d, st, n);
    PRX_LAUNCH_CHECK();
}
template void launch_axpy_alpha<double>(double*, const double*, const double*, const CgState*, int64_t, hipStream_t);
template void launch_axpy_alpha<float>(float*, const float*, const float*, const CgState*, int64_t, hipStream_t);

template <typename T>
__global__ void k_flush_x(T* __restrict__ x, const T* __restrict__ d0, const T* __restrict__ d1, const CgState* st, int64_t n) {
    if (!(st->lazy && st->xpend)) return;
    const double ap = st->alpha_pend;
    const T* d = st->dir_buf ? d1 : d0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (T)((double)x[i] + ap * (double)d[i]);
}
template <typename T>
void launch_flush_x(T* x, const T* d0, const T* d1, const CgState* st, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_flush_x<T>, dim3(nblocks_for(n, 4, 256, 65536)), dim3(256), 0, s, x, d0, d1, st, n);
    PRX_LAUNCH_CHECK();
}
template void launch_flush_x<double>(double*, const double*, const double*, const CgState*, int64_t, hipStream_t);
template void launch_flush_x<float>(float*, const float*, const float*, const CgState*, int64_t, hipStream_t);
__global__ void k_clear_xpend(CgState* st) {
    if (threadIdx.x == 0) {
        st->xpend = 0;
        st->xn = 0;
    }
}
void launch_clear_xpend(CgState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_xpend, dim3(1), dim3(64), 0, s, st);
    PRX_LAUNCH_CHECK();
}

// Fused CG x terms (n a multiple of the vector width: the fused step runs on
// tiled levels only): x + xa[0] d[0] + xa[1] d[1] + ... in FP64, rounded
// once.  MODE 0: all CG_XM terms, only when all are pending and the solve
// runs; MODE 1: whatever is pending; MODE 2: the same while not done.  One
// 16-byte vector per thread: on MI355X at 2^27 nodes this streams 3-6%
// faster than two per thread (0.92 against 0.95 ms in FP32).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_x_terms(T* __restrict__ x, XDirs<T> dd, const CgState* st, int64_t nv) {
    if (MODE != 1 && st->done) return;
    const int xn = st->xn;
    if (MODE == 0 ? xn != CG_XM : xn == 0) return;
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T);
    double a[CG_XM];
#pragma unroll
    for (int j = 0; j < CG_XM; ++j) a[j] = st->xa[j];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nv) return;
    const int64_t g = i * VEC;
    const NV xv = ld_nt(x + g);
    NV dv[CG_XM];
#pragma unroll
    for (int j = 0; j < CG_XM; ++j)
        if (j < xn && dd.d[j]) dv[j] = ld_nt(dd.d[j] + g);
    NV xo;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        double v = (double)xv[e];
#pragma unroll
        for (int j = 0; j < CG_XM; ++j)
            if (j < xn && dd.d[j]) v += a[j] * (double)dv[j][e];
        xo[e] = (T)v;
    }
    st_nt(x + g, xo);
}
template <typename T>
static void launch_x_terms(int mode, T* x, const XDirs<T>& d, const CgState* st, int64_t n, hipStream_t s) {
    constexpr int VEC = 16 / (int)sizeof(T);
    if (n % VEC) throw Error(PRX_ERR_INVALID, "fused CG vectors must be a multiple of the vector width");
    const int64_t nv = n / VEC;
    const int64_t nb64 = std::max<int64_t>(1, (nv + 255) / 256);
    if (nb64 > INT32_MAX) throw Error(PRX_ERR_INVALID, "fused CG vector too long for one launch");
    const int nb = (int)nb64;
    if (mode == 0) hipLaunchKernelGGL((k_x_terms<T, 0>), dim3(nb), dim3(256), 0, s, x, d, st, nv);
    else if (mode == 1) hipLaunchKernelGGL((k_x_terms<T, 1>), dim3(nb), dim3(256), 0, s, x, d, st, nv);
    else hipLaunchKernelGGL((k_x_terms<T, 2>), dim3(nb), dim3(256), 0, s, x, d, st, nv);
    PRX_LAUNCH_CHECK();
}
template <typename T>
void launch_x_terms_full(T* x, const XDirs<T>& d, const CgState* st, int64_t n, hipStream_t s) {
    launch_x_terms<T>(0, x, d, st, n, s);
}
template <typename T>
void launch_x_flush(T* x, const XDirs<T>& d, const CgState* st, int64_t n, bool if_running, hipStream_t s) {
    launch_x_terms<T>(if_running ? 2 : 1, x, d, st, n, s);
}
template void launch_x_terms_full<double>(double*, const XDirs<double>&, const CgState*, int64_t, hipStream_t);
template void launch_x_terms_full<float>(float*, const XDirs<float>&, const CgState*, int64_t, hipStream_t);
template void launch_x_flush<double>(double*, const XDirs<double>&, const CgState*, int64_t, bool, hipStream_t);
template void launch_x_flush<float>(float*, const XDirs<float>&, const CgState*, int64_t, bool, hipStream_t);

// d = r + beta d   (src/solvers.rs:64)
template <typename T>
__global__ void k_xpby_beta(T* __restrict__ d, const T* __restrict__ r, const CgState* st, int64_t n) {
    if (st->done) return;
    const double b = st->beta;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (T)((double)r[i] + b * (double)d[i]);
}
template <typename T>
void launch_xpby_beta(T* d, const T* r, const CgState* st, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_xpby_beta<T>, dim3(nblocks_for(n, 4, 256, 65536)), dim3(256), 0, s, d, r, st, n);
    PRX_LAUNCH_CHECK();
}
template void launch_xpby_beta<double>(double*, const double*, const CgState*, int64_t, hipStream_t);
template void launch_xpby_beta<float>(float*, const float*, const CgState*, int64_t, hipStream_t);

// r = b - q ; partial = r.r   (src/solvers.rs:33, :51)
template <typename T>
__global__ __launch_bounds__(256) void k_sub(T* __restrict__ r, const T* __restrict__ b, const T* __restrict__ q, int64_t n,
                                             double* partial, const CgState* st) {
    __shared__ double ws[4];
    if (st && st->done) return;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T v = b[i] - q[i];
        r[i] = v;
        acc += (double)v * (double)v;
    }
    double s = block_sum_256(acc, ws);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
template <typename T>
void launch_sub(T* r, const T* b, const T* q, int64_t n, double* partial, int* nblk, const CgState* st, hipStream_t s) {
    int nb = red_blocks(n);
    hipLaunchKernelGGL(k_sub<T>, dim3(nb), dim3(256), 0, s, r, b, q, n, partial, st);
    PRX_LAUNCH_CHECK();
    *nblk = nb;
}
template void launch_sub<double>(double*, const double*, const double*, int64_t, double*, int*, const CgState*, hipStream_t);
template void launch_sub<float>(float*, const float*, const float*, int64_t, double*, int*, const CgState*, hipStream_t);

__global__ void k_cast_scale_d2f(float* dst, const double* src, double scale, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)(src[i] * scale);
}
void launch_cast_scale_d2f(float* dst, const double* src, double scale, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_cast_scale_d2f, dim3(nblocks_for(n, 4, 256, 65536)), dim3(256), 0, s, dst, src, scale, n);
    PRX_LAUNCH_CHECK();
}
__global__ void k_add_f2d(double* x, const float* e, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] += (double)e[i];
}
void launch_add_f2d(double* x, const float* e, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_add_f2d, dim3(nblocks_for(n, 4, 256, 65536)), dim3(256), 0, s, x, e, n);
    PRX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ coarse level
// RHS of build_system_at_depth (src/fem_tree.rs:386-402): the fine nodes of
// one coarse cell are the 8^(D-d) consecutive Morton indices under it, summed
// in FEM index order as (v.x + v.y) + v.z, then + wh3*0.5.
__global__ void k_coarse_rhs(uint32_t D, uint32_t d, const double* __restrict__ nf, double add, int has_add, double* rhs) {
    const uint64_t nc = 1ull << (3 * d), nfine = 1ull << (3 * D);
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint64_t per = 1ull << (3 * (D - d));
    double s = 0.0;
    uint64_t k = 0;
    // batches of 8 nodes: all 24 loads are issued before the (ordered) adds
    for (; k + 8 <= per; k += 8) {
        double a[8], b[8], e[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t i = c * per + k + u;
            a[u] = nf[i];
            b[u] = nf[nfine + i];
            e[u] = nf[2 * nfine + i];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s += a[u] + b[u] + e[u];
    }
    for (; k < per; ++k) {
        const uint64_t i = c * per + k;
        s += nf[i] + nf[nfine + i] + nf[2 * nfine + i];
    }
    if (has_add) s += add;
    rhs[c] = s;
}
// The same sums over the splat list's nodes only.  The normal field is +0.0
// at every other node (k_splat writes only the list's nodes), such a node
// adds (0 + 0) + 0 = +0 to a sum that is never -0, so leaving it out changes
// no bit; the list is sorted by node, so its nodes under one coarse cell come
// in FEM index order.  brick8[b .. b+1] is the list's range of 8^3 brick b.
__global__ void k_coarse_rhs_list(uint32_t D, uint32_t d, const double* __restrict__ nf, const uint32_t* __restrict__ node,
                                  const uint32_t* __restrict__ brick8, double add, int has_add, double* rhs) {
    const uint64_t nc = 1ull << (3 * d), nfine = 1ull << (3 * D);
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int sh = 3 * (int)(D - d);  // log2 of fine nodes per coarse cell
    const uint64_t f0 = c << sh, f1 = (c + 1) << sh;
    uint32_t lo = brick8[f0 >> 9], hi = brick8[(f1 + 511) >> 9];
    if (sh < 9) {  // several coarse cells share the 8^3 brick: narrow to [f0, f1)
        uint32_t a = lo, b = hi;
        while (a < b) {
            const uint32_t m = (a + b) >> 1;
            if ((uint64_t)node[m] < f0) a = m + 1;
            else b = m;
        }
        lo = a;
        b = hi;
        while (a < b) {
            const uint32_t m = (a + b) >> 1;
            if ((uint64_t)node[m] < f1) a = m + 1;
            else b = m;
        }
        hi = a;
    }
    double s = 0.0;
    for (uint32_t t = lo; t < hi; ++t) {
        const uint64_t i = node[t];
        s += nf[i] + nf[nfine + i] + nf[2 * nfine + i];
    }
    if (has_add) s += add;
    rhs[c] = s;
}
void launch_coarse_rhs(uint32_t D, uint32_t d, const double* nf, double add, double* rhs, const NodeList* splat, hipStream_t s) {
    int64_t nc = (int64_t)1 << (3 * d);
    if (splat && splat->brick8 && D >= 3 && d < D)
        hipLaunchKernelGGL(k_coarse_rhs_list, dim3(nblocks_for(nc)), dim3(256), 0, s, D, d, nf, splat->node, splat->brick8, add,
                           add > 0.0 ? 1 : 0, rhs);
    else
        hipLaunchKernelGGL(k_coarse_rhs, dim3(nblocks_for(nc)), dim3(256), 0, s, D, d, nf, add, add > 0.0 ? 1 : 0, rhs);
    PRX_LAUNCH_CHECK();
}

// prolongate (src/fem_tree.rs:310-343) as a gather: fine node f receives
// v_c * w from every coarse c with f = 2c + delta - 1, in increasing
// row-major c (z, then y, then x), skipping v == 0.
__global__ void k_prolongate(uint32_t D, uint32_t d, const double* __restrict__ coarse, double* fine) {
    const uint32_t N = 1u << (3 * D);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t cres = 1u << d;
    const double bw[3] = {0.5, 1.0, 0.5};
    uint32_t f[3];
    demorton3(i, f[0], f[1], f[2]);
    uint32_t cand[3][2], delta[3][2];
    int nc[3];
    for (int a = 0; a < 3; ++a) {
        nc[a] = 0;
        if ((f[a] & 1u) == 0) {
            uint32_t c = f[a] >> 1;
            if (c < cres) { cand[a][nc[a]] = c; delta[a][nc[a]] = 1; nc[a]++; }
        } else {
            uint32_t c0 = (f[a] - 1) >> 1, c1 = (f[a] + 1) >> 1;
            if (c0 < cres) { cand[a][nc[a]] = c0; delta[a][nc[a]] = 2; nc[a]++; }
            if (c1 < cres) { cand[a][nc[a]] = c1; delta[a][nc[a]] = 0; nc[a]++; }
        }
    }
    double acc = 0.0;
    for (int iz = 0; iz < nc[2]; ++iz)
        for (int iy = 0; iy < nc[1]; ++iy)
            for (int ix = 0; ix < nc[0]; ++ix) {
                double v = coarse[morton3(cand[0][ix], cand[1][iy], cand[2][iz])];
                if (v == 0.0) continue;
                double w = bw[delta[0][ix]] * bw[delta[1][iy]] * bw[delta[2][iz]];
                acc += v * w;
            }
    fine[i] = acc;
}
void launch_prolongate(uint32_t D, uint32_t d, const double* coarse, double* fine, hipStream_t s) {
    int64_t n = (int64_t)1 << (3 * D);
    hipLaunchKernelGGL(k_prolongate, dim3(nblocks_for(n)), dim3(256), 0, s, D, d, coarse, fine);
    PRX_LAUNCH_CHECK();
}

__global__ void k_m2r(uint32_t d, const double* src, double* dst, int dir) {
    const uint32_t res = 1u << d;
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (uint64_t)res * res * res) return;
    uint32_t x = r % res, y = (r / res) % res, z = r / ((uint64_t)res * res);
    uint32_t m = morton3(x, y, z);
    if (dir == 0) dst[r] = src[m];
    else dst[m] = src[r];
}
void launch_morton_to_rowmajor(uint32_t d, const double* src, double* dst, hipStream_t s) {
    int64_t n = (int64_t)1 << (3 * d);
    hipLaunchKernelGGL(k_m2r, dim3(nblocks_for(n)), dim3(256), 0, s, d, src, dst, 0);
    PRX_LAUNCH_CHECK();
}
void launch_rowmajor_to_morton(uint32_t d, const double* src, double* dst, hipStream_t s) {
    int64_t n = (int64_t)1 << (3 * d);
    hipLaunchKernelGGL(k_m2r, dim3(nblocks_for(n)), dim3(256), 0, s, d, src, dst, 1);
    PRX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ evaluate
// FEMTree::evaluate (src/fem_tree.rs:490-513) with bspline1 (:640-642) and
// contributing_offsets (:645-649); loop order ox, oy, oz and zero-weight skips
// as there, so the result is bit-identical for the same solution vector.
__device__ inline double bspline1(double t) {
    if (t < 0.0 || t > 2.0) return 0.0;
    if (t <= 1.0) return t;
    return 2.0 - t;
}
__device__ inline int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ double eval_field(uint32_t D, const double* __restrict__ sol, double px, double py, double pz) {
    const int64_t res = (int64_t)1 << D;
    const double h = 1.0 / (double)res;
    const double s[3] = {px / h, py / h, pz / h};
    int64_t o[3][2];
    double b[3][2];
    for (int k = 0; k < 3; ++k) {
        int64_t f = sat_i64(floor(s[k]));
        o[k][0] = clampi(wadd(f, -1), 0, res - 1);
        o[k][1] = clampi(f, 0, res - 1);
        b[k][0] = bspline1(s[k] - (double)o[k][0]);
        b[k][1] = bspline1(s[k] - (double)o[k][1]);
    }
    double v = 0.0;
    for (int a = 0; a < 2; ++a) {
        if (b[0][a] == 0.0) continue;
        for (int c = 0; c < 2; ++c) {
            if (b[1][c] == 0.0) continue;
            for (int e = 0; e < 2; ++e) {
                if (b[2][e] == 0.0) continue;
                uint32_t idx = morton3((uint32_t)o[0][a], (uint32_t)o[1][c], (uint32_t)o[2][e]);
                v += sol[idx] * b[0][a] * b[1][c] * b[2][e];
            }
        }
    }
    return v;
}

__global__ void k_evaluate(uint32_t D, const double* __restrict__ sol, const double* __restrict__ q, int64_t m, double* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    out[i] = eval_field(D, sol, q[3 * i], q[3 * i + 1], q[3 * i + 2]);
}
void launch_evaluate(uint32_t D, const double* sol, const double* q, int64_t m, double* out, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_evaluate, dim3(nblocks_for(m)), dim3(256), 0, s, D, sol, q, m, out);
    PRX_LAUNCH_CHECK();
}
__global__ void k_evaluate_soa(uint32_t D, const double* __restrict__ sol, const double* __restrict__ pos, int64_t P, double* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    out[i] = eval_field(D, sol, pos[i], pos[P + i], pos[2 * P + i]);
}
void launch_evaluate_pointsSoA(uint32_t D, const double* sol, const double* pos, int64_t P, double* out, hipStream_t s) {
    if (P <= 0) return;
    hipLaunchKernelGGL(k_evaluate_soa, dim3(nblocks_for(P)), dim3(256), 0, s, D, sol, pos, P, out);
    PRX_LAUNCH_CHECK();
}

// Corner positions are (off + c) * w with w = 1/res (src/fem_tree.rs:596-602).
// evaluate() at the corner k*h (k in [0,res]^3).  There s = k exactly, so
// every basis weight is N1(1) = 1 or N1(0) = 0: the node offset is k-1 on
// each axis, k = 0 gives weight 0 on both offsets (F = 0), and k = res
// clamps both offsets to res-1 with weight 1 (the node counted twice,
// src/fem_tree.rs:645-649).  F is then x at one node added m = prod(1 or 2)
// times in the order evaluate() adds its terms, which reproduces eval_field
// bit for bit with one read per corner.
__global__ void k_evaluate_corners(uint32_t D, const double* __restrict__ sol, double* out) {
    const uint64_t res = 1ull << D, r1 = res + 1;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r1 * r1 * r1) return;
    const uint64_t kx = i % r1, ky = (i / r1) % r1, kz = i / (r1 * r1);
    double v = 0.0;
    if (kx > 0 && ky > 0 && kz > 0) {
        const double x = sol[morton3((uint32_t)(kx - 1), (uint32_t)(ky - 1), (uint32_t)(kz - 1))];
        const int m = (kx == res ? 2 : 1) * (ky == res ? 2 : 1) * (kz == res ? 2 : 1);
        for (int t = 0; t < m; ++t) v += x;
    }
    out[i] = v;
}
// the general evaluate() at the same corners (parity check of the form above)
__global__ void k_evaluate_corners_general(uint32_t D, const double* __restrict__ sol, double* out) {
    const uint64_t r1 = (1ull << D) + 1;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r1 * r1 * r1) return;
    const uint64_t kx = i % r1, ky = (i / r1) % r1, kz = i / (r1 * r1);
    const double w = 1.0 / (double)(1ull << D);
    out[i] = eval_field(D, sol, (double)kx * w, (double)ky * w, (double)kz * w);
}
void launch_evaluate_corners_general(uint32_t D, const double* sol, double* out, hipStream_t s) {
    int64_t r1 = ((int64_t)1 << D) + 1;
    hipLaunchKernelGGL(k_evaluate_corners_general, dim3(nblocks_for(r1 * r1 * r1)), dim3(256), 0, s, D, sol, out);
    PRX_LAUNCH_CHECK();
}
void launch_evaluate_corners(uint32_t D, const double* sol, double* out, hipStream_t s) {
    int64_t r1 = ((int64_t)1 << D) + 1;
    hipLaunchKernelGGL(k_evaluate_corners, dim3(nblocks_for(r1 * r1 * r1)), dim3(256), 0, s, D, sol, out);
    PRX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ generic CSR
// multiply_vector (src/sparse.rs:105-117): one row per thread, summed from 0.0
// in stored column order.
__global__ void k_csr_spmv(int64_t n, const int64_t* __restrict__ off, const int64_t* __restrict__ col,
                           const double* __restrict__ val, const double* __restrict__ x, double* y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t k = off[i]; k < off[i + 1]; ++k) s = s + val[k] * x[col[k]];
    y[i] = s;
}
void launch_csr_spmv(int64_t n, const int64_t* off, const int64_t* col, const double* val, const double* x, double* y, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_csr_spmv, dim3(nblocks_for(n)), dim3(256), 0, s, n, off, col, val, x, y);
    PRX_LAUNCH_CHECK();
}

}  // namespace prx
