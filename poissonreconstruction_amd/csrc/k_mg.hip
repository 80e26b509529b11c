// Grid transfers and the coarsest solve of the geometric multigrid
// preconditioner (a V-cycle over the dense Morton grids of depths dc..D).
//
// The hat basis of depth d-1 is spanned by the hats of depth d:
// phi^{d-1}_I = sum_i P_iI phi^d_i with, per axis, weight 1 at the fine node
// 2I and 1/2 at 2I-1 and 2I+1 (trilinear interpolation).  Nodes outside
// [0, res) do not exist on either grid (the operator's rows are truncated
// there), so their terms are dropped.  Restriction is exactly P^T, which
// keeps the V-cycle symmetric.
#include <algorithm>

#include "kernels.hpp"

namespace prx {

template <typename T>
__global__ void k_mg_scale(T* __restrict__ x, const T* __restrict__ b, double w, int64_t n, const int32_t* done) {
    if (done && *done) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (T)(w * (double)b[i]);
}
template <typename T>
void launch_mg_scale(T* x, const T* b, double w, int64_t n, hipStream_t s, const int32_t* done) {
    const int64_t nb = std::min<int64_t>(65536, std::max<int64_t>(1, (n + 1023) / 1024));
    hipLaunchKernelGGL(k_mg_scale<T>, dim3((unsigned)nb), dim3(256), 0, s, x, b, w, n, done);
    PRX_LAUNCH_CHECK();
}

// b_c[I] = sum_{delta in {-1,0,1}^3} 2^-(number of nonzero delta) r_f[2I + delta]
template <typename T>
__global__ void k_mg_restrict(uint32_t dc, const T* __restrict__ rf, T* __restrict__ bc, const int32_t* done) {
    if (done && *done) return;
    const uint32_t nc = 1u << (3 * dc);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int fres = 2 << dc;
    uint32_t X, Y, Z;
    demorton3(i, X, Y, Z);
    double acc = 0.0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int fz = 2 * (int)Z + dz;
        if ((unsigned)fz >= (unsigned)fres) continue;
        const double wz = dz ? 0.5 : 1.0;
        for (int dy = -1; dy <= 1; ++dy) {
            const int fy = 2 * (int)Y + dy;
            if ((unsigned)fy >= (unsigned)fres) continue;
            const double wy = dy ? 0.5 : 1.0;
            for (int dx = -1; dx <= 1; ++dx) {
                const int fx = 2 * (int)X + dx;
                if ((unsigned)fx >= (unsigned)fres) continue;
                const double w = (dx ? 0.5 : 1.0) * wy * wz;
                acc += w * (double)rf[morton3((uint32_t)fx, (uint32_t)fy, (uint32_t)fz)];
            }
        }
    }
    bc[i] = (T)acc;
}
// Tiled form for fine grids of 16^3 bricks: a workgroup takes one fine brick
// (a contiguous Morton range) plus the one-node layer below it in x, y and z
// into LDS, and writes the aligned 8^3 coarse brick (also contiguous).
template <typename T>
__global__ __launch_bounds__(256) void k_mg_restrict_tile(uint32_t df, const T* __restrict__ rf, T* __restrict__ bc,
                                                          const int32_t* done) {
    if (done && *done) return;
    constexpr int E = 17;  // local fine coordinates -1..15
    __shared__ T f[E * E * E];
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T), NVEC = 4096 / (256 * VEC);
    const int tid = threadIdx.x;
    const uint32_t fb = blockIdx.x;
    const uint64_t base = (uint64_t)fb << 12;
    uint32_t bx, by, bz;
    demorton3(fb, bx, by, bz);
    const int ox = (int)bx * 16, oy = (int)by * 16, oz = (int)bz * 16;
#pragma unroll
    for (int v = 0; v < NVEC; ++v) {
        const uint32_t m = (uint32_t)(v * 256 + tid) * VEC;
        const NV q = *(const NV*)(rf + base + m);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            uint32_t lx, ly, lz;
            demorton3(m + (uint32_t)e, lx, ly, lz);
            f[((lz + 1) * E + ly + 1) * E + lx + 1] = q[e];
        }
    }
    // the lower layer: local coordinate -1 in at least one axis (17^3 - 16^3 nodes)
    for (int i = tid; i < E * E * E - 4096; i += 256) {
        // enumerate the layer as the E^3 cube minus the 16^3 block above it
        int j = i, hx, hy, hz;
        if (j < E * E) { hz = -1; hy = j / E - 1; hx = j % E - 1; }
        else {
            j -= E * E;  // hz in 0..15: the 16 x 17 + 17 - ... rows with hx or hy = -1
            hz = j / (2 * E - 1);
            const int k = j % (2 * E - 1);
            if (k < E) { hy = -1; hx = k - 1; } else { hx = -1; hy = k - E; }
        }
        const int gx = ox + hx, gy = oy + hy, gz = oz + hz;
        T v = T(0);
        if (gx >= 0 && gy >= 0 && gz >= 0) v = rf[morton3((uint32_t)gx, (uint32_t)gy, (uint32_t)gz)];
        f[((hz + 1) * E + hy + 1) * E + hx + 1] = v;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const uint32_t m = (uint32_t)(u * 256 + tid);
        uint32_t cx, cy, cz;
        demorton3(m, cx, cy, cz);
        // fine local 2c-1..2c+1 -> LDS index 2c..2c+2
        double acc = 0.0;
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const double w = (dx == 1 ? 1.0 : 0.5) * (dy == 1 ? 1.0 : 0.5) * (dz == 1 ? 1.0 : 0.5);
                    acc += w * (double)f[((2 * cz + dz) * E + 2 * cy + dy) * E + 2 * cx + dx];
                }
        bc[((uint64_t)fb << 9) + m] = (T)acc;
    }
    (void)df;
}

template <typename T>
void launch_mg_restrict(uint32_t dc, const T* rf, T* bc, hipStream_t s, const int32_t* done) {
    const int64_t nc = (int64_t)1 << (3 * dc);
    if (dc >= 3) {  // fine grid of at least 16^3
        hipLaunchKernelGGL(k_mg_restrict_tile<T>, dim3((unsigned)(nc >> 9)), dim3(256), 0, s, dc + 1, rf, bc, done);
    } else {
        hipLaunchKernelGGL(k_mg_restrict<T>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, dc, rf, bc, done);
    }
    PRX_LAUNCH_CHECK();
}

// x_f[i] += sum_I P_iI x_c[I]
template <typename T>
__global__ void k_mg_prolong_add(uint32_t df, const T* __restrict__ xc, T* __restrict__ xf, const int32_t* done) {
    if (done && *done) return;
    const uint32_t nf = 1u << (3 * df);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const uint32_t cres = 1u << (df - 1);
    uint32_t f[3];
    demorton3(i, f[0], f[1], f[2]);
    uint32_t c[3][2];
    double w[3][2];
    int nc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if ((f[a] & 1u) == 0) {
            c[a][0] = f[a] >> 1;
            w[a][0] = 1.0;
            nc[a] = 1;
        } else {
            c[a][0] = (f[a] - 1) >> 1;
            w[a][0] = 0.5;
            c[a][1] = (f[a] + 1) >> 1;
            w[a][1] = 0.5;
            nc[a] = c[a][1] < cres ? 2 : 1;
        }
    }
    double acc = 0.0;
    for (int iz = 0; iz < nc[2]; ++iz)
        for (int iy = 0; iy < nc[1]; ++iy)
            for (int ix = 0; ix < nc[0]; ++ix)
                acc += (w[0][ix] * w[1][iy] * w[2][iz]) * (double)xc[morton3(c[0][ix], c[1][iy], c[2][iz])];
    xf[i] = (T)((double)xf[i] + acc);
}
// Tiled form: a workgroup takes the 8^3 coarse brick under one 16^3 fine
// brick plus the coarse layer above it (coordinate 8; absent at the grid's
// upper end) into LDS and updates the fine brick with 16-byte accesses.
template <typename T>
__global__ __launch_bounds__(256) void k_mg_prolong_tile(uint32_t df, const T* __restrict__ xc, T* __restrict__ xf,
                                                         const int32_t* done) {
    if (done && *done) return;
    constexpr int E = 9;  // local coarse coordinates 0..8
    __shared__ T c[E * E * E];
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T), NVEC = 4096 / (256 * VEC);
    const int tid = threadIdx.x;
    const uint32_t fb = blockIdx.x;
    const uint32_t cres = 1u << (df - 1);
    uint32_t bx, by, bz;
    demorton3(fb, bx, by, bz);
    const uint32_t ox = bx * 8, oy = by * 8, oz = bz * 8;
    for (int i = tid; i < E * E * E; i += 256) {
        const int lx = i % E, ly = (i / E) % E, lz = i / (E * E);
        const uint32_t gx = ox + (uint32_t)lx, gy = oy + (uint32_t)ly, gz = oz + (uint32_t)lz;
        T v = T(0);
        if (gx < cres && gy < cres && gz < cres) v = xc[morton3(gx, gy, gz)];
        c[i] = v;
    }
    __syncthreads();
    const uint64_t base = (uint64_t)fb << 12;
#pragma unroll
    for (int v = 0; v < NVEC; ++v) {
        const uint32_t m = (uint32_t)(v * 256 + tid) * VEC;
        NV q = *(const NV*)(xf + base + m);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            uint32_t fx, fy, fz;
            demorton3(m + (uint32_t)e, fx, fy, fz);
            // per axis: even fine -> coarse f/2 (weight 1); odd -> f/2 and f/2+1 (1/2 each)
            const int x0 = (int)(fx >> 1), y0 = (int)(fy >> 1), z0 = (int)(fz >> 1);
            const int nx = (int)(fx & 1u), ny = (int)(fy & 1u), nz = (int)(fz & 1u);
            double acc = 0.0;
            for (int dz = 0; dz <= nz; ++dz)
                for (int dy = 0; dy <= ny; ++dy)
                    for (int dx = 0; dx <= nx; ++dx) {
                        const double w = (nx ? 0.5 : 1.0) * (ny ? 0.5 : 1.0) * (nz ? 0.5 : 1.0);
                        acc += w * (double)c[((z0 + dz) * E + y0 + dy) * E + x0 + dx];
                    }
            q[e] = (T)((double)q[e] + acc);
        }
        *(NV*)(xf + base + m) = q;
    }
}

template <typename T>
void launch_mg_prolong_add(uint32_t df, const T* xc, T* xf, hipStream_t s, const int32_t* done) {
    const int64_t nf = (int64_t)1 << (3 * df);
    if (df >= 4) {
        hipLaunchKernelGGL(k_mg_prolong_tile<T>, dim3((unsigned)(nf >> 12)), dim3(256), 0, s, df, xc, xf, done);
    } else {
        hipLaunchKernelGGL(k_mg_prolong_add<T>, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, df, xc, xf, done);
    }
    PRX_LAUNCH_CHECK();
}

// x = Ainv b, one workgroup; b staged in LDS, FP64 products
template <typename T>
__global__ __launch_bounds__(256) void k_mg_dense(int n, const double* __restrict__ ainv, const T* __restrict__ b, T* __restrict__ x,
                                                  const int32_t* done) {
    __shared__ double sb[512];
    if (done && *done) return;
    for (int j = threadIdx.x; j < n; j += blockDim.x) sb[j] = (double)b[j];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double* row = ainv + (int64_t)i * n;
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += row[j] * sb[j];
        x[i] = (T)acc;
    }
}
template <typename T>
void launch_mg_dense(int n, const double* ainv, const T* b, T* x, hipStream_t s, const int32_t* done) {
    if (n > 512) throw Error(PRX_ERR_INVALID, "dense coarse solve limited to 512 nodes");
    hipLaunchKernelGGL(k_mg_dense<T>, dim3(1), dim3(256), 0, s, n, ainv, b, x, done);
    PRX_LAUNCH_CHECK();
}

template void launch_mg_scale<double>(double*, const double*, double, int64_t, hipStream_t, const int32_t*);
template void launch_mg_scale<float>(float*, const float*, double, int64_t, hipStream_t, const int32_t*);
template void launch_mg_restrict<double>(uint32_t, const double*, double*, hipStream_t, const int32_t*);
template void launch_mg_restrict<float>(uint32_t, const float*, float*, hipStream_t, const int32_t*);
template void launch_mg_prolong_add<double>(uint32_t, const double*, double*, hipStream_t, const int32_t*);
template void launch_mg_prolong_add<float>(uint32_t, const float*, float*, hipStream_t, const int32_t*);
template void launch_mg_dense<double>(int, const double*, const double*, double*, hipStream_t, const int32_t*);
template void launch_mg_dense<float>(int, const double*, const float*, float*, hipStream_t, const int32_t*);

}  // namespace prx
