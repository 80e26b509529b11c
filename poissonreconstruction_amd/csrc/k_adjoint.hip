// Adjoint (vjp) kernels: gradient of a loss on the field at the splatted
// points with respect to the points' positions and normals.  The forward path
// is the reference's: splat (src/fem_tree.rs:130-178), RHS and screening
// (src/fem_tree.rs:199-299), solve, evaluate (src/fem_tree.rs:490-513).  With
// A symmetric, dL/dtheta = gF.dF/dtheta + lam^T (db/dtheta - dA/dtheta x),
// A lam = dL/dx; the kernels below form dL/dx, the nf-adjoint of the RHS
// stencil, and the per-point terms.  Weight derivatives are taken on the
// open pieces of the piecewise-linear weights.
#include "kernels.hpp"

namespace prx {

static inline int nblk(int64_t n) { return (int)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

__global__ void k_adj_load(int64_t T, const uint32_t* __restrict__ node, const uint32_t* __restrict__ off,
                           const uint32_t* __restrict__ code, const double* __restrict__ phi, const double* __restrict__ gF,
                           double* gx) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    double acc = 0.0;
    for (uint32_t j = off[t]; j < off[t + 1]; ++j) acc += gF[code[j] >> 3] * phi[j];
    gx[node[t]] = acc;
}
void launch_adj_load(const NodeList& E, const double* gF, double* gx, hipStream_t s) {
    if (E.T <= 0) return;
    hipLaunchKernelGGL(k_adj_load, dim3(nblk(E.T)), dim3(256), 0, s, E.T, E.node, E.off, E.code, E.phi, gF, gx);
    PRX_LAUNCH_CHECK();
}

__device__ inline double mass1(int d) { return d == 0 ? 2. / 3. : 1. / 6.; }
__device__ inline double gmass1(int d) { return d == 0 ? 0. : (d == 1 ? -0.5 : 0.5); }

__global__ void k_adj_gnf(uint32_t D, int64_t T, const uint32_t* __restrict__ node, const double* __restrict__ lam, double* gnf) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const int64_t res = (int64_t)1 << D;
    const double h = 1.0 / (double)res;
    uint32_t jx, jy, jz;
    demorton3(node[t], jx, jy, jz);
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const int64_t ix = (int64_t)jx - dx, iy = (int64_t)jy - dy, iz = (int64_t)jz - dz;
                if (ix < 0 || ix >= res || iy < 0 || iy >= res || iz < 0 || iz >= res) continue;
                const double l = lam[morton3((uint32_t)ix, (uint32_t)iy, (uint32_t)iz)] * (h * h);
                g0 += l * gmass1(dx) * mass1(dy) * mass1(dz);
                g1 += l * mass1(dx) * gmass1(dy) * mass1(dz);
                g2 += l * mass1(dx) * mass1(dy) * gmass1(dz);
            }
    gnf[3 * t] = g0;
    gnf[3 * t + 1] = g1;
    gnf[3 * t + 2] = g2;
}
void launch_adj_gnf(uint32_t D, const NodeList& F, const double* lam, double* gnf, hipStream_t s) {
    if (F.T <= 0) return;
    hipLaunchKernelGGL(k_adj_gnf, dim3(nblk(F.T)), dim3(256), 0, s, D, F.T, F.node, lam, gnf);
    PRX_LAUNCH_CHECK();
}

__device__ inline double dbspline1(double t) {
    if (t > 0.0 && t < 1.0) return 1.0;
    if (t > 1.0 && t < 2.0) return -1.0;
    return 0.0;
}
__device__ inline double bsp1(double t) {
    if (t < 0.0 || t > 2.0) return 0.0;
    if (t <= 1.0) return t;
    return 2.0 - t;
}
__device__ inline int64_t clampr(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline int64_t find_node(const NodeList& L, uint32_t i) {
    if (L.T == 0) return -1;
    const uint32_t b = i >> L.brick_log;
    int64_t lo = L.brick[b], hi = L.brick[b + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        const uint32_t v = L.node[mid];
        if (v == i) return mid;
        if (v < i) lo = mid + 1;
        else hi = mid;
    }
    return -1;
}

__global__ void k_adj_points(uint32_t D, int64_t P, const double* __restrict__ pos, const double* __restrict__ nrm, NodeList F,
                             const double* __restrict__ gF, const double* __restrict__ x, const double* __restrict__ lam,
                             const double* __restrict__ nf, const double* __restrict__ w, const double* __restrict__ gnf,
                             double wh3, double* gpos, double* gnrm) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int64_t res = (int64_t)1 << D, N = res * res * res;
    const double h = 1.0 / (double)res;
    const double q[3] = {pos[p], pos[P + p], pos[2 * P + p]};
    const double nv[3] = {nrm[p], nrm[P + p], nrm[2 * P + p]};
    double gq[3] = {0.0, 0.0, 0.0}, gn[3] = {0.0, 0.0, 0.0};
    // explicit dF/dq at the evaluate corners
    {
        int64_t o[3][2];
        double b[3][2], db[3][2];
        for (int a = 0; a < 3; ++a) {
            const double s = q[a] / h;
            const int64_t f = sat_i64(floor(s));
            o[a][0] = clampr(wadd(f, -1), 0, res - 1);
            o[a][1] = clampr(f, 0, res - 1);
            for (int k = 0; k < 2; ++k) {
                b[a][k] = bsp1(s - (double)o[a][k]);
                db[a][k] = dbspline1(s - (double)o[a][k]) / h;
            }
        }
        const double g = gF[p];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) {
                    const double xv = x[morton3((uint32_t)o[0][i], (uint32_t)o[1][j], (uint32_t)o[2][k])];
                    gq[0] += g * xv * db[0][i] * b[1][j] * b[2][k];
                    gq[1] += g * xv * b[0][i] * db[1][j] * b[2][k];
                    gq[2] += g * xv * b[0][i] * b[1][j] * db[2][k];
                }
    }
    // splat (floor cell) and screening (trunc cell) corners
    for (int mode = 0; mode < 2; ++mode) {
        if (mode == 1 && !(wh3 > 0.0)) break;
        int64_t c[3];
        double t[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = mode == 0 ? sat_i64(floor(q[a] / h)) : sat_i64(q[a] / h);
            t[a] = q[a] / h - (double)c[a];
        }
        double px = 0.0, pl = 0.0, dpx[3] = {0.0, 0.0, 0.0}, dpl[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 8; ++k) {
            const int d[3] = {k & 1, (k >> 1) & 1, k >> 2};
            uint32_t o[3];
            bool ok = true;
            for (int a = 0; a < 3; ++a) {
                if (mode == 0) {
                    const uint64_t u = (uint64_t)wadd(c[a], d[a]);
                    ok = ok && u < (uint64_t)res;
                    o[a] = (uint32_t)u;
                } else {
                    const uint32_t u = (uint32_t)(uint64_t)wadd(c[a], d[a]);
                    ok = ok && u < (uint32_t)res;
                    o[a] = u;
                }
            }
            if (!ok) continue;
            double bb[3], dbb[3];
            for (int a = 0; a < 3; ++a) {
                bb[a] = d[a] ? t[a] : 1.0 - t[a];
                dbb[a] = (d[a] ? 1.0 : -1.0) / h;
            }
            const double wk = bb[0] * bb[1] * bb[2];
            const double dw[3] = {dbb[0] * bb[1] * bb[2], bb[0] * dbb[1] * bb[2], bb[0] * bb[1] * dbb[2]};
            const uint32_t i = morton3(o[0], o[1], o[2]);
            if (mode == 0) {
                const double Wi = w[i];
                if (!(Wi > 0.0)) continue;
                const int64_t ti = find_node(F, i);
                if (ti < 0) continue;
                const double g3[3] = {gnf[3 * ti], gnf[3 * ti + 1], gnf[3 * ti + 2]};
                double sdot = 0.0;
                for (int cc = 0; cc < 3; ++cc) {
                    gn[cc] += -(wk / Wi) * g3[cc];
                    sdot += g3[cc] * (-(nv[cc] + nf[(int64_t)cc * N + i]) / Wi);
                }
                for (int a = 0; a < 3; ++a) gq[a] += sdot * dw[a];
            } else {
                const double xv = x[i], lv = lam[i];
                px += wk * xv;
                pl += wk * lv;
                for (int a = 0; a < 3; ++a) {
                    dpx[a] += dw[a] * xv;
                    dpl[a] += dw[a] * lv;
                }
            }
        }
        if (mode == 1)
            for (int a = 0; a < 3; ++a) gq[a] += wh3 * 0.5 * dpl[a] - wh3 * (dpl[a] * px + pl * dpx[a]);
    }
    for (int a = 0; a < 3; ++a) {
        gpos[3 * p + a] = gq[a];
        gnrm[3 * p + a] = gn[a];
    }
}
void launch_adj_points(uint32_t D, const DevPoints& pts, const NodeList& F, const double* gF, const double* x,
                       const double* lam, const double* nf, const double* w, const double* gnf, double wh3, double* gpos,
                       double* gnrm, hipStream_t s) {
    if (pts.P <= 0) return;
    hipLaunchKernelGGL(k_adj_points, dim3(nblk(pts.P)), dim3(256), 0, s, D, pts.P, pts.pos, pts.nrm, F, gF, x, lam, nf, w,
                       gnf, wh3, gpos, gnrm);
    PRX_LAUNCH_CHECK();
}

}  // namespace prx
