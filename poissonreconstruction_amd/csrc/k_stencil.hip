// Matrix-free screened Poisson operator on the Morton-ordered dense grid.
//
// A = L + wg*I + S, with L the 21-point trilinear-hat Laplacian
// (src/fem_tree.rs:44-48, :186-197: centre 8h/3, 12 edges -h/6, 8 corners
// -h/12, faces exactly 0.0, rows truncated at the boundary), wg the global
// screening mass on the diagonal (src/fem_tree.rs:203-204) and S the sum of
// the per-point 8x8 blocks wh3*phi*phi^T (src/fem_tree.rs:206-226).
//
// L is applied in difference form with zero padding outside the grid:
//   (Lx)_i = ce * sum_edges (x_j - x_i) + cc * sum_corners (x_j - x_i),
// which equals the reference rows (interior row sum c0 + 12ce + 8cc = 0; a
// missing neighbour contributes ce*(0 - x_i) or cc*(0 - x_i), i.e. exactly the
// truncated row).  Every term is a difference, so the FP32 operator keeps the
// tiny screening terms that a direct sum would round away.
//
// Tiled kernel: one workgroup per aligned B^3 brick, which is a contiguous
// Morton range.  The brick is read with 16-byte coalesced loads, the one-node
// halo is gathered from the 26 neighbour bricks, all of it is staged in LDS,
// and each thread sweeps one (x,y) column along z with a 3-plane register
// window (9 LDS reads per output).  Outputs go back through LDS so the store
// is coalesced too.  Screening: points were pre-reduced to s_p =
// wh3*sum(phi*x) (k_screen_gather); the brick's touched nodes add
// sum(phi*s_p) in a fixed order, so the result is deterministic.
#include "kernels.hpp"

namespace prx {

template <typename T>
struct LevelArgs {
    uint32_t D, res;
    T ce, cc, wg, c0;  // c0 = -(12ce + 8cc), the diagonal of L implied by the difference form
    int has_screen;
    int brick_log;
    const uint32_t* brick;
    const uint32_t* node;
    const uint32_t* off;
    const uint32_t* sp;
    const T* phi;
    const double* diag;  // sum phi^2 per touched node
    T wh3;
};

template <typename T>
static LevelArgs<T> make_args(const LevelView& L, int tile_log) {
    LevelArgs<T> a{};
    a.D = L.D;
    a.res = L.res;
    if constexpr (sizeof(T) == 8) {
        a.ce = L.k64.ce; a.cc = L.k64.cc; a.wg = L.k64.wg; a.wh3 = L.k64.wh3;
    } else {
        a.ce = L.k32.ce; a.cc = L.k32.cc; a.wg = L.k32.wg; a.wh3 = L.k32.wh3;
    }
    a.c0 = -(T(12) * a.ce + T(8) * a.cc);
    a.has_screen = (L.scr && L.scr->T > 0) ? 1 : 0;
    if (a.has_screen) {
        const NodeList& S = *L.scr;
        a.brick_log = S.brick_log;
        a.brick = S.brick;
        if (tile_log == 9 && S.brick8) { a.brick = S.brick8; a.brick_log = 9; }
        a.node = S.node;
        a.off = S.off;
        a.sp = S.sp;
        if constexpr (sizeof(T) == 8) a.phi = (const T*)S.phi;
        else a.phi = (const T*)S.phi32;
        a.diag = S.diag;
    }
    return a;
}

template <typename T> struct V16;
template <> struct V16<float> { using type = float4; static constexpr int n = 4; };
template <> struct V16<double> { using type = double2; static constexpr int n = 2; };

__device__ inline double block_sum_red(double v, double* wsum) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wsum[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) s += wsum[w];
    return s;
}

// halo shell index -> (hx,hy,hz) in [-1,B]^3 \ [0,B)^3
template <int B>
__device__ inline void shell_pos(int i, int& hx, int& hy, int& hz) {
    constexpr int E = B + 2, P2 = E * E, R = 4 * B + 4;
    if (i < 2 * P2) {
        hz = i < P2 ? -1 : B;
        const int j = i < P2 ? i : i - P2;
        hy = j / E - 1;
        hx = j % E - 1;
    } else {
        i -= 2 * P2;
        hz = i / R;
        int j = i % R;
        if (j < E) { hy = -1; hx = j - 1; }
        else if (j < 2 * E) { hy = B; hx = j - E - 1; }
        else { j -= 2 * E; hy = j >> 1; hx = (j & 1) ? B : -1; }
    }
}

template <typename T, int B, int OP>
__global__ __launch_bounds__(B* B) void k_stencil_tile(LevelArgs<T> L, StencilIO<T> io, uint32_t nbricks, int swz) {
    if (io.done && *io.done) return;
    constexpr int NT = B * B, B3 = B * B * B, E = B + 2, E3 = E * E * E;
    using VT = typename V16<T>::type;
    constexpr int VEC = V16<T>::n, PER = B3 / NT, NV = PER / VEC;
    __shared__ T tile[E3];
    __shared__ double wsum[(NT + 63) / 64];
    const int tid = threadIdx.x;
    uint32_t bid = blockIdx.x;
    if (swz) {
        const uint32_t per = nbricks >> 3;  // XCD-contiguous brick ranges
        bid = (bid & 7u) * per + (bid >> 3);
    }
    const uint64_t base = (uint64_t)bid * B3;
    uint32_t ox, oy, oz;
    demorton3(bid, ox, oy, oz);
    ox *= B; oy *= B; oz *= B;

    // ---- brick interior: 16-byte coalesced loads, scattered into LDS
    T xin[PER];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const uint32_t m = (uint32_t)(v * NT + tid) * VEC;
        if constexpr (OP == OP_RESID_UPD) {
            const double2 xx = *(const double2*)(io.x64 + base + m);
            const float2 ee = *(const float2*)(io.e32 + base + m);
            xin[v * VEC] = xx.x + (double)ee.x;
            xin[v * VEC + 1] = xx.y + (double)ee.y;
        } else {
            const VT q = *(const VT*)(io.x + base + m);
            const T* qq = (const T*)&q;
#pragma unroll
            for (int e = 0; e < VEC; ++e) xin[v * VEC + e] = qq[e];
        }
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            uint32_t lx, ly, lz;
            demorton3(m + e, lx, ly, lz);
            tile[((lz + 1) * E + ly + 1) * E + lx + 1] = xin[v * VEC + e];
        }
    }
    // ---- one-node halo from the neighbouring bricks (zero outside the grid)
    for (int i = tid; i < E3 - B3; i += NT) {
        int hx, hy, hz;
        shell_pos<B>(i, hx, hy, hz);
        const int gx = (int)ox + hx, gy = (int)oy + hy, gz = (int)oz + hz;
        T v = 0;
        if ((unsigned)gx < L.res && (unsigned)gy < L.res && (unsigned)gz < L.res) {
            const uint32_t g = morton3((uint32_t)gx, (uint32_t)gy, (uint32_t)gz);
            if constexpr (OP == OP_RESID_UPD) v = io.x64[g] + (double)io.e32[g];
            else v = io.x[g];
        }
        tile[((hz + 1) * E + hy + 1) * E + hx + 1] = v;
    }
    __syncthreads();

    // ---- z sweep with a 3-plane window
    const int tx = tid % B, ty = tid / B;
    T out[B];
    {
        const T ce = L.ce, cc = L.cc, c4 = T(4) * (L.ce + L.cc), wg = L.wg;
        T Cm, Fm, Dm, Cc, Fc, Dc;
        auto plane = [&](int p, T& C, T& F, T& Dg) {
            const T* r0 = tile + (p * E + ty) * E + tx;
            const T* r1 = r0 + E;
            const T* r2 = r1 + E;
            C = r1[1];
            F = ((r0[1] - C) + (r2[1] - C)) + ((r1[0] - C) + (r1[2] - C));
            Dg = ((r0[0] - C) + (r0[2] - C)) + ((r2[0] - C) + (r2[2] - C));
        };
        plane(0, Cm, Fm, Dm);
        plane(1, Cc, Fc, Dc);
#pragma unroll
        for (int z = 0; z < B; ++z) {
            T Cp, Fp, Dp;
            plane(z + 2, Cp, Fp, Dp);
            T lap = ce * (Dc + Fm + Fp) + cc * (Dm + Dp) + c4 * ((Cm - Cc) + (Cp - Cc));
            out[z] = lap + wg * Cc;
            Cm = Cc; Fm = Fc; Dm = Dc;
            Cc = Cp; Fc = Fp; Dc = Dp;
        }
    }
    __syncthreads();
#pragma unroll
    for (int z = 0; z < B; ++z) tile[morton3((uint32_t)tx, (uint32_t)ty, (uint32_t)z)] = out[z];
    __syncthreads();

    // ---- screening: touched nodes of this brick, fixed summation order
    if (L.has_screen) {
        const uint32_t t0 = L.brick[bid], t1 = L.brick[bid + 1];
        if (t0 < t1) {
            for (uint32_t t = t0 + tid; t < t1; t += NT) {
                const uint32_t loc = (uint32_t)(L.node[t] - base);
                T acc = 0;
                for (uint32_t j = L.off[t]; j < L.off[t + 1]; ++j) acc += L.phi[j] * io.s[L.sp[j]];
                tile[loc] += acc;
            }
            __syncthreads();
        }
    }

    // ---- coalesced output + reduction partial
    double acc = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const uint32_t m = (uint32_t)(v * NT + tid) * VEC;
        if constexpr (OP == OP_APPLY) {
            VT q;
            T* qq = (T*)&q;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                qq[e] = tile[m + e];
                acc += (double)xin[v * VEC + e] * (double)qq[e];
            }
            *(VT*)(io.y + base + m) = q;
        } else if constexpr (OP == OP_RESID) {
            const VT bb = *(const VT*)(io.b + base + m);
            const T* bq = (const T*)&bb;
            VT q;
            T* qq = (T*)&q;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                qq[e] = bq[e] - tile[m + e];
                acc += (double)qq[e] * (double)qq[e];
            }
            *(VT*)(io.y + base + m) = q;
        } else {
            const double2 bb = *(const double2*)((const double*)io.b + base + m);
            double2 xo;
            xo.x = xin[v * VEC];
            xo.y = xin[v * VEC + 1];
            *(double2*)(io.x_out + base + m) = xo;
            const double r0 = bb.x - tile[m], r1 = bb.y - tile[m + 1];
            float2 rr;
            rr.x = (float)(r0 * io.rscale);
            rr.y = (float)(r1 * io.rscale);
            *(float2*)(io.r32 + base + m) = rr;
            acc += r0 * r0 + r1 * r1;
        }
    }
    const double s = block_sum_red(acc, wsum);
    if (tid == 0 && io.partial) io.partial[bid] = s;
}

// ------------------------------------------------------------------ per-node form (small grids, GS)
template <typename T>
__device__ inline int64_t touched_index(const LevelArgs<T>& L, uint32_t i) {
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

// (A x)_i with x read from global memory; also returns the diagonal a_ii.
template <typename T, int MODE>
__device__ inline T apply_node(const LevelArgs<T>& L, const T* x, const double* x64, const float* e32, const T* s, uint32_t ix,
                               uint32_t iy, uint32_t iz, uint32_t i, T* diag) {
    auto X = [&](int jx, int jy, int jz) -> T {
        if ((unsigned)jx >= L.res || (unsigned)jy >= L.res || (unsigned)jz >= L.res) return T(0);
        const uint32_t j = morton3((uint32_t)jx, (uint32_t)jy, (uint32_t)jz);
        if constexpr (MODE == 1) return (T)(x64[j] + (double)e32[j]);
        else return x[j];
    };
    const T xc = X(ix, iy, iz);
    T se = 0, sc = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nz = (dx != 0) + (dy != 0) + (dz != 0);
                if (nz < 2) continue;
                const T d = X((int)ix + dx, (int)iy + dy, (int)iz + dz) - xc;
                if (nz == 2) se += d;
                else sc += d;
            }
    T a = L.ce * se + L.cc * sc + L.wg * xc;
    T dg = L.c0 + L.wg;
    if (L.has_screen) {
        const int64_t t = touched_index(L, i);
        if (t >= 0) {
            T acc = 0;
            for (uint32_t j = L.off[t]; j < L.off[t + 1]; ++j) acc += L.phi[j] * s[L.sp[j]];
            a += acc;
            dg += L.wh3 * (T)L.diag[t];
        }
    }
    if (diag) *diag = dg;
    return a;
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_stencil_node(LevelArgs<T> L, StencilIO<T> io, uint32_t N) {
    __shared__ double wsum[4];
    if (io.done && *io.done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (i < N) {
        uint32_t ix, iy, iz;
        demorton3(i, ix, iy, iz);
        if constexpr (OP == OP_RESID_UPD) {
            const double a = apply_node<T, 1>(L, nullptr, io.x64, io.e32, io.s, ix, iy, iz, i, nullptr);
            io.x_out[i] = io.x64[i] + (double)io.e32[i];
            const double r = io.b[i] - a;
            io.r32[i] = (float)(r * io.rscale);
            acc = r * r;
        } else {
            const T a = apply_node<T, 0>(L, io.x, nullptr, nullptr, io.s, ix, iy, iz, i, nullptr);
            if constexpr (OP == OP_APPLY) {
                io.y[i] = a;
                acc = (double)io.x[i] * (double)a;
            } else {
                const T r = io.b[i] - a;
                io.y[i] = r;
                acc = (double)r * (double)r;
            }
        }
    }
    const double sum = block_sum_red(acc, wsum);
    if (threadIdx.x == 0 && io.partial) io.partial[blockIdx.x] = sum;
}

template <typename T>
void launch_stencil(const LevelView& Lv, int op, const StencilIO<T>& io, int* nblk, int tile, hipStream_t s) {
    const int64_t N = Lv.N;
    int B = 0;
    if (Lv.D >= 4) B = (tile == 8) ? 8 : 16;
    else if (Lv.D == 3) B = 8;
    if (B == 0) {
        LevelArgs<T> a = make_args<T>(Lv, 12);
        const int nb = (int)((N + 255) / 256);
        switch (op) {
            case OP_APPLY: hipLaunchKernelGGL((k_stencil_node<T, OP_APPLY>), dim3(nb), dim3(256), 0, s, a, io, (uint32_t)N); break;
            case OP_RESID: hipLaunchKernelGGL((k_stencil_node<T, OP_RESID>), dim3(nb), dim3(256), 0, s, a, io, (uint32_t)N); break;
            default:
                if constexpr (sizeof(T) == 8)
                    hipLaunchKernelGGL((k_stencil_node<T, OP_RESID_UPD>), dim3(nb), dim3(256), 0, s, a, io, (uint32_t)N);
        }
        PRX_LAUNCH_CHECK();
        *nblk = nb;
        return;
    }
    const int tl = B == 16 ? 12 : 9;
    LevelArgs<T> a = make_args<T>(Lv, tl);
    const uint32_t nbricks = (uint32_t)(N >> tl);
    const int swz = (nbricks >= 64 && (nbricks & 7u) == 0) ? 1 : 0;
#define PRX_TILE_LAUNCH(BB, OPV) \
    hipLaunchKernelGGL((k_stencil_tile<T, BB, OPV>), dim3(nbricks), dim3(BB * BB), 0, s, a, io, nbricks, swz)
    if (B == 16) {
        if (op == OP_APPLY) PRX_TILE_LAUNCH(16, OP_APPLY);
        else if (op == OP_RESID) PRX_TILE_LAUNCH(16, OP_RESID);
        else if constexpr (sizeof(T) == 8) PRX_TILE_LAUNCH(16, OP_RESID_UPD);
    } else {
        if (op == OP_APPLY) PRX_TILE_LAUNCH(8, OP_APPLY);
        else if (op == OP_RESID) PRX_TILE_LAUNCH(8, OP_RESID);
        else if constexpr (sizeof(T) == 8) PRX_TILE_LAUNCH(8, OP_RESID_UPD);
    }
#undef PRX_TILE_LAUNCH
    PRX_LAUNCH_CHECK();
    *nblk = (int)nbricks;
}
template void launch_stencil<double>(const LevelView&, int, const StencilIO<double>&, int*, int, hipStream_t);
template void launch_stencil<float>(const LevelView&, int, const StencilIO<float>&, int*, int, hipStream_t);

// ------------------------------------------------------------------ multicolour Gauss-Seidel
// gauss_seidel_sweep (src/solvers.rs:164-175) recoloured: the 27-point
// stencil and the 2x2x2 screening blocks couple only nodes whose coordinate
// parities differ, so the 8 parity classes are each updated in parallel,
// colour by colour: x_i += (b_i - (Ax)_i) / a_ii.
template <typename T>
__global__ __launch_bounds__(256) void k_gs_color(LevelArgs<T> L, int color, T* x, const T* b, const T* s, uint32_t nper) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nper) return;
    uint32_t X, Y, Z;
    demorton3(j, X, Y, Z);
    const uint32_t ix = 2 * X + (color & 1), iy = 2 * Y + ((color >> 1) & 1), iz = 2 * Z + (color >> 2);
    if (ix >= L.res || iy >= L.res || iz >= L.res) return;
    const uint32_t i = morton3(ix, iy, iz);
    T dg;
    const T a = apply_node<T, 0>(L, x, nullptr, nullptr, s, ix, iy, iz, i, &dg);
    if (dg != T(0)) x[i] = x[i] + (b[i] - a) / dg;
}
template <typename T>
void launch_gs_color(const LevelView& Lv, int color, T* x, const T* b, const T* sbuf, hipStream_t s) {
    LevelArgs<T> a = make_args<T>(Lv, 12);
    const uint32_t half = Lv.res >= 2 ? Lv.res / 2 : 1;
    const uint32_t nper = half * half * half;
    hipLaunchKernelGGL((k_gs_color<T>), dim3((nper + 255) / 256), dim3(256), 0, s, a, color, x, b, sbuf, nper);
    PRX_LAUNCH_CHECK();
}
template void launch_gs_color<double>(const LevelView&, int, double*, const double*, const double*, hipStream_t);
template void launch_gs_color<float>(const LevelView&, int, float*, const float*, const float*, hipStream_t);

}  // namespace prx
