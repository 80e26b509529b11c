// Matrix-free screened Poisson operator on the Morton-ordered dense grid.
//
// A = L + wg*I + S, with L the 21-point trilinear-hat Laplacian
// (src/fem_tree.rs:44-48, :186-197: centre 8h/3, 12 edges -h/6, 8 corners
// -h/12, faces exactly 0.0, rows truncated at the boundary), wg the global
// screening mass on the diagonal (src/fem_tree.rs:203-204) and S the sum of
// the per-point 8x8 blocks wh3*phi*phi^T (src/fem_tree.rs:206-226).
//
// L is applied with exact integer weights: ce = -2s, cc = -s, c0 = 32s with
// s = h/12, so every row of the zero-padded 27-point stencil sums to exactly
// zero, and (Lx)_i is unchanged when a constant is subtracted from x.  Each
// brick therefore works on x - r0, r0 = the brick's first node value, with
// nodes outside the grid stored as 0 - r0 (this reproduces the reference's
// truncated boundary rows).  Rounding then scales with the variation of x
// inside the brick instead of |x|, which keeps the tiny screening terms
// resolvable in FP32.  The weights equal the reference's K(d) to 1 ulp.
//
// Tiled kernel: one workgroup per aligned B^3 brick, which is a contiguous
// Morton range.  The brick is read with 16-byte coalesced loads, the one-node
// halo is gathered from the 26 neighbour bricks, all of it is staged in LDS
// (interior LDS offsets are linear in the Morton bits, so no per-element
// decode), and each thread sweeps one (x,y) column along z with a 3-plane
// register window (9 LDS reads per output).  Outputs go back through LDS so the store
// is coalesced too, and leave as non-temporal stores (measured on MI355X at 2^27 nodes:
// 2-3% faster than plain stores; non-temporal interior loads instead lose 17-20%, since
// they evict the lines that neighbouring bricks read as halo).  Screening: points were pre-reduced to s_p =
// wh3*sum(phi*x) (k_screen_gather); the brick's touched nodes add
// sum(phi*s_p) in a fixed order, so the result is deterministic.
//
// The operand is formed per node as the pass needs it: x (apply, residual,
// smoothing), x64 + e32 (the mixed solve's FP64 refinement), r + beta d_old
// (the two-pass CG direction) or, for the fused CG step, (r - alpha q) +
// beta d_old, with the same expression for interior, halo and gathered
// values so that every brick sees one value per node.  Workgroups go to the
// XCDs in chunks of consecutive bricks (xcd_chunk).
#include "kernels.hpp"

namespace prx {

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
    a.s = -a.cc;
    a.ce = T(-2) * a.s;
    a.c0 = T(32) * a.s;
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
        if (S.spc && L.pts && L.pts->P > 0) {
            a.spc = S.spc;
            if constexpr (sizeof(T) == 8) a.tt = (const T*)L.pts->t64;
            else a.tt = (const T*)L.pts->t32;
            a.P = L.pts->P;
        }
    }
    return a;
}

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

// NS block sums at once (each folded in wave order, as block_sum_red)
template <int NS>
__device__ inline void block_sum_redn(const double* v, double* wsum, double* out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)((blockDim.x + 63) >> 6);
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        double u = v[j];
        for (int o = 32; o > 0; o >>= 1) u += __shfl_xor(u, o, 64);
        if (lane == 0) wsum[j * nw + wave] = u;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        double s = 0.0;
        if (threadIdx.x == 0)
            for (int w = 0; w < nw; ++w) s += wsum[j * nw + w];
        out[j] = s;
    }
}

// ------------------------------------------------------------------ tile index tables
// Morton bit helpers usable at compile time.
constexpr uint32_t spread3c(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}
constexpr uint32_t compact3c(uint32_t m) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((m >> (3 * b)) & 1u) << b;
    return r;
}
// LDS offset of local Morton index m in an E^3 tile (before the +1 halo
// shift).  It is linear in the bits of m, so the offset of m = a | b | c with
// disjoint bit fields is the sum of the three offsets: each thread computes
// its own part once and adds compile-time constants per element.
template <int E>
constexpr int lds_lin(uint32_t m) {
    return (int)(compact3c(m) + E * compact3c(m >> 1) + E * E * compact3c(m >> 2));
}

// LDS placement of the output staging (flat local Morton index m): 4-element
// groups stay contiguous (16-byte reads), and for 16^3 bricks Morton bits 6
// and 9 are XORed into the group index so that the column-wise writes of a
// half-wave (x bits 0,3,6,9 and the row pair) spread over all 32 banks.
template <int B>
__device__ __forceinline__ uint32_t stage_at(uint32_t m) {
    if constexpr (B == 16) return m ^ ((((m >> 6) & 1u) | (((m >> 9) & 1u) << 2)) << 2);
    else return m;
}
// thread -> (x,y) column of the z sweep.  For 16^3 bricks the two rows read by
// one half-wave are 8 apart: with the 18-word row pitch they start 144 = 16
// (mod 32) banks apart, so the 9 window reads of a z step are conflict-free.
template <int B>
__device__ __forceinline__ void column_of(int tid, int& tx, int& ty) {
    if constexpr (B == 16) {
        tx = tid & 15;
        const int q = tid >> 4;
        ty = ((q & 1) << 3) | (q >> 1);
    } else {
        tx = tid % B;
        ty = tid / B;
    }
}

// The x neighbours of a 16^3 brick's column sweep come from the adjacent
// lanes: with column_of<16>, lane & 15 is x, so each 16-lane DPP row holds
// one x row.  row_from_left(v, edge) is v of lane x - 1 (row_shr:1) and
// row_from_right v of lane x + 1 (row_shl:1); lane 0 / 15 of the row gets
// `edge` (the halo column) instead.
template <typename T>
__device__ __forceinline__ T dpp_row(T v, T edge, int ctrl);
template <>
__device__ __forceinline__ float dpp_row<float>(float v, float edge, int ctrl) {
    const int e = __builtin_bit_cast(int, edge), s = __builtin_bit_cast(int, v);
    return ctrl == 0x111 ? __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(e, s, 0x111, 0xF, 0xF, false))
                         : __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(e, s, 0x101, 0xF, 0xF, false));
}
template <>
__device__ __forceinline__ double dpp_row<double>(double v, double edge, int ctrl) {
    const long long e = __builtin_bit_cast(long long, edge), s = __builtin_bit_cast(long long, v);
    int lo, hi;
    if (ctrl == 0x111) {
        lo = __builtin_amdgcn_update_dpp((int)e, (int)s, 0x111, 0xF, 0xF, false);
        hi = __builtin_amdgcn_update_dpp((int)(e >> 32), (int)(s >> 32), 0x111, 0xF, 0xF, false);
    } else {
        lo = __builtin_amdgcn_update_dpp((int)e, (int)s, 0x101, 0xF, 0xF, false);
        hi = __builtin_amdgcn_update_dpp((int)(e >> 32), (int)(s >> 32), 0x101, 0xF, 0xF, false);
    }
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <typename T>
__device__ __forceinline__ T row_from_left(T v, T edge) { return dpp_row<T>(v, edge, 0x111); }
template <typename T>
__device__ __forceinline__ T row_from_right(T v, T edge) { return dpp_row<T>(v, edge, 0x101); }

// One z plane of a 16^3 brick's 18^3 tile at column (tx, ty): C the node,
// F its 4 in-plane face neighbours, Dg its 4 in-plane corner neighbours, in
// the sums' fixed order.  Six LDS reads per plane instead of nine: the
// three row values (rows y-1, y, y+1) and the three halo-column values that
// only the edge lanes use (read by every lane, which keeps the code
// branch-free); the x neighbours come through DPP.
template <typename T>
__device__ __forceinline__ void plane16(const T* tile, int p, int tx, int ty, T& C, T& F, T& Dg) {
    constexpr int E = 18;
    const T* a0 = tile + (p * E + ty) * E + tx;
    const T* a1 = a0 + E;
    const T* a2 = a1 + E;
    const T c0 = a0[1], c1 = a1[1], c2 = a2[1];
    const int o = tx == 0 ? 0 : 2;
    const T e0 = a0[o], e1 = a1[o], e2 = a2[o];
    const T l0 = row_from_left(c0, e0), r0 = row_from_right(c0, e0);
    const T l1 = row_from_left(c1, e1), r1 = row_from_right(c1, e1);
    const T l2 = row_from_left(c2, e2), r2 = row_from_right(c2, e2);
    C = c1;
    F = (c0 + c2) + (l1 + r1);
    Dg = (l0 + r0) + (l2 + r2);
}

// halo shell index -> (hx,hy,hz) in [-1,B]^3 \\ [0,B)^3
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

// n consecutive elements (an aligned run) as one vector load
template <typename U, int n>
__device__ __forceinline__ void ldrun(const U* p, U* o) {
    if constexpr (n == 1) {
        o[0] = *p;
    } else {
        typedef U vt __attribute__((ext_vector_type(n)));
        const vt v = *(const vt*)p;
#pragma unroll
        for (int e = 0; e < n; ++e) o[e] = v[e];
    }
}
// the operand value of the pass at n nodes starting at g, formed as the
// interior is formed (x ; x64 + e32 ; fma(beta, d_old, r))
template <typename T, int OP, int n>
__device__ __forceinline__ void halo_run(const StencilIO<T>& io, uint64_t g, double bt, double al, int rp, T* o) {
    if constexpr (OP == OP_CG) {
        T r[n], d[n], q[n];
        ldrun<T, n>(io.r + g, r);
        ldrun<T, n>(io.dold + g, d);
        if (rp) ldrun<T, n>(io.q + g, q);
#pragma unroll
        for (int k = 0; k < n; ++k) o[k] = cg_form(bt, al, rp, d[k], r[k], rp ? q[k] : T(0));
    } else if constexpr (OP == OP_RESID_UPD) {
        double a[n];
        float e[n];
        ldrun<double, n>(io.x64 + g, a);
        ldrun<float, n>(io.e32 + g, e);
#pragma unroll
        for (int k = 0; k < n; ++k) o[k] = a[k] + (double)e[k];
    } else if constexpr (OP == OP_DIR_APPLY) {
        T r[n], d[n];
        ldrun<T, n>(io.r + g, r);
        ldrun<T, n>(io.dold + g, d);
#pragma unroll
        for (int k = 0; k < n; ++k) o[k] = dir_form(bt, d[k], r[k]);
    } else {
        ldrun<T, n>(io.x + g, o);
    }
}

// BATCH: the grid is a concatenation of instances (see BatchTab); the
// brick's instance supplies res, the weights, beta and the run flag, and
// halo addresses are offset by the instance's first node.
template <typename T, int B, int OP, bool BATCH = false>
__global__ __launch_bounds__(B* B) void k_stencil_tile(LevelArgs<T> L, StencilIO<T> io, uint32_t nbricks, int swz, BatchTab tab) {
    if constexpr (!BATCH) {
        if (io.done && *io.done) return;
    }
    constexpr int NT = B * B, B3 = B * B * B, E = B + 2, E3 = E * E * E, S = E3 - B3;
    using NV = typename NV16<T>::type;
    typedef float nf2 __attribute__((ext_vector_type(2)));
    typedef double nd2 __attribute__((ext_vector_type(2)));
    constexpr int VEC = 16 / (int)sizeof(T), PER = B3 / NT, NVEC = PER / VEC, NH = (S + NT - 1) / NT;
    __shared__ T tile[E3];
    __shared__ double wsum[4 * ((NT + 63) / 64)];
    __shared__ T gdiag[OP == OP_GS ? B3 / 8 : 1];  // screening diagonal of the colour's touched nodes
    const int tid = threadIdx.x;
    if constexpr (OP == OP_GS)
        for (int i = tid; i < B3 / 8; i += NT) gdiag[i] = T(0);
    uint32_t bid = blockIdx.x;
    if (swz) {
        // workgroups go to the 8 XCDs round-robin: XCD x takes the chunks of
        // swz consecutive bricks numbered x, x+8, x+16, ... in turn
        const uint32_t j = bid >> 3, c = (uint32_t)swz;
        bid = (j / c) * 8u * c + (bid & 7u) * c + j % c;
    }
    const uint64_t base = (uint64_t)bid * B3;
    uint32_t lb = bid, res = L.res;
    uint64_t ib = 0;  // first node of the brick's instance
    double bt = 0.0, al = 0.0;
    int rp = 0;
    if constexpr (BATCH) {
        const uint32_t inst = tab.brick_inst[base >> 12];
        if (tab.run[(int64_t)inst * tab.run_stride] != tab.run_when) return;
        const BCoef c = tab.coef[inst];
        res = c.res;
        if constexpr (sizeof(T) == 8) { L.s = (T)c.s64; L.wg = (T)c.wg64; L.wh3 = (T)c.wh3_64; }
        else { L.s = (T)c.s32; L.wg = (T)c.wg32; L.wh3 = (T)c.wh3_32; }
        if constexpr (OP == OP_RESID_UPD) io.rscale = (double)c.res;  // 1/h of the instance (FP32 system scale)
        ib = tab.off[inst];
        lb = bid - (uint32_t)(ib / B3);
        if constexpr (OP == OP_DIR_APPLY) bt = tab.st[inst].beta;
        if constexpr (OP == OP_CG) {
            bt = tab.st[inst].beta;
            al = tab.st[inst].alpha;
            rp = tab.st[inst].rpend;
        }
    } else {
        if constexpr (OP == OP_DIR_APPLY) bt = *io.beta;
        if constexpr (OP == OP_CG) {
            bt = io.cst->beta;
            al = io.cst->alpha;
            rp = io.cst->rpend;
        }
    }
    uint32_t ox, oy, oz;
    demorton3(lb, ox, oy, oz);
    ox *= B; oy *= B; oz *= B;
    // the brick's touched-node range (screening), needed early by the prefetch below
    uint32_t st0 = 0, st1 = 0;
    if (L.has_screen) {
        st0 = L.brick[bid];
        st1 = L.brick[bid + 1];
    }

    // ---- 1. issue the brick's interior loads (16-byte, coalesced)
    NV qa[NVEC], qb[NVEC], qc[OP == OP_CG ? NVEC : 1];
    nf2 ee[OP == OP_RESID_UPD ? NVEC : 1];
#pragma unroll
    for (int v = 0; v < NVEC; ++v) {
        const uint64_t g = base + (uint32_t)(v * NT + tid) * VEC;
        if constexpr (OP == OP_RESID_UPD) {
            qa[v] = ld_v((const T*)io.x64 + g);
            ee[v] = *(const nf2*)(io.e32 + g);
        } else if constexpr (OP == OP_DIR_APPLY) {
            qa[v] = ld_v(io.r + g);
            qb[v] = ld_v(io.dold + g);
        } else if constexpr (OP == OP_CG) {
            qa[v] = ld_v(io.r + g);
            qb[v] = ld_v(io.dold + g);
            if (rp) qc[v] = ld_v(io.q + g);
            else qc[v] = NV{};
        } else {
            qa[v] = ld_v(io.x + g);
        }
    }
    // ---- 2. issue the one-node halo loads (x = 0 outside the grid).  16^3
    // bricks: face by face from the six face-, twelve edge- and eight
    // corner-neighbour bricks.  In Morton order a z face (local z = 15 or 0)
    // is 64 aligned runs of 4 nodes (2x2 in x,y) and a y face 128 runs of 2
    // (x pairs), so they load as vectors; x faces, edges and corners are
    // single nodes.  A face lies in one neighbour brick, so the grid test is
    // per brick.
    T zf[4], yf[2][2], xf[2], ev;
    int zi = 0, yi[2], xi[2], ei = -1;
    T hv[B == 16 ? 1 : NH];
    int hl[B == 16 ? 1 : NH];
    bool hin[B == 16 ? 1 : NH];
    if constexpr (B == 16) {
        const int bx = (int)(ox >> 4), by = (int)(oy >> 4), bz = (int)(oz >> 4), nbr = (int)(res >> 4);
        auto nb_base = [&](int dx, int dy, int dz, bool& ok) -> uint64_t {
            const int cx = bx + dx, cy = by + dy, cz = bz + dz;
            ok = (unsigned)cx < (unsigned)nbr && (unsigned)cy < (unsigned)nbr && (unsigned)cz < (unsigned)nbr;
            return ok ? ib + (uint64_t)morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz) * B3 : 0;
        };
        auto tix = [](int hx, int hy, int hz) { return (hx + 1) + E * (hy + 1) + E * E * (hz + 1); };
#pragma unroll
        for (int k = 0; k < 4; ++k) zf[k] = T(0);
        if (tid < 128) {  // z faces: one run of 4 per thread
            const int side = tid >> 6, a = tid & 7, b = (tid >> 3) & 7;
            bool ok;
            const uint64_t sb = nb_base(0, 0, side ? 1 : -1, ok);
            const uint32_t m0 = spread3(2 * a) | (spread3(2 * b) << 1) | (side ? 0u : (spread3(15u) << 2));
            if (ok) halo_run<T, OP, 4>(io, sb + m0, bt, al, rp, zf);
            zi = tix(2 * a, 2 * b, side ? B : -1);
        }
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {  // y faces: threads 128..255, a run of 2 on each side
            yf[sd][0] = yf[sd][1] = T(0);
            const int t = tid & 127, a = t & 7, c = t >> 3;
            yi[sd] = tix(2 * a, sd ? B : -1, c);
            if (tid >= 128) {
                bool ok;
                const uint64_t sb = nb_base(0, sd ? 1 : -1, 0, ok);
                const uint32_t m0 = spread3(2 * a) | ((sd ? 0u : spread3(15u)) << 1) | (spread3((uint32_t)c) << 2);
                if (ok) halo_run<T, OP, 2>(io, sb + m0, bt, al, rp, yf[sd]);
            }
        }
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {  // x faces: one node on each side per thread
            const int yy = tid & 15, zz = tid >> 4;
            xf[sd] = T(0);
            bool ok;
            const uint64_t sb = nb_base(sd ? 1 : -1, 0, 0, ok);
            const uint32_t m = (sd ? 0u : spread3(15u)) | (spread3((uint32_t)yy) << 1) | (spread3((uint32_t)zz) << 2);
            if (ok) halo_run<T, OP, 1>(io, sb + m, bt, al, rp, &xf[sd]);
            xi[sd] = tix(sd ? B : -1, yy, zz);
        }
        ev = T(0);
        if (tid < 200) {  // 12 edges of 16 nodes (threads 0..191), 8 corners (192..199)
            int hx, hy, hz;
            if (tid < 192) {
                const int e = tid >> 4, u = tid & 15, axis = e >> 2, s1 = (e & 1) ? B : -1, s2 = (e & 2) ? B : -1;
                if (axis == 0) { hx = u; hy = s1; hz = s2; }
                else if (axis == 1) { hy = u; hx = s1; hz = s2; }
                else { hz = u; hx = s1; hy = s2; }
            } else {
                const int c = tid - 192;
                hx = (c & 1) ? B : -1;
                hy = (c & 2) ? B : -1;
                hz = (c & 4) ? B : -1;
            }
            bool ok;
            const uint64_t sb = nb_base(hx < 0 ? -1 : (hx >= B ? 1 : 0), hy < 0 ? -1 : (hy >= B ? 1 : 0), hz < 0 ? -1 : (hz >= B ? 1 : 0), ok);
            const uint32_t m = spread3((uint32_t)(hx & 15)) | (spread3((uint32_t)(hy & 15)) << 1) | (spread3((uint32_t)(hz & 15)) << 2);
            if (ok) halo_run<T, OP, 1>(io, sb + m, bt, al, rp, &ev);
            ei = tix(hx, hy, hz);
        }
    } else {
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            const int i = tid + k * NT;
            hl[k] = -1;
            hin[k] = false;
            hv[k] = T(0);
            if (i < S) {
                int hx, hy, hz;
                shell_pos<B>(i, hx, hy, hz);
                hl[k] = ((hz + 1) * E + hy + 1) * E + hx + 1;
                const int gx = (int)ox + hx, gy = (int)oy + hy, gz = (int)oz + hz;
                if ((unsigned)gx < res && (unsigned)gy < res && (unsigned)gz < res) {
                    const uint64_t g = ib + morton3((uint32_t)gx, (uint32_t)gy, (uint32_t)gz);
                    hin[k] = true;
                    if constexpr (OP == OP_RESID_UPD) hv[k] = io.x64[g] + (double)io.e32[g];
                    else if constexpr (OP == OP_DIR_APPLY) hv[k] = dir_form(bt, io.dold[g], io.r[g]);
                    else if constexpr (OP == OP_CG) hv[k] = cg_form(bt, al, rp, io.dold[g], io.r[g], rp ? io.q[g] : T(0));
                    else hv[k] = io.x[g];
                }
            }
        }
    }
    // ---- 2b. screening prefetch (bricks with touched nodes): the first
    // group's node slots and the first chunk of its entries are requested
    // now, their point gathers before the sweep, so the chain of dependent
    // loads overlaps the stencil work instead of following it
    constexpr int CH = 4 * NT;
    const bool pre = L.has_screen && L.spc && st0 < st1;
    uint32_t pob[4], poe[4], pnd[4], pspv[4], pe0 = 0, pe1 = 0;
    T ppv[4];
    if (pre) {
        const uint32_t g1 = min(st1, st0 + (uint32_t)CH);
        pe0 = L.off[st0];
        pe1 = L.off[g1];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t t = st0 + k * NT + tid;
            pob[k] = t < g1 ? L.off[t] : 0u;
            poe[k] = t < g1 ? L.off[t + 1] : 0u;
            pnd[k] = t < g1 ? L.node[t] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t j = pe0 + u * NT + tid;
            pspv[u] = j < pe1 ? L.spc[j] : 0u;
        }
    }
    // brick reference value r0 (the brick's first node, formed as the node is)
    T r0;
    if constexpr (OP == OP_RESID_UPD) r0 = io.x64[base] + (double)io.e32[base];
    else if constexpr (OP == OP_DIR_APPLY) r0 = dir_form(bt, io.dold[base], io.r[base]);
    else if constexpr (OP == OP_CG) r0 = cg_form(bt, al, rp, io.dold[base], io.r[base], rp ? io.q[base] : T(0));
    else r0 = io.x[base];
    // ---- 3. interior values into LDS (offset linear in the Morton bits)
    T xin[PER];
    T rk[OP == OP_CG ? PER : 1];  // OP_CG: the updated residual, for r'.y
    const int lbase = (E * E + E + 1) + (int)(compact3(tid * VEC) + E * compact3((tid * VEC) >> 1) + E * E * compact3((tid * VEC) >> 2));
#pragma unroll
    for (int v = 0; v < NVEC; ++v) {
        const uint64_t g = base + (uint32_t)(v * NT + tid) * VEC;
        if constexpr (OP == OP_RESID_UPD) {
            xin[v * VEC] = qa[v][0] + (double)ee[v][0];
            xin[v * VEC + 1] = qa[v][1] + (double)ee[v][1];
        } else if constexpr (OP == OP_DIR_APPLY) {
            // d_new = r + beta*d_old, the same fma wherever a node is formed
            NV nq;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                nq[e] = dir_form(bt, (T)qb[v][e], (T)qa[v][e]);
                xin[v * VEC + e] = nq[e];
            }
            st_nt(io.dnew + g, nq);
        } else if constexpr (OP == OP_CG) {
            // r' = r - alpha q (the previous iteration's update), d' = r' + beta d_old
            NV nr, nq;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                nr[e] = rp ? r_form(al, (T)qc[v][e], (T)qa[v][e]) : (T)qa[v][e];
                nq[e] = dir_form(bt, (T)qb[v][e], nr[e]);
                rk[v * VEC + e] = nr[e];
                xin[v * VEC + e] = nq[e];
            }
            st_nt(io.rout + g, nr);
            st_nt(io.dnew + g, nq);
        } else {
#pragma unroll
            for (int e = 0; e < VEC; ++e) xin[v * VEC + e] = qa[v][e];
        }
#pragma unroll
        for (int e = 0; e < VEC; ++e)
            tile[lbase + lds_lin<E>((uint32_t)(v * NT * VEC)) + lds_lin<E>((uint32_t)e)] = xin[v * VEC + e] - r0;
    }
    if constexpr (B == 16) {
        if (tid < 128) {
            tile[zi] = zf[0] - r0;
            tile[zi + 1] = zf[1] - r0;
            tile[zi + E] = zf[2] - r0;
            tile[zi + E + 1] = zf[3] - r0;
        } else {
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
                tile[yi[sd]] = yf[sd][0] - r0;
                tile[yi[sd] + 1] = yf[sd][1] - r0;
            }
        }
        tile[xi[0]] = xf[0] - r0;
        tile[xi[1]] = xf[1] - r0;
        if (ei >= 0) tile[ei] = ev - r0;
    } else {
#pragma unroll
        for (int k = 0; k < NH; ++k)
            if (hl[k] >= 0) tile[hl[k]] = hin[k] ? hv[k] - r0 : -r0;
    }
    if (pre) {  // the prefetched chunk's point gathers (products phi * s_p)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t q = pspv[u] >> 3, k = pspv[u] & 7u;
            const T tx = L.tt[q], ty = L.tt[L.P + q], tz = L.tt[2 * L.P + q];
            const T bx = (k & 1u) ? tx : T(1) - tx;
            const T by = (k & 2u) ? ty : T(1) - ty;
            const T bz = (k & 4u) ? tz : T(1) - tz;
            const T ph = bx * by * bz;
            ppv[u] = pe0 + u * NT + tid < pe1 ? ph * io.s[q] : T(0);
        }
    }
    __syncthreads();

    // ---- 4. z sweep with a 3-plane register window (values are x - r0)
    int tx, ty;
    column_of<B>(tid, tx, ty);
    T out[B];
    {
        const T sc = L.s, wg = L.wg;
        T Cm, Fm, Dm, Cc, Fc, Dc;
        auto plane = [&](int p, T& C, T& F, T& Dg) {
            if constexpr (B == 16) {
                plane16<T>(tile, p, tx, ty, C, F, Dg);
            } else {
                const T* a0 = tile + (p * E + ty) * E + tx;
                const T* a1 = a0 + E;
                const T* a2 = a1 + E;
                C = a1[1];
                F = (a0[1] + a2[1]) + (a1[0] + a1[2]);
                Dg = (a0[0] + a0[2]) + (a2[0] + a2[2]);
            }
        };
        plane(0, Cm, Fm, Dm);
        plane(1, Cc, Fc, Dc);
#pragma unroll
        for (int z = 0; z < B; ++z) {
            T Cp, Fp, Dp;
            plane(z + 2, Cp, Fp, Dp);
            const T lap = sc * (T(32) * Cc - T(2) * ((Dc + Fm) + Fp) - (Dm + Dp));
            out[z] = lap + wg * (Cc + r0);
            Cm = Cc; Fm = Fc; Dm = Dc;
            Cc = Cp; Fc = Fp; Dc = Dp;
        }
    }
    __syncthreads();
    {
        const uint32_t mxy = spread3((uint32_t)tx) | (spread3((uint32_t)ty) << 1);
#pragma unroll
        for (int z = 0; z < B; ++z) tile[stage_at<B>(mxy | (spread3c((uint32_t)z) << 2))] = out[z];
    }
    __syncthreads();

    // ---- 5. screening: touched nodes of this brick.  The entries (one per
    // point and corner, grouped by node) are read in chunks of 4*NT with
    // independent loads per thread, their products phi*s_p staged in the LDS
    // beyond the output staging, and each node then adds its own entries in
    // ascending order: the sum is deterministic and no thread waits on a
    // chain of dependent global loads.
    if (L.has_screen) {
        const uint32_t t0 = st0, t1 = st1;
        if (t0 < t1) {
            static_assert(E3 - B3 >= CH, "the spare LDS holds one chunk");
            T* prod = tile + B3;
            for (uint32_t g0 = t0; g0 < t1; g0 += CH) {
                const uint32_t g1 = min(t1, g0 + (uint32_t)CH);
                const bool first = pre && g0 == t0;  // uniform
                uint32_t ob[4], oe[4];
                T acc[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t t = g0 + k * NT + tid;
                    ob[k] = first ? pob[k] : (t < g1 ? L.off[t] : 0u);
                    oe[k] = first ? poe[k] : (t < g1 ? L.off[t + 1] : 0u);
                    acc[k] = T(0);
                }
                const uint32_t e0 = first ? pe0 : L.off[g0], e1 = first ? pe1 : L.off[g1];
                for (uint32_t c0 = e0; c0 < e1; c0 += CH) {
                    uint32_t spv[4];
                    T phv[4], pv[4];
                    if (first && c0 == e0) {
#pragma unroll
                        for (int u = 0; u < 4; ++u) pv[u] = ppv[u];
                    } else if (L.spc) {
                        // phi = bx*by*bz from the point's in-cell coordinates, as the gather forms it
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const uint32_t j = c0 + u * NT + tid;
                            spv[u] = j < e1 ? L.spc[j] : 0u;
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const uint32_t q = spv[u] >> 3, k = spv[u] & 7u;
                            const T tx = L.tt[q], ty = L.tt[L.P + q], tz = L.tt[2 * L.P + q];
                            const T bx = (k & 1u) ? tx : T(1) - tx;
                            const T by = (k & 2u) ? ty : T(1) - ty;
                            const T bz = (k & 4u) ? tz : T(1) - tz;
                            phv[u] = bx * by * bz;
                            pv[u] = c0 + u * NT + tid < e1 ? phv[u] * io.s[q] : T(0);
                        }
                    } else {
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const uint32_t j = c0 + u * NT + tid;
                            spv[u] = j < e1 ? L.sp[j] : 0u;
                            phv[u] = j < e1 ? L.phi[j] : T(0);
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) pv[u] = c0 + u * NT + tid < e1 ? phv[u] * io.s[spv[u]] : T(0);
                    }
                    __syncthreads();  // the previous chunk has been read
#pragma unroll
                    for (int u = 0; u < 4; ++u) prod[u * NT + tid] = pv[u];
                    __syncthreads();
                    const uint32_t c1 = min(e1, c0 + (uint32_t)CH);
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        for (uint32_t j = max(ob[k], c0); j < min(oe[k], c1); ++j) acc[k] += prod[j - c0];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t t = g0 + k * NT + tid;
                    if (t < g1) {
                        const uint32_t loc = (uint32_t)((first ? pnd[k] : L.node[t]) - base);
                        tile[stage_at<B>(loc)] += acc[k];
                        if constexpr (OP == OP_GS)
                            if ((int)(loc & 7u) == io.color) gdiag[loc >> 3] = L.wh3 * (T)L.diag[t];
                    }
                }
            }
            __syncthreads();
        }
    }

    // ---- 6. coalesced stores + reduction partial (per-thread terms in T,
    // folded in FP64 across the block and the grid)
    T tacc = T(0);
    double c3[4] = {0.0, 0.0, 0.0, 0.0};  // OP_CG: d'.y, r'.y, y.y, r'.r'
#pragma unroll
    for (int v = 0; v < NVEC; ++v) {
        const uint32_t m = (uint32_t)(v * NT + tid) * VEC;
        const uint64_t g = base + m;
        if constexpr (OP == OP_CG) {
            NV q;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                q[e] = tile[stage_at<B>(m) + e];
                c3[0] += (double)xin[v * VEC + e] * (double)q[e];
                c3[1] += (double)rk[v * VEC + e] * (double)q[e];
                c3[2] += (double)q[e] * (double)q[e];
                c3[3] += (double)rk[v * VEC + e] * (double)rk[v * VEC + e];
            }
            st_nt(io.y + g, q);
        } else if constexpr (OP == OP_APPLY || OP == OP_DIR_APPLY) {
            NV q;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                q[e] = tile[stage_at<B>(m) + e];
                tacc += xin[v * VEC + e] * q[e];
            }
            st_nt(io.y + g, q);
        } else if constexpr (OP == OP_RESID) {
            const NV bb = ld_v(io.b + g);
            NV q;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                q[e] = bb[e] - tile[stage_at<B>(m) + e];
                tacc += q[e] * q[e];
            }
            st_nt(io.y + g, q);
        } else if constexpr (OP == OP_JACOBI) {
            const NV bb = ld_v(io.b + g);
            const T jw = (T)io.jw;
            NV q;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                q[e] = xin[v * VEC + e] + jw * (bb[e] - tile[stage_at<B>(m) + e]);
                tacc += bb[e] * q[e];
            }
            st_nt(io.y + g, q);
        } else if constexpr (OP == OP_GS) {
            // the colour's nodes only (local Morton bits 0..2 are the coordinate parities)
            const T dg0 = T(32) * L.s + L.wg;  // the diagonal of L + wg I
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const uint32_t me = m + (uint32_t)e;
                if ((int)(me & 7u) == io.color) {
                    const T dg = dg0 + gdiag[me >> 3];
                    if (dg != T(0)) io.y[g + e] = xin[v * VEC + e] + (io.b[g + e] - tile[stage_at<B>(m) + e]) / dg;
                }
            }
        } else if constexpr (OP == OP_JACOBI2) {
            const T j2 = (T)(io.jw + io.jw2), jj = (T)(io.jw * io.jw2);
            NV q;
#pragma unroll
            for (int e = 0; e < VEC; ++e) q[e] = j2 * xin[v * VEC + e] - jj * tile[stage_at<B>(m) + e];
            st_nt(io.y + g, q);
        } else {
            const nd2 bb = *(const nd2*)((const double*)io.b + g);
            nd2 xo;
            xo[0] = xin[v * VEC];
            xo[1] = xin[v * VEC + 1];
            __builtin_nontemporal_store(xo, (nd2*)(io.x_out + g));
            const double q0 = bb[0] - tile[stage_at<B>(m)], q1 = bb[1] - tile[stage_at<B>(m) + 1];
            nf2 rr;
            rr[0] = (float)(q0 * io.rscale);
            rr[1] = (float)(q1 * io.rscale);
            __builtin_nontemporal_store(rr, (nf2*)(io.r32 + g));
            tacc += q0 * q0 + q1 * q1;
        }
    }
    if constexpr (OP == OP_JACOBI2 || OP == OP_GS) return;
    if constexpr (OP == OP_CG) {
        double s4[4];
        block_sum_redn<4>(c3, wsum, s4);
        if (tid == 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) io.partial[4 * (uint64_t)bid + j] = s4[j];
        }
        return;
    }
    if constexpr (OP == OP_JACOBI) {
        if (!io.partial) return;  // uniform over the block
    }
    const double sacc = block_sum_red((double)tacc, wsum);
    if (tid == 0 && io.partial) io.partial[bid] = sacc;
}


// ------------------------------------------------------------------ fused CG step, face form
// The fused CG step of OP_CG without a stored q = A d.  Pass k needs
// q = A d_old only to form r' = r - alpha q.  Inside the brick it is
// recomputed from d_old: the pass loads d_old with its halo anyway, and the
// previous pass computed q there from the same LDS tile with the same code,
// so the values are bit for bit those it would have stored.  On the one-node
// halo, q is read from the neighbours' faces, which every pass writes: per
// brick the six 16x16 boundary planes of its A d' (FACE_VALS values; face
// f = 2 * axis + (plane at 15), index a + 16 b over the other two axes in
// x, y, z order).  The screening term needs q at the touched nodes too: the
// pass writes q' there into a node-indexed vector (io.qt, only touched nodes
// written), from which the per-point gather forms the next operand's s_p as
// it does for the stored-q form.  The recomputed q at a touched node is then
// simply read back from io.qt (the brick itself wrote it, from the same tile,
// in the previous pass), so only q' adds the screening sums, which come
// precomputed per touched node (io.sc, launch_cell_screening).  Per iteration this reads r and d_old
// (with halo) and the halo's faces and writes r', d' and the faces: 16 bytes
// per node of vectors plus 3 of faces in FP32, where the stored-q form
// moves 24.  The halo's d' is formed as soon as its loads arrive, so only
// the interior's r and d_old stay in registers across the first sweep.
template <typename T, bool BATCH = false>
__global__ __launch_bounds__(256) void k_cg_faces(LevelArgs<T> L, StencilIO<T> io, uint32_t nbricks, int swz, BatchTab tab) {
    if constexpr (!BATCH) {
        if (io.done && *io.done) return;
    }
    constexpr int B = 16, NT = 256, B3 = 4096, E = B + 2, E3 = E * E * E;
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T), PER = B3 / NT, NVEC = PER / VEC;
    __shared__ T tile[E3];
    __shared__ T stg[B3];
    __shared__ double wsum[4 * (NT / 64)];
    const int tid = threadIdx.x;
    uint32_t bid = blockIdx.x;
    if (swz) {
        const uint32_t j = bid >> 3, c = (uint32_t)swz;
        bid = (j / c) * 8u * c + (bid & 7u) * c + j % c;
    }
    const uint64_t base = (uint64_t)bid * B3;
    uint32_t lb = bid, res = L.res;
    uint64_t ib = 0;
    double bt, al;
    int rp;
    if constexpr (BATCH) {
        const uint32_t inst = tab.brick_inst[base >> 12];
        if (tab.run[(int64_t)inst * tab.run_stride] != tab.run_when) return;
        const BCoef c = tab.coef[inst];
        res = c.res;
        if constexpr (sizeof(T) == 8) { L.s = (T)c.s64; L.wg = (T)c.wg64; }
        else { L.s = (T)c.s32; L.wg = (T)c.wg32; }
        ib = tab.off[inst];
        lb = bid - (uint32_t)(ib / B3);
        bt = tab.st[inst].beta;
        al = tab.st[inst].alpha;
        rp = tab.st[inst].rpend;
    } else {
        bt = io.cst->beta;
        al = io.cst->alpha;
        rp = io.cst->rpend;
    }
    uint32_t ox, oy, oz;
    demorton3(lb, ox, oy, oz);
    ox *= B; oy *= B; oz *= B;

    // the brick's touched nodes (screening)
    uint32_t st0 = 0, st1 = 0;
    if (L.has_screen) {
        st0 = L.brick[bid];
        st1 = L.brick[bid + 1];
    }

    // ---- 1. interior d_old (16-byte, coalesced); the interior r is loaded
    // after the first sweep, which does not need it (fewer registers live
    // across the sweep)
    NV ra[NVEC], da[NVEC];
#pragma unroll
    for (int v = 0; v < NVEC; ++v) da[v] = ld_v(io.dold + base + (uint32_t)(v * NT + tid) * VEC);
    const T r0d = io.dold[base];
    // ---- 2. halo r, d_old from the neighbour bricks (the OP_CG tile's face /
    // edge / corner split) and, when an update is pending, q from the
    // neighbours' faces; d_old - r0d goes to the tile (for the recomputed q),
    // d' is formed at once and kept
    const int bx = (int)(ox >> 4), by = (int)(oy >> 4), bz = (int)(oz >> 4), nbr = (int)(res >> 4);
    auto nb_base = [&](int dx, int dy, int dz, bool& ok) -> uint64_t {
        const int cx = bx + dx, cy = by + dy, cz = bz + dz;
        ok = (unsigned)cx < (unsigned)nbr && (unsigned)cy < (unsigned)nbr && (unsigned)cz < (unsigned)nbr;
        return ok ? ib + (uint64_t)morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz) * B3 : 0;
    };
    auto faces_of = [&](uint64_t nb) { return io.q + (nb >> 12) * FACE_VALS; };
    auto tix = [](int hx, int hy, int hz) { return (hx + 1) + E * (hy + 1) + E * E * (hz + 1); };
    // every halo load is issued before any is used
    T zr[4], zd[4], zq[4], yr[2][2], yd[2][2], yq[2][2], xr[2], xd[2], xq[2], er = T(0), ed = T(0), eq = T(0);
    int zi = 0, yi[2], xi[2], ei = -1;
#pragma unroll
    for (int k = 0; k < 4; ++k) zr[k] = zd[k] = zq[k] = T(0);
    if (tid < 128) {  // z faces: a 2x2 run per thread
        const int side = tid >> 6, a = tid & 7, b = (tid >> 3) & 7;
        bool ok;
        const uint64_t sb0 = nb_base(0, 0, side ? 1 : -1, ok);
        const uint32_t m0 = spread3(2 * a) | (spread3(2 * b) << 1) | (side ? 0u : (spread3(15u) << 2));
        if (ok) {
            ldrun<T, 4>(io.r + sb0 + m0, zr);
            ldrun<T, 4>(io.dold + sb0 + m0, zd);
            if (rp) {
                const T* f = faces_of(sb0) + (side ? 4 : 5) * 256 + 2 * a + 32 * b;
                ldrun<T, 2>(f, zq);
                ldrun<T, 2>(f + 16, zq + 2);
            }
        }
        zi = tix(2 * a, 2 * b, side ? B : -1);
    }
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {  // y faces: threads 128..255, an x pair on each side
        yr[sd][0] = yr[sd][1] = yd[sd][0] = yd[sd][1] = yq[sd][0] = yq[sd][1] = T(0);
        const int t = tid & 127, a = t & 7, c = t >> 3;
        yi[sd] = tix(2 * a, sd ? B : -1, c);
        if (tid >= 128) {
            bool ok;
            const uint64_t sb0 = nb_base(0, sd ? 1 : -1, 0, ok);
            const uint32_t m0 = spread3(2 * a) | ((sd ? 0u : spread3(15u)) << 1) | (spread3((uint32_t)c) << 2);
            if (ok) {
                ldrun<T, 2>(io.r + sb0 + m0, yr[sd]);
                ldrun<T, 2>(io.dold + sb0 + m0, yd[sd]);
                if (rp) ldrun<T, 2>(faces_of(sb0) + (sd ? 2 : 3) * 256 + 2 * a + 16 * c, yq[sd]);
            }
        }
    }
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {  // x faces: one node on each side per thread
        const int yy = tid & 15, zz = tid >> 4;
        xr[sd] = xd[sd] = xq[sd] = T(0);
        bool ok;
        const uint64_t sb0 = nb_base(sd ? 1 : -1, 0, 0, ok);
        const uint32_t m = (sd ? 0u : spread3(15u)) | (spread3((uint32_t)yy) << 1) | (spread3((uint32_t)zz) << 2);
        if (ok) {
            xr[sd] = io.r[sb0 + m];
            xd[sd] = io.dold[sb0 + m];
            if (rp) xq[sd] = faces_of(sb0)[(sd ? 0 : 1) * 256 + yy + 16 * zz];
        }
        xi[sd] = tix(sd ? B : -1, yy, zz);
    }
    if (tid < 200) {  // 12 edges of 16 nodes (threads 0..191), 8 corners (192..199)
        int hx, hy, hz;
        if (tid < 192) {
            const int e = tid >> 4, u = tid & 15, axis = e >> 2, s1 = (e & 1) ? B : -1, s2 = (e & 2) ? B : -1;
            if (axis == 0) { hx = u; hy = s1; hz = s2; }
            else if (axis == 1) { hy = u; hx = s1; hz = s2; }
            else { hz = u; hx = s1; hy = s2; }
        } else {
            const int c = tid - 192;
            hx = (c & 1) ? B : -1;
            hy = (c & 2) ? B : -1;
            hz = (c & 4) ? B : -1;
        }
        bool ok;
        const uint64_t sb0 = nb_base(hx < 0 ? -1 : (hx >= B ? 1 : 0), hy < 0 ? -1 : (hy >= B ? 1 : 0), hz < 0 ? -1 : (hz >= B ? 1 : 0), ok);
        const int lx = hx & 15, ly = hy & 15, lz = hz & 15;
        const uint32_t m = spread3((uint32_t)lx) | (spread3((uint32_t)ly) << 1) | (spread3((uint32_t)lz) << 2);
        if (ok) {
            er = io.r[sb0 + m];
            ed = io.dold[sb0 + m];
            if (rp) {
                // an x-side node lies on the neighbour's x face, an x-axis edge node on its y face
                const bool xs = hx < 0 || hx >= B;
                eq = faces_of(sb0)[xs ? (hx < 0 ? 1 : 0) * 256 + ly + 16 * lz : (hy < 0 ? 3 : 2) * 256 + lx + 16 * lz];
            }
        }
        ei = tix(hx, hy, hz);
    }
    // d_old - r0d of the halo into the tile (for the recomputed q); the halo's d' is formed now and kept
    if (rp) {
        if (tid < 128) {
            tile[zi] = zd[0] - r0d;
            tile[zi + 1] = zd[1] - r0d;
            tile[zi + E] = zd[2] - r0d;
            tile[zi + E + 1] = zd[3] - r0d;
        } else {
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
                tile[yi[sd]] = yd[sd][0] - r0d;
                tile[yi[sd] + 1] = yd[sd][1] - r0d;
            }
        }
        tile[xi[0]] = xd[0] - r0d;
        tile[xi[1]] = xd[1] - r0d;
        if (ei >= 0) tile[ei] = ed - r0d;
    }
    T hz4[4], hy2[2][2], hx1[2], he;
#pragma unroll
    for (int k = 0; k < 4; ++k) hz4[k] = cg_form(bt, al, rp, zd[k], zr[k], zq[k]);
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        hy2[sd][0] = cg_form(bt, al, rp, yd[sd][0], yr[sd][0], yq[sd][0]);
        hy2[sd][1] = cg_form(bt, al, rp, yd[sd][1], yr[sd][1], yq[sd][1]);
        hx1[sd] = cg_form(bt, al, rp, xd[sd], xr[sd], xq[sd]);
    }
    he = cg_form(bt, al, rp, ed, er, eq);

    // ---- helpers: interior placement, z sweep into the staging, screening
    // (the tile holds x - r0, as the OP_CG tile)
    const int lbase = (E * E + E + 1) + (int)(compact3(tid * VEC) + E * compact3((tid * VEC) >> 1) + E * E * compact3((tid * VEC) >> 2));
    int tx, ty;
    column_of<B>(tid, tx, ty);
    const uint32_t mxy = spread3((uint32_t)tx) | (spread3((uint32_t)ty) << 1);
    // (L + wg I) x on the interior, from the tile, into stg
    auto sweep = [&](T r0) {
        const T sc = L.s, wg = L.wg;
        T Cm, Fm, Dm, Cc, Fc, Dc;
        plane16<T>(tile, 0, tx, ty, Cm, Fm, Dm);
        plane16<T>(tile, 1, tx, ty, Cc, Fc, Dc);
#pragma unroll
        for (int z = 0; z < B; ++z) {
            T Cp, Fp, Dp;
            plane16<T>(tile, z + 2, tx, ty, Cp, Fp, Dp);
            const T lap = sc * (T(32) * Cc - T(2) * ((Dc + Fm) + Fp) - (Dm + Dp));
            stg[stage_at<B>(mxy | (spread3c((uint32_t)z) << 2))] = lap + wg * (Cc + r0);
            Cm = Cc; Fm = Fc; Dm = Dc;
            Cc = Cp; Fc = Fp; Dc = Dp;
        }
    };
    // q = A d_old at the touched nodes: the values the previous pass wrote
    // (4 nodes per thread with all loads issued first)
    auto touched_from_qt = [&]() {
        for (uint32_t t0 = st0; t0 < st1; t0 += 4 * NT) {
            uint32_t loc[4];
            T v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t t = t0 + k * NT + tid;
                loc[k] = t < st1 ? (uint32_t)(L.node[t] - base) : 0u;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = t0 + k * NT + tid < st1 ? io.qt[base + loc[k]] : T(0);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (t0 + k * NT + tid < st1) stg[stage_at<B>(loc[k])] = v[k];
        }
    };
    // S d' at the touched nodes added to stg, and the results to io.qt
    auto screening = [&]() {
        for (uint32_t t0 = st0; t0 < st1; t0 += 4 * NT) {
            uint32_t loc[4];
            T sv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t t = t0 + k * NT + tid;
                const bool in = t < st1;
                loc[k] = in ? (uint32_t)(L.node[t] - base) : 0u;
                sv[k] = in ? io.sc[t] : T(0);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (t0 + k * NT + tid < st1) {
                    const uint32_t at = stage_at<B>(loc[k]);
                    const T v = stg[at] + sv[k];
                    stg[at] = v;
                    io.qt[base + loc[k]] = v;
                }
        }
    };

    // ---- 3. q = A d_old on the interior (when an update is pending), then
    // r' = r - alpha q and d' = r' + beta d_old on the interior
    T rn[PER], dn[PER];
    T r0n;
    auto load_r = [&]() {
#pragma unroll
        for (int v = 0; v < NVEC; ++v) ra[v] = ld_v(io.r + base + (uint32_t)(v * NT + tid) * VEC);
    };
    if (rp) {
#pragma unroll
        for (int v = 0; v < NVEC; ++v)
#pragma unroll
            for (int e = 0; e < VEC; ++e)
                tile[lbase + lds_lin<E>((uint32_t)(v * NT * VEC)) + lds_lin<E>((uint32_t)e)] = (T)da[v][e] - r0d;
        __syncthreads();
        sweep(r0d);
        load_r();
        __syncthreads();
        if (L.has_screen) touched_from_qt();
        __syncthreads();
        r0n = dir_form(bt, r0d, r_form(al, stg[stage_at<B>(0u)], io.r[base]));
#pragma unroll
        for (int v = 0; v < NVEC; ++v) {
            const uint32_t m = (uint32_t)(v * NT + tid) * VEC;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                rn[v * VEC + e] = r_form(al, stg[stage_at<B>(m) + e], (T)ra[v][e]);
                dn[v * VEC + e] = dir_form(bt, (T)da[v][e], rn[v * VEC + e]);
            }
        }
    } else {
        load_r();
        r0n = dir_form(bt, r0d, io.r[base]);
#pragma unroll
        for (int v = 0; v < NVEC; ++v)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                rn[v * VEC + e] = (T)ra[v][e];
                dn[v * VEC + e] = dir_form(bt, (T)da[v][e], rn[v * VEC + e]);
            }
    }
    double c3[4] = {0.0, 0.0, 0.0, 0.0};  // d'.q', r'.q', q'.q', r'.r'
#pragma unroll
    for (int v = 0; v < NVEC; ++v) {
        const uint64_t g = base + (uint32_t)(v * NT + tid) * VEC;
        NV nr, nd;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            nr[e] = rn[v * VEC + e];
            nd[e] = dn[v * VEC + e];
            c3[3] += (double)nr[e] * (double)nr[e];
        }
        st_nt(io.rout + g, nr);
        st_nt(io.dnew + g, nd);
    }
    // the tile now takes d' (every read of d_old's tile ended at the barrier above)
#pragma unroll
    for (int v = 0; v < NVEC; ++v)
#pragma unroll
        for (int e = 0; e < VEC; ++e)
            tile[lbase + lds_lin<E>((uint32_t)(v * NT * VEC)) + lds_lin<E>((uint32_t)e)] = dn[v * VEC + e] - r0n;
    if (tid < 128) {
        tile[zi] = hz4[0] - r0n;
        tile[zi + 1] = hz4[1] - r0n;
        tile[zi + E] = hz4[2] - r0n;
        tile[zi + E + 1] = hz4[3] - r0n;
    } else {
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {
            tile[yi[sd]] = hy2[sd][0] - r0n;
            tile[yi[sd] + 1] = hy2[sd][1] - r0n;
        }
    }
    tile[xi[0]] = hx1[0] - r0n;
    tile[xi[1]] = hx1[1] - r0n;
    if (ei >= 0) tile[ei] = he - r0n;
    __syncthreads();

    // ---- 4. q' = A d' on the interior
    sweep(r0n);
    __syncthreads();
    if (L.has_screen) screening();
    __syncthreads();

    // ---- 5. the pass's dot partials and the faces of q'
#pragma unroll
    for (int v = 0; v < NVEC; ++v) {
        const uint32_t m = (uint32_t)(v * NT + tid) * VEC;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            const T q = stg[stage_at<B>(m) + e];
            c3[0] += (double)dn[v * VEC + e] * (double)q;
            c3[1] += (double)rn[v * VEC + e] * (double)q;
            c3[2] += (double)q * (double)q;
        }
    }
    {
        T* fo = io.y + (uint64_t)bid * FACE_VALS;
        const uint32_t a = (uint32_t)tid & 15u, b = (uint32_t)tid >> 4;
        fo[tid] = stg[stage_at<B>(morton3(0u, a, b))];
        fo[256 + tid] = stg[stage_at<B>(morton3(15u, a, b))];
        fo[512 + tid] = stg[stage_at<B>(morton3(a, 0u, b))];
        fo[768 + tid] = stg[stage_at<B>(morton3(a, 15u, b))];
        fo[1024 + tid] = stg[stage_at<B>(morton3(a, b, 0u))];
        fo[1280 + tid] = stg[stage_at<B>(morton3(a, b, 15u))];
    }
    double s4[4];
    block_sum_redn<4>(c3, wsum, s4);
    if (tid == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) io.partial[4 * (uint64_t)bid + j] = s4[j];
    }
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
// ib: first node of the instance (batched grids; 0 otherwise); i is the
// global node id, (ix,iy,iz) the coordinates inside the instance.
template <typename T, int MODE>  // MODE 0: x ; 1: x64 + e32 ; 2: r + beta*d (x = r, x64 unused, d = dd)
__device__ inline T apply_node(const LevelArgs<T>& L, const T* x, const double* x64, const float* e32, const T* s, uint32_t ix,
                               uint32_t iy, uint32_t iz, uint32_t i, T* diag, const T* dd = nullptr, double bt = 0.0,
                               uint64_t ib = 0) {
    auto X = [&](int jx, int jy, int jz) -> T {
        if ((unsigned)jx >= L.res || (unsigned)jy >= L.res || (unsigned)jz >= L.res) return T(0);
        const uint64_t j = ib + morton3((uint32_t)jx, (uint32_t)jy, (uint32_t)jz);
        if constexpr (MODE == 1) return (T)(x64[j] + (double)e32[j]);
        else if constexpr (MODE == 2) return dir_form(bt, dd[j], x[j]);
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
        } else if constexpr (OP == OP_DIR_APPLY) {
            const double bt = *io.beta;
            const T a = apply_node<T, 2>(L, io.r, nullptr, nullptr, io.s, ix, iy, iz, i, nullptr, io.dold, bt);
            const T dn = dir_form(bt, io.dold[i], io.r[i]);
            io.dnew[i] = dn;
            io.y[i] = a;
            acc = (double)dn * (double)a;
        } else {
            const T a = apply_node<T, 0>(L, io.x, nullptr, nullptr, io.s, ix, iy, iz, i, nullptr);
            if constexpr (OP == OP_APPLY) {
                io.y[i] = a;
                acc = (double)io.x[i] * (double)a;
            } else if constexpr (OP == OP_JACOBI) {
                const T y = io.x[i] + (T)io.jw * (io.b[i] - a);
                io.y[i] = y;
                acc = (double)io.b[i] * (double)y;
            } else if constexpr (OP == OP_JACOBI2) {
                io.y[i] = (T)(io.jw + io.jw2) * io.x[i] - (T)(io.jw * io.jw2) * a;
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

// bricks per XCD chunk of the tile launch (0: no remapping).  Chunks of
// PRX_XCD_CHUNK consecutive Morton bricks keep a brick's near neighbours in
// its XCD's L2, while the eight XCDs sweep the grid together so that a far
// neighbour (the next brick row or plane in Morton order) was read shortly
// before by some XCD and is still in the last-level cache.  Measured on the
// depth-9 grid (32768 16^3 bricks), FP32 direction pass: chunk 1 0.557 ms,
// 8 0.462, 32 0.440, 64 0.435, 128 0.440, 4096 (one contiguous octant per
// XCD) 0.465; FP32 fused CG step: chunk 16 0.656 ms, 32 0.645, 64 0.637,
// 128 0.643.
#ifndef PRX_XCD_CHUNK
#define PRX_XCD_CHUNK 64
#endif
static int xcd_chunk(uint32_t nbricks) {
    if (nbricks < 64 || (nbricks & 7u)) return 0;
    uint32_t c = nbricks >> 3;
    while (c > (uint32_t)PRX_XCD_CHUNK && (c & 1u) == 0) c >>= 1;
    return (int)c;
}

template <typename T>
void launch_stencil(const LevelView& Lv, int op, const StencilIO<T>& io, int* nblk, int tile, hipStream_t s) {
    const int64_t N = Lv.N;
    int B = 0;
    if (Lv.D >= 4) B = (tile == 8) ? 8 : 16;
    else if (Lv.D == 3) B = 8;
    if (B == 0) {
        if (op == OP_CG) throw Error(PRX_ERR_INVALID, "the fused CG pass needs a tiled level (depth >= 3)");
        LevelArgs<T> a = make_args<T>(Lv, 12);
        const int nb = (int)((N + 255) / 256);
        switch (op) {
            case OP_APPLY: hipLaunchKernelGGL((k_stencil_node<T, OP_APPLY>), dim3(nb), dim3(256), 0, s, a, io, (uint32_t)N); break;
            case OP_DIR_APPLY:
                hipLaunchKernelGGL((k_stencil_node<T, OP_DIR_APPLY>), dim3(nb), dim3(256), 0, s, a, io, (uint32_t)N);
                break;
            case OP_RESID: hipLaunchKernelGGL((k_stencil_node<T, OP_RESID>), dim3(nb), dim3(256), 0, s, a, io, (uint32_t)N); break;
            case OP_JACOBI: hipLaunchKernelGGL((k_stencil_node<T, OP_JACOBI>), dim3(nb), dim3(256), 0, s, a, io, (uint32_t)N); break;
            case OP_JACOBI2: hipLaunchKernelGGL((k_stencil_node<T, OP_JACOBI2>), dim3(nb), dim3(256), 0, s, a, io, (uint32_t)N); break;
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
    const int swz = xcd_chunk(nbricks);
#define PRX_TILE_LAUNCH(BB, OPV) \
    hipLaunchKernelGGL((k_stencil_tile<T, BB, OPV>), dim3(nbricks), dim3(BB * BB), 0, s, a, io, nbricks, swz, BatchTab{})
    if (B == 16) {
        if (op == OP_APPLY) PRX_TILE_LAUNCH(16, OP_APPLY);
        else if (op == OP_DIR_APPLY) PRX_TILE_LAUNCH(16, OP_DIR_APPLY);
        else if (op == OP_RESID) PRX_TILE_LAUNCH(16, OP_RESID);
        else if (op == OP_JACOBI) PRX_TILE_LAUNCH(16, OP_JACOBI);
        else if (op == OP_JACOBI2) PRX_TILE_LAUNCH(16, OP_JACOBI2);
        else if (op == OP_GS) PRX_TILE_LAUNCH(16, OP_GS);
        else if (op == OP_CG) PRX_TILE_LAUNCH(16, OP_CG);
        else if constexpr (sizeof(T) == 8) PRX_TILE_LAUNCH(16, OP_RESID_UPD);
    } else {
        if (op == OP_APPLY) PRX_TILE_LAUNCH(8, OP_APPLY);
        else if (op == OP_DIR_APPLY) PRX_TILE_LAUNCH(8, OP_DIR_APPLY);
        else if (op == OP_RESID) PRX_TILE_LAUNCH(8, OP_RESID);
        else if (op == OP_JACOBI) PRX_TILE_LAUNCH(8, OP_JACOBI);
        else if (op == OP_JACOBI2) PRX_TILE_LAUNCH(8, OP_JACOBI2);
        else if (op == OP_GS) PRX_TILE_LAUNCH(8, OP_GS);
        else if (op == OP_CG) PRX_TILE_LAUNCH(8, OP_CG);
        else if constexpr (sizeof(T) == 8) PRX_TILE_LAUNCH(8, OP_RESID_UPD);
    }
#undef PRX_TILE_LAUNCH
    PRX_LAUNCH_CHECK();
    *nblk = (int)nbricks;
}
template void launch_stencil<double>(const LevelView&, int, const StencilIO<double>&, int*, int, hipStream_t);
template void launch_stencil<float>(const LevelView&, int, const StencilIO<float>&, int*, int, hipStream_t);


template <typename T>
void launch_cg_faces(const LevelView& Lv, const StencilIO<T>& io, int* nblk, hipStream_t s) {
    if (Lv.D < 4) throw Error(PRX_ERR_INVALID, "the face form of the fused CG pass needs 16^3 bricks (depth >= 4)");
    const LevelArgs<T> a = make_args<T>(Lv, 12);
    if (a.has_screen && !(io.sc && io.qt)) throw Error(PRX_ERR_STATE, "the face form needs sc and qt");
    const uint32_t nbricks = (uint32_t)(Lv.N >> 12);
    hipLaunchKernelGGL((k_cg_faces<T, false>), dim3(nbricks), dim3(256), 0, s, a, io, nbricks, xcd_chunk(nbricks), BatchTab{});
    PRX_LAUNCH_CHECK();
    *nblk = (int)nbricks;
}
template void launch_cg_faces<double>(const LevelView&, const StencilIO<double>&, int*, hipStream_t);
template void launch_cg_faces<float>(const LevelView&, const StencilIO<float>&, int*, hipStream_t);

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

// ------------------------------------------------------------------ batched grids
template <typename T>
static LevelArgs<T> batch_args(const NodeList* scr, int tile_log) {
    LevelArgs<T> a{};
    a.has_screen = (scr && scr->T > 0) ? 1 : 0;
    if (a.has_screen) {
        a.brick_log = scr->brick_log;
        a.brick = scr->brick;
        if (tile_log == 9 && scr->brick8) { a.brick = scr->brick8; a.brick_log = 9; }
        a.node = scr->node;
        a.off = scr->off;
        a.sp = scr->sp;
        if constexpr (sizeof(T) == 8) a.phi = (const T*)scr->phi;
        else a.phi = (const T*)scr->phi32;
        a.diag = scr->diag;
    }
    return a;
}

template <typename T>
void launch_bstencil(const BatchTab& tab, int64_t Ntot, const NodeList* scr, int op, const StencilIO<T>& io, int tile,
                     hipStream_t s) {
    const int B = tile == 8 ? 8 : 16, tl = B == 16 ? 12 : 9;
    const LevelArgs<T> a = batch_args<T>(scr, tl);
    const uint32_t nbricks = (uint32_t)(Ntot >> tl);
    const int swz = xcd_chunk(nbricks);
#define PRX_BTILE_LAUNCH(BB, OPV) \
    hipLaunchKernelGGL((k_stencil_tile<T, BB, OPV, true>), dim3(nbricks), dim3(BB * BB), 0, s, a, io, nbricks, swz, tab)
    if (B == 16) {
        if (op == OP_APPLY) PRX_BTILE_LAUNCH(16, OP_APPLY);
        else if (op == OP_DIR_APPLY) PRX_BTILE_LAUNCH(16, OP_DIR_APPLY);
        else if (op == OP_RESID) PRX_BTILE_LAUNCH(16, OP_RESID);
        else if (op == OP_GS) PRX_BTILE_LAUNCH(16, OP_GS);
        else if (op == OP_CG) PRX_BTILE_LAUNCH(16, OP_CG);
        else if (op == OP_RESID_UPD) {
            if constexpr (sizeof(T) == 8) PRX_BTILE_LAUNCH(16, OP_RESID_UPD);
            else throw Error(PRX_ERR_INVALID, "the residual update pass is FP64");
        } else throw Error(PRX_ERR_INVALID, "batched operator pass not available");
    } else {
        if (op == OP_APPLY) PRX_BTILE_LAUNCH(8, OP_APPLY);
        else if (op == OP_DIR_APPLY) PRX_BTILE_LAUNCH(8, OP_DIR_APPLY);
        else if (op == OP_RESID) PRX_BTILE_LAUNCH(8, OP_RESID);
        else if (op == OP_GS) PRX_BTILE_LAUNCH(8, OP_GS);
        else if (op == OP_CG) PRX_BTILE_LAUNCH(8, OP_CG);
        else if (op == OP_RESID_UPD) {
            if constexpr (sizeof(T) == 8) PRX_BTILE_LAUNCH(8, OP_RESID_UPD);
            else throw Error(PRX_ERR_INVALID, "the residual update pass is FP64");
        } else throw Error(PRX_ERR_INVALID, "batched operator pass not available");
    }
#undef PRX_BTILE_LAUNCH
    PRX_LAUNCH_CHECK();
}
template void launch_bstencil<double>(const BatchTab&, int64_t, const NodeList*, int, const StencilIO<double>&, int, hipStream_t);

template <typename T>
void launch_bcg_faces(const BatchTab& tab, int64_t Ntot, const NodeList* scr, const StencilIO<T>& io, hipStream_t s) {
    const LevelArgs<T> a = batch_args<T>(scr, 12);
    if (a.has_screen && !(io.sc && io.qt)) throw Error(PRX_ERR_STATE, "the face form needs sc and qt");
    const uint32_t nbricks = (uint32_t)(Ntot >> 12);
    hipLaunchKernelGGL((k_cg_faces<T, true>), dim3(nbricks), dim3(256), 0, s, a, io, nbricks, xcd_chunk(nbricks), tab);
    PRX_LAUNCH_CHECK();
}
template void launch_bcg_faces<double>(const BatchTab&, int64_t, const NodeList*, const StencilIO<double>&, hipStream_t);
template void launch_bcg_faces<float>(const BatchTab&, int64_t, const NodeList*, const StencilIO<float>&, hipStream_t);
template void launch_bstencil<float>(const BatchTab&, int64_t, const NodeList*, int, const StencilIO<float>&, int, hipStream_t);


}  // namespace prx
