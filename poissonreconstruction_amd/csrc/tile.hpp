// The brick-tiled operator pass (shared by the one-launch-per-pass kernels
// and the persistent batched solve): level arguments, LDS tile helpers and
// the pass over one 16^3 or 8^3 brick, stencil_tile_brick.  The algorithm
// and its accuracy are described at the top of k_stencil.hip.
#pragma once

#include "kernels.hpp"

namespace prx {


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
//
// The pass over one brick is a device function of the workgroup (block id
// vbid before the XCD remapping; LDS passed in), so that a persistent kernel
// can run it over many bricks; k_stencil_tile is the one-brick-per-workgroup
// launch of it.
template <typename T, int B, int OP>
struct TileLds {
    static constexpr int E = B + 2, NT = B * B, B3 = B * B * B;
    T tile[E * E * E];
    double wsum[4 * ((NT + 63) / 64)];
    T gdiag[OP == OP_GS ? B3 / 8 : 1];  // screening diagonal of the colour's touched nodes
};
// LOOPED: the caller runs the pass over many bricks in one workgroup (a
// persistent kernel).  The thread index is then passed through an empty asm,
// so the compiler cannot hoist the brick-independent index arithmetic out of
// the caller's loops and keep it live across them (measured: 173 instead of
// 127 VGPRs for a loop over this pass).
template <typename T, int B, int OP, bool BATCH = false, bool LOOPED = false>
__device__ __forceinline__ void stencil_tile_brick(const LevelArgs<T>& L, const StencilIO<T>& io, int swz, const BatchTab& tab,
                                                   uint32_t vbid, TileLds<T, B, OP>& lds) {
    if constexpr (!BATCH) {
        if (io.done && *io.done) return;
    }
    constexpr int NT = B * B, B3 = B * B * B, E = B + 2, E3 = E * E * E, S = E3 - B3;
    using NV = typename NV16<T>::type;
    typedef float nf2 __attribute__((ext_vector_type(2)));
    typedef double nd2 __attribute__((ext_vector_type(2)));
    constexpr int VEC = 16 / (int)sizeof(T), PER = B3 / NT, NVEC = PER / VEC, NH = (S + NT - 1) / NT;
    T* tile = lds.tile;
    double* wsum = lds.wsum;
    T* gdiag = lds.gdiag;
    int tid = threadIdx.x;
    if constexpr (LOOPED) asm volatile("" : "+v"(tid));
    if constexpr (OP == OP_GS)
        for (int i = tid; i < B3 / 8; i += NT) gdiag[i] = T(0);
    // the weights and the residual scale (a batch instance's own replace the level's)
    T Ls = L.s, Lwg = L.wg, Lwh3 = L.wh3;
    double rscale = io.rscale;
    uint32_t bid = vbid;
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
        if constexpr (sizeof(T) == 8) { Ls = (T)c.s64; Lwg = (T)c.wg64; Lwh3 = (T)c.wh3_64; }
        else { Ls = (T)c.s32; Lwg = (T)c.wg32; Lwh3 = (T)c.wh3_32; }
        if constexpr (OP == OP_RESID_UPD) rscale = (double)c.res;  // 1/h of the instance (FP32 system scale)
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
        const T sc = Ls, wg = Lwg;
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
                            if ((int)(loc & 7u) == io.color) gdiag[loc >> 3] = Lwh3 * (T)L.diag[t];
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
            const T dg0 = T(32) * Ls + Lwg;  // the diagonal of L + wg I
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
            rr[0] = (float)(q0 * rscale);
            rr[1] = (float)(q1 * rscale);
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
template <typename T, int B, int OP, bool BATCH = false>
__global__ __launch_bounds__(B* B) void k_stencil_tile(LevelArgs<T> L, StencilIO<T> io, uint32_t nbricks, int swz, BatchTab tab) {
    __shared__ TileLds<T, B, OP> lds;
    stencil_tile_brick<T, B, OP, BATCH>(L, io, swz, tab, blockIdx.x, lds);
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
inline int xcd_chunk(uint32_t nbricks) {
    if (nbricks < 64 || (nbricks & 7u)) return 0;
    uint32_t c = nbricks >> 3;
    while (c > (uint32_t)PRX_XCD_CHUNK && (c & 1u) == 0) c >>= 1;
    return (int)c;
}

// level arguments of a batched level (weights come per instance from the BatchTab)
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

}  // namespace prx
