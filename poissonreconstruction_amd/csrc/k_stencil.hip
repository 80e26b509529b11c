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
#include "tile.hpp"

namespace prx {

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
