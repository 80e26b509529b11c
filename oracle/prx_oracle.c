/*
 * prx_oracle.c -- FP64 CPU restatement of the reference's screened-Poisson
 * solve path.  TEST INFRASTRUCTURE ONLY: the parity tests, smoke() and the
 * benchmark's cpu_baseline leg load it as the checker / timed CPU baseline.
 * The product library never links or calls it.
 *
 * Every function follows the reference file:line named in its comment, with
 * the same loop order and the same floating-point evaluation order (compiled
 * with -ffp-contract=off, so a*b+c is never fused, as in Rust).  Parallel
 * regions sit exactly where the reference uses rayon (point transform, CSR
 * SpMV, dot/square_norm, iso mean); everything else is serial, as there.
 *
 * Parity status of each function is recorded in the header of oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* The case table is pinned to the reference's TRI_TABLE / EDGE_TABLE by a
 * digest fixture (tests/golden/mc_tables_digest.json). */
#include "../poissonreconstruction_amd/csrc/mc_cases.h"

typedef struct { double px, py, pz, nx, ny, nz; } orc_point;

/* Rust `f as isize`: saturating, NaN -> 0. */
static inline int64_t sat_i64(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v; /* truncation toward zero */
}
/* wrapping isize add (release-mode Rust) */
static inline int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }

/* ---------------------------------------------------------------- geometry
 * BBox::extend (src/geometry.rs:356-363) uses f64::min/max, which ignore NaN.
 * unit_cube_transform (src/geometry.rs:384-402), Mat4::transform_point/vector
 * (src/geometry.rs:256-269), Vec3::normalize (src/geometry.rs:49-56).
 * Applied as in src/reconstructor.rs:31-40.  m is column-major m[col*4+row]. */
static inline double fmin_rs(double a, double b) { if (a != a) return b; if (b != b) return a; return a < b ? a : b; }
static inline double fmax_rs(double a, double b) { if (a != a) return b; if (b != b) return a; return a > b ? a : b; }

void orc_unit_transform(const orc_point* pts, int64_t n, double scale, double* m, double* minv) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i) {
        const double p[3] = {pts[i].px, pts[i].py, pts[i].pz};
        for (int k = 0; k < 3; ++k) { mn[k] = fmin_rs(mn[k], p[k]); mx[k] = fmax_rs(mx[k], p[k]); }
    }
    double c[3], sz[3];
    for (int k = 0; k < 3; ++k) { c[k] = (mn[k] + mx[k]) * 0.5; sz[k] = mx[k] - mn[k]; }
    double md = fmax_rs(fmax_rs(sz[0], sz[1]), sz[2]);
    double s = (md > 0.0) ? scale / md : 1.0;
    memset(m, 0, 16 * sizeof(double));
    m[0] = s; m[5] = s; m[10] = s;
    m[12] = -s * c[0] + 0.5; m[13] = -s * c[1] + 0.5; m[14] = -s * c[2] + 0.5; m[15] = 1.0;
    if (minv) {
        /* Mat4::inverse (src/geometry.rs:274-316) */
        double a = m[0], b = m[4], cc = m[8], d = m[1], e = m[5], f = m[9], g = m[2], h = m[6], ii = m[10];
        double det = a * (e * ii - f * h) - b * (d * ii - f * g) + cc * (d * h - e * g);
        memset(minv, 0, 16 * sizeof(double));
        if (fabs(det) < 1e-30) { minv[0] = minv[5] = minv[10] = minv[15] = 1.0; return; }
        double id = 1.0 / det;
        double ai = (e * ii - f * h) * id, bi = (cc * h - b * ii) * id, ci = (b * f - cc * e) * id;
        double di = (f * g - d * ii) * id, ei = (a * ii - cc * g) * id, fi = (cc * d - a * f) * id;
        double gi = (d * h - e * g) * id, hi = (b * g - a * h) * id, iii = (a * e - b * d) * id;
        double tx = m[12], ty = m[13], tz = m[14];
        minv[0] = ai; minv[1] = di; minv[2] = gi;
        minv[4] = bi; minv[5] = ei; minv[6] = hi;
        minv[8] = ci; minv[9] = fi; minv[10] = iii;
        minv[12] = -(ai * tx + bi * ty + ci * tz);
        minv[13] = -(di * tx + ei * ty + fi * tz);
        minv[14] = -(gi * tx + hi * ty + iii * tz);
        minv[15] = 1.0;
    }
}

void orc_apply_transform(const orc_point* pts, int64_t n, const double* m, orc_point* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const orc_point p = pts[i];
        double x = m[0] * p.px + m[4] * p.py + m[8] * p.pz + m[12];
        double y = m[1] * p.px + m[5] * p.py + m[9] * p.pz + m[13];
        double z = m[2] * p.px + m[6] * p.py + m[10] * p.pz + m[14];
        double vx = m[0] * p.nx + m[4] * p.ny + m[8] * p.nz;
        double vy = m[1] * p.nx + m[5] * p.ny + m[9] * p.nz;
        double vz = m[2] * p.nx + m[6] * p.ny + m[10] * p.nz;
        double nn = sqrt(vx * vx + vy * vy + vz * vz);
        if (nn > 0.0) { vx = vx / nn; vy = vy / nn; vz = vz / nn; } else { vx = vy = vz = 0.0; }
        out[i].px = x; out[i].py = y; out[i].pz = z;
        out[i].nx = vx; out[i].ny = vy; out[i].nz = vz;
    }
}

/* ---------------------------------------------------------------- tree
 * FEMTree::initialize (src/fem_tree.rs:83-93) refines every node to depth D;
 * child i of a node at offset o has offset 2o + ((i>>d)&1) per axis d
 * (src/octree.rs:145-153).  SortedTreeNodes::collect_nodes
 * (src/octree.rs:235-248) visits nodes in DFS pre-order, bucketed by depth;
 * FEMTree::finalize (src/fem_tree.rs:96-126) numbers the depth-D leaves in
 * that order and fills grid_to_idx[z*res*res + y*res + x] and idx_to_offset.
 * This restatement walks the same DFS; it does not use a Morton formula. */
typedef struct { uint32_t D; int64_t res; int64_t next; int64_t* g2i; uint32_t* i2o; } tree_ctx;

static void tree_visit(tree_ctx* t, uint32_t d, uint32_t ox, uint32_t oy, uint32_t oz) {
    if (d == t->D) {
        int64_t idx = t->next++;
        if (t->g2i) t->g2i[(int64_t)oz * t->res * t->res + (int64_t)oy * t->res + ox] = idx;
        if (t->i2o) { t->i2o[3 * idx] = ox; t->i2o[3 * idx + 1] = oy; t->i2o[3 * idx + 2] = oz; }
        return;
    }
    for (uint32_t i = 0; i < 8; ++i)
        tree_visit(t, d + 1, ox * 2 + (i & 1), oy * 2 + ((i >> 1) & 1), oz * 2 + ((i >> 2) & 1));
}

int64_t orc_tree(uint32_t D, int64_t* g2i, uint32_t* i2o) {
    tree_ctx t = {D, (int64_t)1 << D, 0, g2i, i2o};
    int64_t n = t.res * t.res * t.res;
    if (g2i) for (int64_t i = 0; i < n; ++i) g2i[i] = -1;
    tree_visit(&t, 0, 0, 0, 0);
    return t.next;
}

/* FEMTree::idx_at (src/fem_tree.rs:76-81) */
static inline int64_t idx_at(const int64_t* g2i, int64_t res, uint64_t x, uint64_t y, uint64_t z) {
    if (x >= (uint64_t)res || y >= (uint64_t)res || z >= (uint64_t)res) return -1;
    return g2i[(int64_t)z * res * res + (int64_t)y * res + (int64_t)x];
}

/* ---------------------------------------------------------------- stencil constants
 * src/fem_tree.rs:39-48 */
static inline double mass_1d(int64_t d) { int64_t a = d < 0 ? -d : d; return a == 0 ? 2. / 3. : (a == 1 ? 1. / 6. : 0.0); }
static inline double stiff_1d(int64_t d) { int64_t a = d < 0 ? -d : d; return a == 0 ? 2. : (a == 1 ? -1. : 0.0); }
static inline double deriv_mass_1d(int64_t d) { return d == 0 ? 0. : (d == 1 ? -0.5 : (d == -1 ? 0.5 : 0.0)); }
static inline double stiffness_entry(double h, int64_t dx, int64_t dy, int64_t dz) {
    return (stiff_1d(dx) * mass_1d(dy) * mass_1d(dz) + mass_1d(dx) * stiff_1d(dy) * mass_1d(dz) +
            mass_1d(dx) * mass_1d(dy) * stiff_1d(dz)) * h;
}
double orc_stiffness_entry(double h, int64_t dx, int64_t dy, int64_t dz) { return stiffness_entry(h, dx, dy, dz); }

/* ---------------------------------------------------------------- splat
 * FEMTree::splat_normal_field (src/fem_tree.rs:130-178): floor-based cell,
 * trilinear weights to c+{0,1}^3 (loop dz, dy, dx), out-of-range skipped,
 * serial accumulation in point order, then nf = -nf * (1/w) where w > 0.
 * nf is AoS [3N]. */
void orc_splat(uint32_t D, const int64_t* g2i, const orc_point* pts, int64_t P, double* nf, double* w) {
    const int64_t res = (int64_t)1 << D, n = res * res * res;
    const double h = 1.0 / (double)res;
    memset(nf, 0, 3 * n * sizeof(double));
    memset(w, 0, n * sizeof(double));
    for (int64_t p = 0; p < P; ++p) {
        const orc_point pt = pts[p];
        int64_t cx = sat_i64(floor(pt.px / h)), cy = sat_i64(floor(pt.py / h)), cz = sat_i64(floor(pt.pz / h));
        double tx = pt.px / h - (double)cx, ty = pt.py / h - (double)cy, tz = pt.pz / h - (double)cz;
        double bx[2] = {1.0 - tx, tx}, by[2] = {1.0 - ty, ty}, bz[2] = {1.0 - tz, tz};
        for (int dz = 0; dz <= 1; ++dz) for (int dy = 0; dy <= 1; ++dy) for (int dx = 0; dx <= 1; ++dx) {
            uint64_t ox = (uint64_t)wadd(cx, dx), oy = (uint64_t)wadd(cy, dy), oz = (uint64_t)wadd(cz, dz);
            if (ox >= (uint64_t)res || oy >= (uint64_t)res || oz >= (uint64_t)res) continue;
            int64_t idx = idx_at(g2i, res, ox, oy, oz);
            if (idx < 0) continue;
            double ww = bx[dx] * by[dy] * bz[dz];
            nf[3 * idx] += pt.nx * ww;
            nf[3 * idx + 1] += pt.ny * ww;
            nf[3 * idx + 2] += pt.nz * ww;
            w[idx] += ww;
        }
    }
    for (int64_t i = 0; i < n; ++i) {
        if (w[i] > 0.0) {
            double inv = 1.0 / w[i];
            nf[3 * i] = -nf[3 * i] * inv;
            nf[3 * i + 1] = -nf[3 * i + 1] * inv;
            nf[3 * i + 2] = -nf[3 * i + 2] * inv;
        }
    }
}

/* ---------------------------------------------------------------- CSR
 * SparseMatrix (src/sparse.rs:24-84): row offsets + (col, value) entries. */
typedef struct { int64_t n; int64_t* off; int64_t* col; double* val; } orc_csr;

typedef struct { int64_t col; double val; } ent;

/* stable merge sort by column (Rust's sort_by_key is a stable merge sort) */
static void merge_sort(ent* r, ent* tmp, int64_t len) {
    if (len < 24) {
        for (int64_t i = 1; i < len; ++i) { /* insertion sort: stable */
            ent e = r[i];
            int64_t j = i - 1;
            while (j >= 0 && r[j].col > e.col) { r[j + 1] = r[j]; --j; }
            r[j + 1] = e;
        }
        return;
    }
    int64_t h = len / 2;
    merge_sort(r, tmp, h);
    merge_sort(r + h, tmp, len - h);
    int64_t i = 0, j = h, k = 0;
    while (i < h && j < len) tmp[k++] = (r[j].col < r[i].col) ? r[j++] : r[i++];
    while (i < h) tmp[k++] = r[i++];
    while (j < len) tmp[k++] = r[j++];
    memcpy(r, tmp, len * sizeof(ent));
}

/* sort_by_key (stable) then sequential dedupe-sum (src/fem_tree.rs:229-239). */
static int64_t sort_dedupe(ent* r, int64_t len) {
    if (len >= 24) {
        ent* tmp = (ent*)malloc(len * sizeof(ent));
        merge_sort(r, tmp, len);
        free(tmp);
    } else {
        merge_sort(r, NULL, len);
    }
    int64_t j = 0, u = 0;
    while (j < len) {
        int64_t c = r[j].col;
        double s = 0.0;
        while (j < len && r[j].col == c) { s += r[j].val; ++j; }
        r[u].col = c; r[u].val = s; ++u;
    }
    return u;
}

static orc_csr* csr_from_rows(int64_t n, int64_t* cnt, int64_t* start, ent* buf) {
    orc_csr* m = (orc_csr*)calloc(1, sizeof(orc_csr));
    m->n = n;
    m->off = (int64_t*)malloc((n + 1) * sizeof(int64_t));
    int64_t nnz = 0;
    for (int64_t i = 0; i < n; ++i) { m->off[i] = nnz; nnz += cnt[i]; }
    m->off[n] = nnz;
    m->col = (int64_t*)malloc((nnz ? nnz : 1) * sizeof(int64_t));
    m->val = (double*)malloc((nnz ? nnz : 1) * sizeof(double));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = 0; k < cnt[i]; ++k) { m->col[m->off[i] + k] = buf[start[i] + k].col; m->val[m->off[i] + k] = buf[start[i] + k].val; }
    return m;
}

/* Screening basis of one point (src/fem_tree.rs:206-220): `as isize` truncation
 * (not floor), node offset `(c + d) as u32`, loop dz, dy, dx. */
static int point_basis_trunc(const orc_point* pt, double h, int64_t res, const int64_t* g2i, int64_t* idx, double* phi) {
    int64_t cx = sat_i64(pt->px / h), cy = sat_i64(pt->py / h), cz = sat_i64(pt->pz / h);
    double tx = pt->px / h - (double)cx, ty = pt->py / h - (double)cy, tz = pt->pz / h - (double)cz;
    double bx[2] = {1. - tx, tx}, by[2] = {1. - ty, ty}, bz[2] = {1. - tz, tz};
    int nb = 0;
    for (int dz = 0; dz <= 1; ++dz) for (int dy = 0; dy <= 1; ++dy) for (int dx = 0; dx <= 1; ++dx) {
        uint32_t ox = (uint32_t)(uint64_t)wadd(cx, dx), oy = (uint32_t)(uint64_t)wadd(cy, dy), oz = (uint32_t)(uint64_t)wadd(cz, dz);
        if (ox >= (uint32_t)res || oy >= (uint32_t)res || oz >= (uint32_t)res) continue;
        int64_t id = idx_at(g2i, res, ox, oy, oz);
        if (id < 0) continue;
        idx[nb] = id;
        phi[nb] = bx[dx] * by[dy] * bz[dz];
        ++nb;
    }
    return nb;
}

/* FEMTree::assemble_system (src/fem_tree.rs:180-249).  Push order per row:
 * Laplacian (dx, dy, dz loops, K != 0), then wg on the diagonal, then the
 * 64 screening pairs of every point in input order (no zero filter). */
orc_csr* orc_assemble(uint32_t D, const int64_t* g2i, const uint32_t* i2o, const orc_point* pts, int64_t P, double pw) {
    const int64_t res = (int64_t)1 << D, n = res * res * res;
    const double h = 1.0 / (double)res;
    int64_t* cnt = (int64_t*)calloc(n, sizeof(int64_t));
    /* pass 1: count pushes */
    for (int64_t i = 0; i < n; ++i) {
        int64_t ix = i2o[3 * i], iy = i2o[3 * i + 1], iz = i2o[3 * i + 2];
        for (int64_t dx = -1; dx <= 1; ++dx) for (int64_t dy = -1; dy <= 1; ++dy) for (int64_t dz = -1; dz <= 1; ++dz) {
            int64_t jx = ix + dx, jy = iy + dy, jz = iz + dz;
            if (jx < 0 || jx >= res || jy < 0 || jy >= res || jz < 0 || jz >= res) continue;
            if (idx_at(g2i, res, jx, jy, jz) >= 0 && stiffness_entry(h, dx, dy, dz) != 0.0) cnt[i]++;
        }
    }
    int64_t bidx[8]; double bphi[8];
    if (pw > 0.0) {
        for (int64_t i = 0; i < n; ++i) cnt[i]++;
        for (int64_t p = 0; p < P; ++p) {
            int nb = point_basis_trunc(&pts[p], h, res, g2i, bidx, bphi);
            for (int a = 0; a < nb; ++a) cnt[bidx[a]] += nb;
        }
    }
    int64_t* start = (int64_t*)malloc((n + 1) * sizeof(int64_t));
    int64_t tot = 0;
    for (int64_t i = 0; i < n; ++i) { start[i] = tot; tot += cnt[i]; }
    start[n] = tot;
    ent* buf = (ent*)malloc((tot ? tot : 1) * sizeof(ent));
    int64_t* fill = (int64_t*)calloc(n, sizeof(int64_t));
    /* pass 2: push in the reference's order */
    for (int64_t i = 0; i < n; ++i) {
        int64_t ix = i2o[3 * i], iy = i2o[3 * i + 1], iz = i2o[3 * i + 2];
        for (int64_t dx = -1; dx <= 1; ++dx) for (int64_t dy = -1; dy <= 1; ++dy) for (int64_t dz = -1; dz <= 1; ++dz) {
            int64_t jx = ix + dx, jy = iy + dy, jz = iz + dz;
            if (jx < 0 || jx >= res || jy < 0 || jy >= res || jz < 0 || jz >= res) continue;
            int64_t j = idx_at(g2i, res, jx, jy, jz);
            if (j < 0) continue;
            double v = stiffness_entry(h, dx, dy, dz);
            if (v != 0.0) { buf[start[i] + fill[i]].col = j; buf[start[i] + fill[i]].val = v; fill[i]++; }
        }
    }
    if (pw > 0.0) {
        double wh3 = pw * (h * h * h) / 1.5;
        double wg = wh3 * 0.005;
        for (int64_t i = 0; i < n; ++i) { buf[start[i] + fill[i]].col = i; buf[start[i] + fill[i]].val = wg; fill[i]++; }
        for (int64_t p = 0; p < P; ++p) {
            int nb = point_basis_trunc(&pts[p], h, res, g2i, bidx, bphi);
            for (int a = 0; a < nb; ++a)
                for (int b = 0; b < nb; ++b) {
                    int64_t i = bidx[a];
                    buf[start[i] + fill[i]].col = bidx[b];
                    buf[start[i] + fill[i]].val = wh3 * bphi[a] * bphi[b];
                    fill[i]++;
                }
        }
    }
    for (int64_t i = 0; i < n; ++i) cnt[i] = sort_dedupe(buf + start[i], fill[i]);
    orc_csr* m = csr_from_rows(n, cnt, start, buf);
    free(cnt); free(start); free(buf); free(fill);
    return m;
}

int64_t orc_csr_rows(const orc_csr* m) { return m->n; }
int64_t orc_csr_nnz(const orc_csr* m) { return m->off[m->n]; }
void orc_csr_export(const orc_csr* m, int64_t* off, int64_t* col, double* val) {
    memcpy(off, m->off, (m->n + 1) * sizeof(int64_t));
    memcpy(col, m->col, m->off[m->n] * sizeof(int64_t));
    memcpy(val, m->val, m->off[m->n] * sizeof(double));
}
orc_csr* orc_csr_import(int64_t n, const int64_t* off, const int64_t* col, const double* val) {
    orc_csr* m = (orc_csr*)calloc(1, sizeof(orc_csr));
    m->n = n;
    int64_t nnz = off[n];
    m->off = (int64_t*)malloc((n + 1) * sizeof(int64_t));
    m->col = (int64_t*)malloc((nnz ? nnz : 1) * sizeof(int64_t));
    m->val = (double*)malloc((nnz ? nnz : 1) * sizeof(double));
    memcpy(m->off, off, (n + 1) * sizeof(int64_t));
    memcpy(m->col, col, nnz * sizeof(int64_t));
    memcpy(m->val, val, nnz * sizeof(double));
    return m;
}
void orc_csr_free(orc_csr* m) { if (!m) return; free(m->off); free(m->col); free(m->val); free(m); }

/* ---------------------------------------------------------------- RHS
 * FEMTree::assemble_rhs (src/fem_tree.rs:251-299). */
void orc_rhs(uint32_t D, const int64_t* g2i, const uint32_t* i2o, const double* nf, const orc_point* pts, int64_t P, double pw, double* rhs) {
    const int64_t res = (int64_t)1 << D, n = res * res * res;
    const double h = 1.0 / (double)res;
    for (int64_t i = 0; i < n; ++i) {
        int64_t ix = i2o[3 * i], iy = i2o[3 * i + 1], iz = i2o[3 * i + 2];
        double sum = 0.0;
        for (int64_t dx = -1; dx <= 1; ++dx) for (int64_t dy = -1; dy <= 1; ++dy) for (int64_t dz = -1; dz <= 1; ++dz) {
            int64_t jx = ix + dx, jy = iy + dy, jz = iz + dz;
            if (jx < 0 || jx >= res || jy < 0 || jy >= res || jz < 0 || jz >= res) continue;
            int64_t j = idx_at(g2i, res, jx, jy, jz);
            if (j < 0) continue;
            const double vx = nf[3 * j], vy = nf[3 * j + 1], vz = nf[3 * j + 2];
            sum += h * h * (vx * deriv_mass_1d(dx) * mass_1d(dy) * mass_1d(dz) +
                            vy * mass_1d(dx) * deriv_mass_1d(dy) * mass_1d(dz) +
                            vz * mass_1d(dx) * mass_1d(dy) * deriv_mass_1d(dz));
        }
        rhs[i] = sum;
    }
    if (pw > 0.0) {
        double wh3 = pw * (h * h * h) / 1.5;
        double target = 0.5;
        int64_t bidx[8]; double bphi[8];
        for (int64_t p = 0; p < P; ++p) {
            /* same basis as the screening block, but the product is
             * ((wh3*target)*bx)*by*bz, so recompute the weights here */
            const orc_point* pt = &pts[p];
            int64_t cx = sat_i64(pt->px / h), cy = sat_i64(pt->py / h), cz = sat_i64(pt->pz / h);
            double tx = pt->px / h - (double)cx, ty = pt->py / h - (double)cy, tz = pt->pz / h - (double)cz;
            double bx[2] = {1. - tx, tx}, by[2] = {1. - ty, ty}, bz[2] = {1. - tz, tz};
            (void)bidx; (void)bphi;
            for (int dz = 0; dz <= 1; ++dz) for (int dy = 0; dy <= 1; ++dy) for (int dx = 0; dx <= 1; ++dx) {
                uint32_t ox = (uint32_t)(uint64_t)wadd(cx, dx), oy = (uint32_t)(uint64_t)wadd(cy, dy), oz = (uint32_t)(uint64_t)wadd(cz, dz);
                if (ox >= (uint32_t)res || oy >= (uint32_t)res || oz >= (uint32_t)res) continue;
                int64_t idx = idx_at(g2i, res, ox, oy, oz);
                if (idx < 0) continue;
                rhs[idx] += wh3 * target * bx[dx] * by[dy] * bz[dz];
            }
        }
    }
}

/* ---------------------------------------------------------------- coarse level
 * FEMTree::build_system_at_depth (src/fem_tree.rs:346-404): dense row-major
 * index (z*r+y)*r+x, loops dz, dy, dx; effective_pw = pw * 4^(D-d); diagonal
 * wh3*1.005; RHS = sum of (v.x+v.y+v.z) over fine nodes (FEM index order) in
 * each coarse cell, then + wh3*0.5. */
orc_csr* orc_coarse_system(uint32_t D, uint32_t d, const uint32_t* i2o, const double* nf, double pw, double* rhs) {
    const int64_t res = (int64_t)1 << d, n = res * res * res;
    const double h = 1.0 / (double)res;
    double sf = (double)((uint64_t)1 << (2 * (D - d)));
    double epw = pw * sf;
    int64_t* cnt = (int64_t*)calloc(n, sizeof(int64_t));
    int64_t* start = (int64_t*)malloc((n + 1) * sizeof(int64_t));
    ent* buf = (ent*)malloc((n * 28 + 1) * sizeof(ent));
    for (int64_t i = 0; i < n; ++i) start[i] = i * 28;
    for (int64_t z = 0; z < res; ++z) for (int64_t y = 0; y < res; ++y) for (int64_t x = 0; x < res; ++x) {
        int64_t i = (z * res + y) * res + x;
        for (int64_t dz = -1; dz <= 1; ++dz) for (int64_t dy = -1; dy <= 1; ++dy) for (int64_t dx = -1; dx <= 1; ++dx) {
            int64_t jx = x + dx, jy = y + dy, jz = z + dz;
            if (jx < 0 || jx >= res || jy < 0 || jy >= res || jz < 0 || jz >= res) continue;
            int64_t j = (jz * res + jy) * res + jx;
            double v = stiffness_entry(h, dx, dy, dz);
            if (v != 0.0) { buf[start[i] + cnt[i]].col = j; buf[start[i] + cnt[i]].val = v; cnt[i]++; }
        }
        if (epw > 0.0) {
            double wh3 = epw * (h * h * h) / 1.5;
            buf[start[i] + cnt[i]].col = i; buf[start[i] + cnt[i]].val = wh3 * 1.005; cnt[i]++;
        }
    }
    for (int64_t i = 0; i < n; ++i) cnt[i] = sort_dedupe(buf + start[i], cnt[i]);
    orc_csr* m = csr_from_rows(n, cnt, start, buf);
    free(cnt); free(start); free(buf);
    const int64_t nfine = (int64_t)1 << (3 * D);
    for (int64_t i = 0; i < n; ++i) rhs[i] = 0.0;
    for (int64_t i = 0; i < nfine; ++i) {
        int64_t cx = i2o[3 * i] >> (D - d), cy = i2o[3 * i + 1] >> (D - d), cz = i2o[3 * i + 2] >> (D - d);
        int64_t ci = (cz * res + cy) * res + cx;
        if (ci < n) rhs[ci] += nf[3 * i] + nf[3 * i + 1] + nf[3 * i + 2];
    }
    if (epw > 0.0) {
        double wh3 = epw * (h * h * h) / 1.5;
        for (int64_t i = 0; i < n; ++i) rhs[i] += wh3 * 0.5;
    }
    return m;
}

/* FEMTree::prolongate (src/fem_tree.rs:310-343): coarse row-major c ->
 * fine 2c+delta-1, weights [0.5,1,0.5]^3, skipping v == 0. */
void orc_prolongate(uint32_t D, uint32_t d, const int64_t* g2i, const double* coarse, double* fine) {
    const int64_t fres = (int64_t)1 << D, cres = (int64_t)1 << d;
    const int64_t nf = fres * fres * fres;
    const double bw[3] = {0.5, 1.0, 0.5};
    for (int64_t i = 0; i < nf; ++i) fine[i] = 0.0;
    for (int64_t cz = 0; cz < cres; ++cz) for (int64_t cy = 0; cy < cres; ++cy) for (int64_t cx = 0; cx < cres; ++cx) {
        int64_t ci = (cz * cres + cy) * cres + cx;
        double v = coarse[ci];
        if (v == 0.0) continue;
        for (int64_t dz = 0; dz <= 2; ++dz) {
            int64_t fz = cz * 2 + dz - 1;
            if (fz < 0 || fz >= fres) continue;
            for (int64_t dy = 0; dy <= 2; ++dy) {
                int64_t fy = cy * 2 + dy - 1;
                if (fy < 0 || fy >= fres) continue;
                for (int64_t dx = 0; dx <= 2; ++dx) {
                    int64_t fx = cx * 2 + dx - 1;
                    if (fx < 0 || fx >= fres) continue;
                    double w = bw[dx] * bw[dy] * bw[dz];
                    int64_t fi = idx_at(g2i, fres, fx, fy, fz);
                    if (fi >= 0) fine[fi] += v * w;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------- numerics
 * SparseMatrix::multiply_vector (src/sparse.rs:105-117): rows in parallel,
 * each row summed sequentially from 0.0 in column order. */
void orc_spmv(const orc_csr* a, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < a->n; ++i) {
        double s = 0.0;
        for (int64_t k = a->off[i]; k < a->off[i + 1]; ++k) s = s + a->val[k] * x[a->col[k]];
        y[i] = s;
    }
}

/* solvers::dot / square_norm (src/solvers.rs:6-15): a parallel sum; the
 * reference's reduction tree is rayon's and is not reproducible itself. */
double orc_dot(const double* a, const double* b, int64_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* solvers::gauss_seidel_sweep (src/solvers.rs:164-175): forward, in index
 * order, row skipped when the diagonal is 0. */
void orc_gs_sweep(const orc_csr* a, const double* b, double* x) {
    for (int64_t i = 0; i < a->n; ++i) {
        double diag = 0.0, sum = 0.0;
        for (int64_t k = a->off[i]; k < a->off[i + 1]; ++k) {
            if (a->col[k] == i) diag = a->val[k];
            else sum += a->val[k] * x[a->col[k]];
        }
        if (diag != 0.0) x[i] = (b[i] - sum) / diag;
    }
}

/* The same sweep with rows visited in a caller-given order.  A multicolour
 * sweep is this with the rows sorted by colour; equality with it pins the
 * recoloured GPU sweep. */
void orc_gs_sweep_order(const orc_csr* a, const double* b, double* x, const int64_t* order) {
    for (int64_t k = 0; k < a->n; ++k) {
        const int64_t i = order[k];
        double diag = 0.0, sum = 0.0;
        for (int64_t q = a->off[i]; q < a->off[i + 1]; ++q) {
            if (a->col[q] == i) diag = a->val[q];
            else sum += a->val[q] * x[a->col[q]];
        }
        if (diag != 0.0) x[i] = (b[i] - sum) / diag;
    }
}

/* 1-D hat-function integrals of src/fem_tree.rs:39-41 at d = -1, 0, 1:
 * out = {M(-1),M(0),M(1), S(-1),S(0),S(1), G(-1),G(0),G(1)}. */
void orc_stencil_1d(double* out) {
    for (int d = -1; d <= 1; ++d) {
        out[d + 1] = mass_1d(d);
        out[3 + d + 1] = stiff_1d(d);
        out[6 + d + 1] = deriv_mass_1d(d);
    }
}

/* solvers::solve_cg (src/solvers.rs:20-69): r refreshed as b-Ax on
 * iterations 10, 20, ...; stop when delta <= delta0*eps^2; returns the
 * iteration count.  `final_rel` (extension) gets sqrt(delta/delta0). */
int64_t orc_cg(const orc_csr* a, const double* b, double* x, int64_t max_iters, double eps, double* final_rel) {
    const int64_t n = a->n;
    double* r = (double*)malloc(n * sizeof(double));
    double* d = (double*)malloc(n * sizeof(double));
    double* q = (double*)malloc(n * sizeof(double));
    int64_t ret = max_iters;
    orc_spmv(a, x, q);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
    memcpy(d, r, n * sizeof(double));
    double dn = orc_dot(r, r, n), d0 = dn;
    if (final_rel) *final_rel = 0.0;
    if (dn == 0.0) { ret = 0; goto done; }
    for (int64_t it = 0; it < max_iters; ++it) {
        orc_spmv(a, d, q);
        double dq = orc_dot(d, q, n);
        if (dq == 0.0) { ret = it + 1; goto done; }
        double alpha = dn / dq;
        for (int64_t i = 0; i < n; ++i) x[i] += alpha * d[i];
        if ((it + 1) % 10 == 0) {
            orc_spmv(a, x, q);
            for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
        } else {
            for (int64_t i = 0; i < n; ++i) r[i] -= alpha * q[i];
        }
        double dold = dn;
        dn = orc_dot(r, r, n);
        if (final_rel) *final_rel = sqrt(dn / d0);
        if (dn <= d0 * eps * eps) { ret = it + 1; goto done; }
        double beta = dn / dold;
        if (isnan(beta) || isinf(beta)) break;
        for (int64_t i = 0; i < n; ++i) d[i] = r[i] + beta * d[i];
    }
    ret = max_iters;
done:
    free(r); free(d); free(q);
    return ret;
}

/* ---------------------------------------------------------------- evaluate
 * FEMTree::evaluate (src/fem_tree.rs:490-513), bspline1 (:640-642),
 * contributing_offsets (:645-649). */
static inline double bspline1(double t) { if (t < 0.0 || t > 2.0) return 0.0; if (t <= 1.0) return t; return 2.0 - t; }
static inline int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

double orc_evaluate1(uint32_t D, const int64_t* g2i, const double* sol, double px, double py, double pz) {
    const int64_t res = (int64_t)1 << D;
    const double h = 1.0 / (double)res;
    double s[3] = {px / h, py / h, pz / h};
    int64_t o[3][2]; double b[3][2];
    for (int k = 0; k < 3; ++k) {
        int64_t f = sat_i64(floor(s[k]));
        o[k][0] = clampi(wadd(f, -1), 0, res - 1);
        o[k][1] = clampi(f, 0, res - 1);
        b[k][0] = bspline1(s[k] - (double)o[k][0]);
        b[k][1] = bspline1(s[k] - (double)o[k][1]);
    }
    double v = 0.0;
    for (int a = 0; a < 2; ++a) { if (b[0][a] == 0.0) continue;
        for (int c = 0; c < 2; ++c) { if (b[1][c] == 0.0) continue;
            for (int e = 0; e < 2; ++e) { if (b[2][e] == 0.0) continue;
                int64_t idx = idx_at(g2i, res, o[0][a], o[1][c], o[2][e]);
                if (idx >= 0) v += sol[idx] * b[0][a] * b[1][c] * b[2][e];
            }
        }
    }
    return v;
}

void orc_evaluate(uint32_t D, const int64_t* g2i, const double* sol, const double* q, int64_t m, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) out[i] = orc_evaluate1(D, g2i, sol, q[3 * i], q[3 * i + 1], q[3 * i + 2]);
}

/* iso = mean of evaluate over the transformed points (src/reconstructor.rs:56). */
double orc_iso(uint32_t D, const int64_t* g2i, const double* sol, const orc_point* pts, int64_t P) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < P; ++i) s += orc_evaluate1(D, g2i, sol, pts[i].px, pts[i].py, pts[i].pz);
    return s / (double)P;
}

/* ---------------------------------------------------------------- driver
 * FEMTree::solve_cascadic (src/fem_tree.rs:458-487) with base = D-2 and
 * gs = 20 from src/reconstructor.rs:51-52.  Fills the fine solution (FEM
 * index order) and reports both CG iteration counts. */
typedef struct {
    double t_assemble, t_coarse, t_fine_gs, t_fine_cg;
    int64_t coarse_iters, fine_iters, nnz;
    double coarse_rel, fine_rel;
} orc_solve_report;

static double now_s(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

void orc_solve_cascadic(uint32_t D, const int64_t* g2i, const uint32_t* i2o, const double* nf, const orc_point* pts, int64_t P,
                        double pw, uint32_t base, int64_t gs, int64_t cg, double eps, double* sol, orc_solve_report* rep) {
    const int64_t n = (int64_t)1 << (3 * D);
    orc_solve_report R;
    memset(&R, 0, sizeof R);
    double t0 = now_s();
    orc_csr* A = orc_assemble(D, g2i, i2o, pts, P, pw);
    double* b = (double*)malloc(n * sizeof(double));
    orc_rhs(D, g2i, i2o, nf, pts, P, pw, b);
    R.nnz = orc_csr_nnz(A);
    R.t_assemble = now_s() - t0;
    if (base >= D) {
        for (int64_t i = 0; i < n; ++i) sol[i] = 0.0;
        t0 = now_s();
        for (int64_t k = 0; k < gs; ++k) orc_gs_sweep(A, b, sol);
        R.t_fine_gs = now_s() - t0; t0 = now_s();
        R.fine_iters = orc_cg(A, b, sol, cg, eps, &R.fine_rel);
        R.t_fine_cg = now_s() - t0;
    } else {
        t0 = now_s();
        const int64_t nc = (int64_t)1 << (3 * base);
        double* cb = (double*)malloc(nc * sizeof(double));
        double* cx = (double*)calloc(nc, sizeof(double));
        orc_csr* C = orc_coarse_system(D, base, i2o, nf, pw, cb);
        for (int64_t k = 0; k < gs; ++k) orc_gs_sweep(C, cb, cx);
        R.coarse_iters = orc_cg(C, cb, cx, cg, eps, &R.coarse_rel);
        orc_prolongate(D, base, g2i, cx, sol);
        R.t_coarse = now_s() - t0;
        orc_csr_free(C); free(cb); free(cx);
        t0 = now_s();
        for (int64_t k = 0; k < gs; ++k) orc_gs_sweep(A, b, sol);
        R.t_fine_gs = now_s() - t0; t0 = now_s();
        R.fine_iters = orc_cg(A, b, sol, cg / 4, eps, &R.fine_rel);
        R.t_fine_cg = now_s() - t0;
    }
    orc_csr_free(A); free(b);
    if (rep) *rep = R;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

/* ---------------------------------------------------------------- adjoint
 * Gradient of L(F) with F_p = evaluate(x, q_p) at the splatted points
 * themselves, x = A^-1 b the converged solution, with respect to the unit-cube
 * positions q_p and normals n_p (the transform is held fixed).  Not part of
 * the reference; it differentiates the reference's own forward path:
 *   splat  nf_i = -N_i/W_i, N_i = sum_p n_p w_ip, W_i = sum_p w_ip  (floor cells)
 *   RHS    b_i = h^2 sum_d nf_{i+d}.g(d) + wh3/2 sum_p phi_ip          (trunc cells)
 *   matrix A = L + wg I + sum_p wh3 phi_p phi_p^T
 *   field  F_p = sum_o x_o N1(s-o) (with the reference's clamping)
 * dL/dtheta = gF.dF/dtheta|_x + lam^T (db/dtheta - dA/dtheta x),  A lam = dL/dx.
 * Derivatives of the piecewise-linear weights are taken on the open pieces
 * (kinks at cell faces have measure zero). */
static inline double dbspline1(double t) { if (t > 0.0 && t < 1.0) return 1.0; if (t > 1.0 && t < 2.0) return -1.0; return 0.0; }

static void eval_basis(int64_t res, double h, const double q[3], int64_t o[3][2], double b[3][2], double db[3][2]) {
    for (int k = 0; k < 3; ++k) {
        double s = q[k] / h;
        int64_t f = sat_i64(floor(s));
        o[k][0] = clampi(wadd(f, -1), 0, res - 1);
        o[k][1] = clampi(f, 0, res - 1);
        for (int a = 0; a < 2; ++a) {
            b[k][a] = bspline1(s - (double)o[k][a]);
            db[k][a] = dbspline1(s - (double)o[k][a]) / h;
        }
    }
}

/* corners of a floor (mode 0) or trunc (mode 1) cell: node, weight, d weight/dq */
static int cell_basis(int mode, const double q[3], double h, int64_t res, const int64_t* g2i, int64_t idx[8], double w[8],
                      double dw[8][3]) {
    int64_t c[3];
    double t[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = mode == 0 ? sat_i64(floor(q[a] / h)) : sat_i64(q[a] / h);
        t[a] = q[a] / h - (double)c[a];
    }
    int nb = 0;
    for (int dz = 0; dz <= 1; ++dz) for (int dy = 0; dy <= 1; ++dy) for (int dx = 0; dx <= 1; ++dx) {
        const int d[3] = {dx, dy, dz};
        int ok = 1;
        uint64_t o[3];
        for (int a = 0; a < 3; ++a) {
            if (mode == 0) { o[a] = (uint64_t)wadd(c[a], d[a]); if (o[a] >= (uint64_t)res) ok = 0; }
            else { uint32_t u = (uint32_t)(uint64_t)wadd(c[a], d[a]); o[a] = u; if (u >= (uint32_t)res) ok = 0; }
        }
        if (!ok) continue;
        double bb[3], db[3];
        for (int a = 0; a < 3; ++a) { bb[a] = d[a] ? t[a] : 1.0 - t[a]; db[a] = (d[a] ? 1.0 : -1.0) / h; }
        idx[nb] = idx_at(g2i, res, o[0], o[1], o[2]);
        w[nb] = bb[0] * bb[1] * bb[2];
        dw[nb][0] = db[0] * bb[1] * bb[2];
        dw[nb][1] = bb[0] * db[1] * bb[2];
        dw[nb][2] = bb[0] * bb[1] * db[2];
        ++nb;
    }
    return nb;
}

int64_t orc_vjp_points(uint32_t D, const int64_t* g2i, const uint32_t* i2o, const orc_point* pts, int64_t P, double pw,
                       const double* x, const double* gF, double cg_eps, double* lam_out, double* gpos, double* gnrm) {
    const int64_t res = (int64_t)1 << D, n = res * res * res;
    const double h = 1.0 / (double)res;
    double* nf = (double*)malloc(3 * n * sizeof(double));
    double* W = (double*)malloc(n * sizeof(double));
    orc_splat(D, g2i, pts, P, nf, W);
    orc_csr* A = orc_assemble(D, g2i, i2o, pts, P, pw);
    double* gx = (double*)calloc(n, sizeof(double));
    for (int64_t p = 0; p < P; ++p) {
        const double q[3] = {pts[p].px, pts[p].py, pts[p].pz};
        int64_t o[3][2]; double b[3][2], db[3][2];
        eval_basis(res, h, q, o, b, db);
        for (int a = 0; a < 2; ++a) for (int c = 0; c < 2; ++c) for (int e = 0; e < 2; ++e) {
            int64_t i = idx_at(g2i, res, o[0][a], o[1][c], o[2][e]);
            gx[i] += gF[p] * b[0][a] * b[1][c] * b[2][e];
        }
    }
    double* lam = (double*)calloc(n, sizeof(double));
    int64_t its = orc_cg(A, gx, lam, 200000, cg_eps, NULL);
    /* d L / d nf_j = h^2 sum_{i = j - d} lam_i g(d) */
    double* gnf = (double*)calloc(3 * n, sizeof(double));
    for (int64_t j = 0; j < n; ++j) {
        int64_t jx = i2o[3 * j], jy = i2o[3 * j + 1], jz = i2o[3 * j + 2];
        for (int64_t dx = -1; dx <= 1; ++dx) for (int64_t dy = -1; dy <= 1; ++dy) for (int64_t dz = -1; dz <= 1; ++dz) {
            int64_t ix = jx - dx, iy = jy - dy, iz = jz - dz;
            if (ix < 0 || ix >= res || iy < 0 || iy >= res || iz < 0 || iz >= res) continue;
            double l = lam[idx_at(g2i, res, ix, iy, iz)] * (h * h);
            gnf[3 * j] += l * deriv_mass_1d(dx) * mass_1d(dy) * mass_1d(dz);
            gnf[3 * j + 1] += l * mass_1d(dx) * deriv_mass_1d(dy) * mass_1d(dz);
            gnf[3 * j + 2] += l * mass_1d(dx) * mass_1d(dy) * deriv_mass_1d(dz);
        }
    }
    const double wh3 = pw > 0.0 ? pw * (h * h * h) / 1.5 : 0.0;
    for (int64_t p = 0; p < P; ++p) {
        const double q[3] = {pts[p].px, pts[p].py, pts[p].pz};
        const double nv[3] = {pts[p].nx, pts[p].ny, pts[p].nz};
        double gq[3] = {0, 0, 0}, gn[3] = {0, 0, 0};
        /* explicit dF/dq */
        int64_t o[3][2]; double b[3][2], db[3][2];
        eval_basis(res, h, q, o, b, db);
        for (int a = 0; a < 2; ++a) for (int c = 0; c < 2; ++c) for (int e = 0; e < 2; ++e) {
            double xv = x[idx_at(g2i, res, o[0][a], o[1][c], o[2][e])];
            gq[0] += gF[p] * xv * db[0][a] * b[1][c] * b[2][e];
            gq[1] += gF[p] * xv * b[0][a] * db[1][c] * b[2][e];
            gq[2] += gF[p] * xv * b[0][a] * b[1][c] * db[2][e];
        }
        /* splat */
        int64_t idx[8]; double w[8], dw[8][3];
        int nb = cell_basis(0, q, h, res, g2i, idx, w, dw);
        for (int k = 0; k < nb; ++k) {
            int64_t i = idx[k];
            if (!(W[i] > 0.0)) continue;
            const double* g = gnf + 3 * i;
            for (int c = 0; c < 3; ++c) gn[c] += -(w[k] / W[i]) * g[c];
            double s = 0.0;
            for (int c = 0; c < 3; ++c) s += g[c] * (-(nv[c] + nf[3 * i + c]) / W[i]);
            for (int a = 0; a < 3; ++a) gq[a] += s * dw[k][a];
        }
        /* screening RHS and matrix */
        if (pw > 0.0) {
            nb = cell_basis(1, q, h, res, g2i, idx, w, dw);
            double px = 0.0, pl = 0.0, dpx[3] = {0, 0, 0}, dpl[3] = {0, 0, 0};
            for (int k = 0; k < nb; ++k) {
                px += w[k] * x[idx[k]];
                pl += w[k] * lam[idx[k]];
                for (int a = 0; a < 3; ++a) { dpx[a] += dw[k][a] * x[idx[k]]; dpl[a] += dw[k][a] * lam[idx[k]]; }
            }
            for (int a = 0; a < 3; ++a) gq[a] += wh3 * 0.5 * dpl[a] - wh3 * (dpl[a] * px + pl * dpx[a]);
        }
        for (int a = 0; a < 3; ++a) { gpos[3 * p + a] = gq[a]; gnrm[3 * p + a] = gn[a]; }
    }
    if (lam_out) memcpy(lam_out, lam, n * sizeof(double));
    free(nf); free(W); free(gx); free(lam); free(gnf); orc_csr_free(A);
    return its;
}

/* F at the points after a full tight solve: the forward of the vjp above. */
void orc_field_at_points(uint32_t D, const int64_t* g2i, const uint32_t* i2o, const orc_point* pts, int64_t P, double pw,
                         double eps, double* x_out, double* F) {
    const int64_t n = (int64_t)1 << (3 * D);
    double* nf = (double*)malloc(3 * n * sizeof(double));
    double* W = (double*)malloc(n * sizeof(double));
    orc_splat(D, g2i, pts, P, nf, W);
    orc_csr* A = orc_assemble(D, g2i, i2o, pts, P, pw);
    double* b = (double*)malloc(n * sizeof(double));
    orc_rhs(D, g2i, i2o, nf, pts, P, pw, b);
    double* x = (double*)calloc(n, sizeof(double));
    orc_cg(A, b, x, 200000, eps, NULL);
    for (int64_t p = 0; p < P; ++p) F[p] = orc_evaluate1(D, g2i, x, pts[p].px, pts[p].py, pts[p].pz);
    if (x_out) memcpy(x_out, x, n * sizeof(double));
    free(nf); free(W); free(b); free(x); orc_csr_free(A);
}


/* ------------------------------------------------------------------ surface extraction
 * FEMTree::extract_surface (src/fem_tree.rs:576-636): leaves in index order
 * (Morton), corner values from evaluate, case bit k when cv[k] < iso, vertices
 * interpolated on the cell's own edge (t = 0.5 when |dv| < 1e-15) and
 * deduplicated by the position quantised as (p*1e9) as u64 (first insertion
 * keeps its index), triangles from the case table.  The case table is the
 * product's (mc_cases.h); it is pinned by its own invariants in the tests. */
static uint64_t rust_as_u64(double v) {
    if (!(v > 0.0)) return 0;                         /* NaN and negatives saturate to 0 */
    if (v >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)v;
}
typedef struct { uint64_t k[3]; int64_t v; } mc_slot;
typedef struct { mc_slot* s; int64_t cap, n; } mc_map;
static uint64_t mc_hash(const uint64_t k[3]) {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 3; ++i) { h ^= k[i] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); h *= 1099511628211ull; }
    return h;
}
static int64_t* mc_find(mc_map* m, const uint64_t k[3]) {
    uint64_t i = mc_hash(k) & (uint64_t)(m->cap - 1);
    for (;;) {
        mc_slot* e = &m->s[i];
        if (e->v < 0) return &e->v;
        if (e->k[0] == k[0] && e->k[1] == k[1] && e->k[2] == k[2]) return &e->v;
        i = (i + 1) & (uint64_t)(m->cap - 1);
    }
}
static void mc_grow(mc_map* m) {
    mc_map g = {(mc_slot*)malloc(sizeof(mc_slot) * (size_t)(m->cap * 2)), m->cap * 2, m->n};
    for (int64_t i = 0; i < g.cap; ++i) g.s[i].v = -1;
    for (int64_t i = 0; i < m->cap; ++i)
        if (m->s[i].v >= 0) { int64_t* slot = mc_find(&g, m->s[i].k); mc_slot* e = (mc_slot*)((char*)slot - offsetof(mc_slot, v)); *e = m->s[i]; }
    free(m->s);
    *m = g;
}

int64_t orc_extract_surface(uint32_t D, const int64_t* g2i, const uint32_t* i2o, const double* sol, double iso,
                            double** verts_out, int64_t** tris_out, int64_t* nt_out) {
    const int64_t res = (int64_t)1 << D, N = res * res * res;
    const double w = 1.0 / (double)res;
    int64_t vcap = 1024, tcap = 1024, nv = 0, nt = 0;
    double* V = (double*)malloc(sizeof(double) * 3 * (size_t)vcap);
    int64_t* Tr = (int64_t*)malloc(sizeof(int64_t) * 3 * (size_t)tcap);
    mc_map map = {(mc_slot*)malloc(sizeof(mc_slot) * 1024), 1024, 0};
    for (int64_t i = 0; i < map.cap; ++i) map.s[i].v = -1;
    for (int64_t idx = 0; idx < N; ++idx) {
        const uint64_t off[3] = {i2o[3 * idx], i2o[3 * idx + 1], i2o[3 * idx + 2]};
        double cv[8];
        int mc = 0;
        for (int k = 0; k < 8; ++k) {
            const double px = (double)(off[0] + kMcCorner[k][0]) * w, py = (double)(off[1] + kMcCorner[k][1]) * w,
                         pz = (double)(off[2] + kMcCorner[k][2]) * w;
            cv[k] = orc_evaluate1(D, g2i, sol, px, py, pz);
            if (cv[k] < iso) mc |= 1 << k;
        }
        int64_t evi[12];
        int any = 0;
        for (int e = 0; e < 12; ++e) {
            const int a = kMcEdge[e][0], b = kMcEdge[e][1];
            evi[e] = -1;
            if (((mc >> a) & 1) == ((mc >> b) & 1)) continue;
            any = 1;
            const double dv = cv[b] - cv[a];
            const double t = fabs(dv) < 1e-15 ? 0.5 : (iso - cv[a]) / dv;
            double pos[3];
            for (int d = 0; d < 3; ++d) {
                const double A = (double)(off[d] + kMcCorner[a][d]) * w;
                pos[d] = A + t * (double)((int)kMcCorner[b][d] - (int)kMcCorner[a][d]) * w;
            }
            const uint64_t key[3] = {rust_as_u64(pos[0] * 1e9), rust_as_u64(pos[1] * 1e9), rust_as_u64(pos[2] * 1e9)};
            int64_t* slot = mc_find(&map, key);
            if (*slot >= 0) { evi[e] = *slot; continue; }
            mc_slot* ent = (mc_slot*)((char*)slot - offsetof(mc_slot, v));
            ent->k[0] = key[0]; ent->k[1] = key[1]; ent->k[2] = key[2];
            ent->v = nv;
            evi[e] = nv;
            if (++map.n * 2 > map.cap) mc_grow(&map);
            if (nv == vcap) { vcap *= 2; V = (double*)realloc(V, sizeof(double) * 3 * (size_t)vcap); }
            V[3 * nv] = pos[0]; V[3 * nv + 1] = pos[1]; V[3 * nv + 2] = pos[2];
            ++nv;
        }
        if (!any) continue;
        const char* cs = kMcCase[mc];
        for (int i = 0; cs[i]; i += 3) {
            if (nt == tcap) { tcap *= 2; Tr = (int64_t*)realloc(Tr, sizeof(int64_t) * 3 * (size_t)tcap); }
            for (int j = 0; j < 3; ++j) {
                const char c = cs[i + j];
                const int e = c <= '9' ? c - '0' : c - 'a' + 10;
                Tr[3 * nt + j] = evi[e];
            }
            ++nt;
        }
    }
    free(map.s);
    *verts_out = V;
    *tris_out = Tr;
    *nt_out = nt;
    return nv;
}
void orc_free(void* p) { free(p); }
