"""Pins of the FP64 oracle (CPU, no GPU): the reference's own known-answer
tests, closed forms and invariants that a plausible restatement slip (a
dropped term, a wrong sign, a transposed index, floor vs trunc) would fail.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle
from poissonreconstruction_amd import synthetic


def morton(x, y, z):
    m = 0
    for b in range(11):
        m |= (((x >> b) & 1) << (3 * b)) | (((y >> b) & 1) << (3 * b + 1)) | (((z >> b) & 1) << (3 * b + 2))
    return m


def unit_points(n, seed=0, kind="sphere"):
    raw = synthetic.sphere(n, seed) if kind == "sphere" else synthetic.torus(n, seed=seed)
    m, _ = oracle.unit_transform(raw, 1.1)
    return oracle.apply_transform(raw, m)


# ----------------------------------------------------------------- tree (a3-a5)
@pytest.mark.parametrize("D", [0, 1, 2, 3, 4])
def test_tree_dfs_numbering_is_morton(D):
    g2i, i2o = oracle.tree(D)
    res = 1 << D
    assert len(i2o) == 8 ** D
    for i in range(8 ** D):
        x, y, z = map(int, i2o[i])
        assert morton(x, y, z) == i
        assert g2i[z * res * res + y * res + x] == i


def test_tree_counts_kat():
    # src/octree.rs:436-450: a tree refined to depth 2 has 1 + 8 + 64 nodes
    assert sum(8 ** d for d in range(3)) == 73
    assert len(oracle.tree(2)[1]) == 64


# ----------------------------------------------------------------- stencil (a6)
def _hat_integrals():
    """Exact integrals of unit-spacing hat functions phi(t) = max(0, 1-|t|):
    M(d) = int phi(t) phi(t-d), S(d) = int phi'(t) phi'(t-d),
    G(d) = int phi'(t) phi(t-d)  (piecewise-linear, integrated exactly)."""
    def pieces(shift):  # phi(t - shift) on [shift-1, shift+1] as (a,b, c0,c1) with value c0 + c1*t
        return [(shift - 1, shift, Fr(1 - shift), Fr(1)), (shift, shift + 1, Fr(1 + shift), Fr(-1))]

    def integ(f, g, df, dg):
        tot = Fr(0)
        for (a1, b1, p0, p1) in pieces(0):
            for (a2, b2, q0, q1) in pieces(f):
                lo, hi = max(a1, a2), min(b1, b2)
                if lo >= hi:
                    continue
                u0, u1 = (p1, Fr(0)) if df else (p0, p1)
                v0, v1 = (q1, Fr(0)) if dg else (q0, q1)
                # integral of (u0+u1 t)(v0+v1 t) on [lo,hi]
                c0, c1, c2 = u0 * v0, u0 * v1 + u1 * v0, u1 * v1
                F = lambda t: c0 * t + c1 * Fr(t) ** 2 / 2 + c2 * Fr(t) ** 3 / 3  # noqa: E731
                tot += F(hi) - F(lo)
        return tot
    M = {d: integ(d, 0, False, False) for d in (-1, 0, 1)}
    S = {d: integ(d, 0, True, True) for d in (-1, 0, 1)}
    G = {d: integ(d, 0, True, False) for d in (-1, 0, 1)}
    return M, S, G


def test_stencil_1d_constants_are_hat_integrals():
    M, S, G = _hat_integrals()
    o = oracle.stencil_1d()
    for k, d in enumerate((-1, 0, 1)):
        assert o["M"][k] == float(M[d])
        assert o["S"][k] == float(S[d])
        assert o["G"][k] == float(G[d])
    # src/fem_tree.rs:41 sign convention: G(+1) = -1/2
    assert o["G"][2] == -0.5 and o["G"][0] == 0.5


def test_bspline_kats():
    # src/bspline.rs:399-437 canonical degree-1 B-spline N1(t) (t in [0,2]) and partition of unity;
    # the oracle's evaluate uses N1, so pin it through evaluate's closed form below.
    def n1(t):
        return 0.0 if (t < 0 or t > 2) else (t if t <= 1 else 2 - t)
    assert n1(0.5) == 0.5 and n1(1.0) == 1.0
    for x in (0.0, 0.25, 0.5, 0.75, 1.0):
        assert abs(sum(n1(x - k + 1) for k in range(-1, 3)) - 1.0) < 1e-12


def test_stiffness_classes():
    L = oracle.lib()
    h = 1 / 512
    for d in [(0, 0, 0)]:
        assert L.orc_stiffness_entry(h, *d) == pytest.approx(8 * h / 3, rel=1e-15)
    for d in [(1, 1, 0), (0, -1, 1), (1, 0, -1)]:
        assert L.orc_stiffness_entry(h, *d) == pytest.approx(-h / 6, rel=1e-15)
    for d in [(1, 1, 1), (-1, 1, -1)]:
        assert L.orc_stiffness_entry(h, *d) == pytest.approx(-h / 12, rel=1e-15)
    for d in [(1, 0, 0), (0, 0, -1)]:
        assert L.orc_stiffness_entry(h, *d) == 0.0  # exactly zero faces (Q6)


# ----------------------------------------------------------------- assembly (a8)
@pytest.mark.parametrize("D", [0, 1, 2, 3, 4])
def test_laplacian_energy_closed_form(D):
    # 1^T K 1 = int |grad u|^2 for u = sum of all truncated hats = 6 h (res - 1/3)^2
    g2i, i2o = oracle.tree(D)
    A = oracle.assemble(D, g2i, i2o, None, 0.0)
    S = A.scipy()
    h, res = 1.0 / (1 << D), 1 << D
    assert S.sum() == pytest.approx(6 * h * (res - 1.0 / 3) ** 2, rel=1e-12)
    # interior rows: 21 entries, zero row sum; K symmetric
    nnz = np.diff(A.arrays()[0])
    inner = [i for i in range(8 ** D) if all(0 < c < res - 1 for c in i2o[i])]
    assert all(nnz[i] == 21 for i in inner)
    rs = np.asarray(S.sum(axis=1)).ravel()
    assert all(abs(rs[i]) < 1e-15 for i in inner)
    assert abs(S - S.T).max() == 0.0


def test_laplacian_kills_linear_functions_in_the_interior():
    D = 4
    g2i, i2o = oracle.tree(D)
    A = oracle.assemble(D, g2i, i2o, None, 0.0)
    x = i2o[:, 0] * 0.3 + i2o[:, 1] * 1.7 - i2o[:, 2] * 0.9
    y = A.spmv(x.astype(np.float64))
    res = 1 << D
    inner = np.all((i2o > 0) & (i2o < res - 1), axis=1)
    assert np.abs(y[inner]).max() < 1e-14


def test_screening_sum_identity_and_rank_one_block():
    D, pw = 4, 4.0
    pts = unit_points(300, seed=1)
    g2i, i2o = oracle.tree(D)
    A0 = oracle.assemble(D, g2i, i2o, None, 0.0).scipy()
    A = oracle.assemble(D, g2i, i2o, pts, pw).scipy()
    h = 1.0 / (1 << D)
    wh3 = pw * h ** 3 / 1.5
    # per-point weight sums with trunc cells (Q5): sum_k phi = prod over axes of
    # (in-range fraction); all-in-range points give exactly 1
    sres = 1 << D
    tot = 0.0
    for p in pts:
        s = np.array([p["px"], p["py"], p["pz"]]) / h
        c = np.trunc(s)
        t = s - c
        w = 1.0
        for a in range(3):
            wa = 0.0
            if 0 <= c[a] < sres:
                wa += 1 - t[a]
            if 0 <= c[a] + 1 < sres:
                wa += t[a]
            w *= wa
        tot += w * w
    assert (A - A0).sum() == pytest.approx(wh3 * tot + 0.005 * wh3 * 8 ** D, rel=1e-10)
    # one interior point: the block is rank one
    p1 = oracle.as_points([[0.53, 0.41, 0.37]], [[0, 0, 1.0]])
    B = (oracle.assemble(D, g2i, i2o, p1, pw).scipy() - A0).toarray()
    B -= np.eye(8 ** D) * 0.005 * wh3
    sv = np.linalg.svd(B[np.ix_(np.any(B != 0, 1), np.any(B != 0, 0))], compute_uv=False)
    assert sv[1] < 1e-12 * sv[0]


def test_screening_truncation_not_floor():
    # Q5: s in (-1, 0) truncates to cell 0, giving node 1 a negative weight
    D = 3
    h = 1.0 / 8
    p = oracle.as_points([[-0.25 * h, 4.5 * h, 4.5 * h]], [[1.0, 0, 0]])
    g2i, i2o = oracle.tree(D)
    B = oracle.assemble(D, g2i, i2o, p, 4.0).scipy() - oracle.assemble(D, g2i, i2o, None, 4.0).scipy()
    i0 = g2i[4 * 64 + 4 * 8 + 0]
    i1 = g2i[4 * 64 + 4 * 8 + 1]
    assert B[i0, i1] < 0  # phi(node 0) > 0, phi(node 1) < 0
    nf, w = oracle.splat(D, g2i, p)  # the splat floors: only node 0 column receives weight
    assert w[i1] == 0.0 and w[i0] > 0


# ----------------------------------------------------------------- splat / RHS (a7, a9)
def test_splat_constant_normal_and_partition_of_unity():
    D = 4
    pts = unit_points(500, seed=3)
    pts["nx"], pts["ny"], pts["nz"] = 0.6, 0.0, 0.8
    g2i, _ = oracle.tree(D)
    nf, w = oracle.splat(D, g2i, pts)
    touched = w > 0
    assert np.allclose(nf[touched], [-0.6, 0.0, -0.8], atol=1e-15)
    inside = np.all([(pts[c] >= 0) & (pts[c] < 1 - 1.0 / 16) for c in ("px", "py", "pz")], axis=0)
    # total weight = sum over points of the product over axes of the in-range
    # 1-D weights (floor cells, Q4): 1 for a point with all 8 nodes in range,
    # a fraction for one near the high faces, 0 for one outside the cube
    assert w.sum() == pytest.approx(_weight_sum(pts, D, cell=np.floor), rel=1e-12)
    assert w.sum() < len(pts) and inside.sum() < len(pts)
    # partition of unity for points whose 8 corners are all in range
    sub = pts[inside]
    _, w2 = oracle.splat(D, g2i, sub)
    assert w2.sum() == pytest.approx(len(sub), rel=1e-12)


def test_rhs_closed_forms():
    D = 4
    g2i, i2o = oracle.tree(D)
    h = 1.0 / (1 << D)
    res = 1 << D
    inner = np.all((i2o > 0) & (i2o < res - 1), axis=1)
    # constant field -> interior RHS is 0 up to rounding of the sequential sum (G antisymmetric)
    nf = np.tile([0.3, -0.2, 0.7], (8 ** D, 1))
    b = oracle.rhs(D, g2i, i2o, nf, None, 0.0)
    assert np.abs(b[inner]).max() <= 1e-15 * h ** 2
    # a sign slip in G (+1/2 at d=+1) would give -2x the value below instead of 0
    nf_lin = np.zeros((8 ** D, 3))
    nf_lin[:, 1] = i2o[:, 1] * h
    b_lin = oracle.rhs(D, g2i, i2o, nf_lin, None, 0.0)
    assert np.allclose(b_lin[inner], -h ** 3, rtol=1e-12)
    # v = (x, 0, 0) -> b = -h^3 in the interior (b_i = -int div(v) phi_i)
    nf = np.zeros((8 ** D, 3))
    nf[:, 0] = i2o[:, 0] * h
    b = oracle.rhs(D, g2i, i2o, nf, None, 0.0)
    assert np.allclose(b[inner], -h ** 3, rtol=1e-12)
    # v = (0, 0, z) the same through the z-derivative
    nf = np.zeros((8 ** D, 3))
    nf[:, 2] = i2o[:, 2] * h
    b = oracle.rhs(D, g2i, i2o, nf, None, 0.0)
    assert np.allclose(b[inner], -h ** 3, rtol=1e-12)
    # screening RHS sums to wh3 * 0.5 * sum of the in-range weights
    pts = unit_points(200, seed=4)
    b0 = oracle.rhs(D, g2i, i2o, np.zeros((8 ** D, 3)), pts, 0.0)
    b1 = oracle.rhs(D, g2i, i2o, np.zeros((8 ** D, 3)), pts, 4.0)
    A0 = oracle.assemble(D, g2i, i2o, None, 4.0).scipy()
    A1 = oracle.assemble(D, g2i, i2o, pts, 4.0).scipy()
    wh3 = 4.0 * h ** 3 / 1.5
    # sum_i b_i = wh3/2 sum_p sum_k phi_pk, and 1^T S 1 = wh3 sum_p (sum_k phi_pk)^2: check b against S
    # on points whose weights sum to 1: b_scr = 0.5 * S 1 restricted to them
    assert (b1 - b0).sum() == pytest.approx(0.5 * wh3 * _weight_sum(pts, D), rel=1e-12)
    assert abs((A1 - A0).sum() - 0.0) > 0


def _weight_sum(pts, D, cell=np.trunc):
    h = 1.0 / (1 << D)
    res = 1 << D
    tot = 0.0
    for p in pts:
        s = np.array([p["px"], p["py"], p["pz"]]) / h
        c = cell(s)
        t = s - c
        w = 1.0
        for a in range(3):
            w *= (1 - t[a] if 0 <= c[a] < res else 0.0) + (t[a] if 0 <= c[a] + 1 < res else 0.0)
        tot += w
    return tot


# ----------------------------------------------------------------- coarse level (a10, a11)
def test_coarse_system_closed_forms():
    D, d, pw = 5, 3, 4.0
    pts = unit_points(1000, seed=5)
    g2i, i2o = oracle.tree(D)
    nf, _ = oracle.splat(D, g2i, pts)
    C, cb = oracle.coarse_system(D, d, i2o, nf, pw)
    hc = 1.0 / (1 << d)
    wh3c = pw * 4 ** (D - d) * hc ** 3 / 1.5
    assert cb.sum() == pytest.approx(nf.sum() + 8 ** d * wh3c * 0.5, rel=1e-10, abs=1e-12)
    S = C.scipy()
    res = 1 << d
    assert (S.sum() - 8 ** d * wh3c * 1.005) == pytest.approx(6 * hc * (res - 1.0 / 3) ** 2, rel=1e-12)
    # row-major order: edge (x+1,y+1) is column i+res+1, corner i+res^2+res+1; face i+1 is absent
    off, col, _ = C.arrays()
    i = (2 * res + 3) * res + 4
    row = col[off[i]:off[i + 1]]
    assert i + res + 1 in row and i + res * res + res + 1 in row and i + 1 not in row


def test_prolongation_closed_form_q11():
    # coarse ones -> fine values prod g(f), g = 1 below res/2 - 1, 1/2 at res/2 - 1, 0 beyond
    D, d = 5, 3
    g2i, i2o = oracle.tree(D)
    fine = oracle.prolongate(D, d, g2i, np.ones(8 ** d))
    res = 1 << D

    def g(f):
        return 1.0 if f < res // 2 - 1 else (0.5 if f == res // 2 - 1 else 0.0)
    expect = np.array([g(x) * g(y) * g(z) for x, y, z in i2o])
    assert np.array_equal(fine, expect)


def test_coarse_rhs_row_major_cell_placement():
    # build_system_at_depth (src/fem_tree.rs:387-397): fine node (fx,fy,fz)
    # adds v.x+v.y+v.z to the coarse cell (f >> (D-d)) at row-major
    # (cz*r+cy)*r+cx.  Distinct coordinates per axis, so an x-major index or a
    # wrong shift moves the value to another cell.
    D, d = 5, 3
    g2i, i2o = oracle.tree(D)
    res, r = 1 << D, 1 << d
    nf = np.zeros((8 ** D, 3))
    for (fx, fy, fz), v in (((5, 9, 22), (0.125, 0.25, 0.5)), ((30, 1, 12), (0.0625, 0.0, 0.0))):
        nf[g2i[(fz * res + fy) * res + fx]] = v
    _, cb = oracle.coarse_system(D, d, i2o, nf, 0.0)
    expect = np.zeros(8 ** d)
    expect[(5 * r + 2) * r + 1] = 0.875   # cell (1, 2, 5)
    expect[(3 * r + 0) * r + 7] = 0.0625  # cell (7, 0, 3)
    assert np.array_equal(cb, expect)


def test_prolongation_impulse_row_major():
    # prolongate (src/fem_tree.rs:310-343): coarse row-major (cz*r+cy)*r+cx
    # -> fine 2c+delta-1 with weights [0.5,1,0.5] per axis (Q11)
    D, d = 5, 3
    g2i, _ = oracle.tree(D)
    res, r = 1 << D, 1 << d
    cx, cy, cz = 1, 2, 3
    coarse = np.zeros(8 ** d)
    coarse[(cz * r + cy) * r + cx] = 1.0
    fine = oracle.prolongate(D, d, g2i, coarse)
    bw = (0.5, 1.0, 0.5)
    expect = np.zeros(8 ** D)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                fx, fy, fz = 2 * cx + dx - 1, 2 * cy + dy - 1, 2 * cz + dz - 1
                expect[g2i[(fz * res + fy) * res + fx]] = bw[dx] * bw[dy] * bw[dz]
    assert np.array_equal(fine, expect)


def test_prolongation_reproduces_a_linear_field():
    # trilinear prolongation is exact on linear fields: with coarse
    # v(c) = 1 + 2cx + 5cy + 11cz, every fine node whose three offsets are
    # <= 2(r-1) (full support under Q11) holds v(f/2) = 1 + fx + 2.5fy + 5.5fz,
    # exactly (all weights are powers of two).  Distinct slopes per axis pin
    # the coarse index order.
    D, d = 5, 3
    g2i, i2o = oracle.tree(D)
    r = 1 << d
    c = np.arange(8 ** d)
    cx, cy, cz = c % r, (c // r) % r, c // (r * r)
    fine = oracle.prolongate(D, d, g2i, 1.0 + 2 * cx + 5 * cy + 11 * cz)
    f = i2o.astype(np.float64)
    full = np.all(i2o <= 2 * (r - 1), axis=1)
    assert full.sum() == (2 * r - 1) ** 3
    assert np.array_equal(fine[full], 1.0 + f[full, 0] + 2.5 * f[full, 1] + 5.5 * f[full, 2])
    beyond = np.any(i2o > 2 * r - 1, axis=1)
    assert np.all(fine[beyond] == 0.0)


def test_splat_weight_orientation_single_point():
    # splat_normal_field (src/fem_tree.rs:137-158): c = floor(s), t = s - c;
    # node c + (dx,dy,dz) gets (dx ? t : 1-t) per axis (node o peaks at s = o, Q4)
    D = 3
    g2i, _ = oracle.tree(D)
    res, h = 1 << D, 1.0 / (1 << D)
    s = np.array([2.3, 3.6, 4.8])
    p = oracle.as_points([s * h], [[0.0, 0.0, 1.0]])
    _, w = oracle.splat(D, g2i, p)
    t = s - np.floor(s)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                e = ((t[0] if dx else 1 - t[0]) * (t[1] if dy else 1 - t[1]) * (t[2] if dz else 1 - t[2]))
                node = g2i[((4 + dz) * res + 3 + dy) * res + 2 + dx]
                assert w[node] == pytest.approx(e, rel=1e-12)
    assert w[g2i[(4 * res + 3) * res + 2]] == pytest.approx(0.7 * 0.4 * 0.2, rel=1e-12)
    assert np.count_nonzero(w) == 8


def test_splat_normal_is_negated_weighted_average():
    # two points with distinct normals share node (3,4,5): nf = -(n1 w1 + n2 w2) / (w1 + w2)
    D = 3
    g2i, _ = oracle.tree(D)
    res, h = 1 << D, 1.0 / (1 << D)
    s1, s2 = np.array([2.3, 3.6, 4.8]), np.array([3.5, 4.25, 5.5])
    p = oracle.as_points([s1 * h, s2 * h], [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    nf, w = oracle.splat(D, g2i, p)
    node = g2i[(5 * res + 4) * res + 3]
    w1 = 0.3 * 0.6 * 0.8
    w2 = 0.5 * 0.75 * 0.5
    assert w[node] == pytest.approx(w1 + w2, rel=1e-12)
    assert nf[node] == pytest.approx([-w1 / (w1 + w2), -w2 / (w1 + w2), 0.0], rel=1e-12, abs=1e-15)
    # node (2,3,4) is reached by the first point only
    assert nf[g2i[(4 * res + 3) * res + 2]] == pytest.approx([-1.0, 0.0, 0.0], abs=1e-15)


# ----------------------------------------------------------------- GS / CG / SpMV (a13-a16)
def test_spmv_kat():
    # src/sparse.rs:174-196
    A = oracle.CSR.from_arrays([0, 2, 3, 4], [0, 1, 1, 2], [2.0, 1.0, 3.0, 4.0])
    assert np.allclose(A.spmv([1.0, 2.0, 3.0]), [4.0, 6.0, 12.0], atol=1e-10)


def test_cg_kat():
    # src/solvers.rs:234-268: tridiag(-1,2,-1), n=10, b=1 -> x_i = (i+1)(n-i)/2
    n = 10
    off, col, val = [0], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                col.append(j)
                val.append(v)
        off.append(len(col))
    A = oracle.CSR.from_arrays(off, col, val)
    x = np.zeros(n)
    it, _ = A.cg(np.ones(n), x, 100, 1e-10)
    assert np.abs(x - [(i + 1) * (n - i) / 2 for i in range(n)]).max() < 1e-6
    assert it <= n


def test_gs_hand_sweep_and_fixed_point():
    # 2x2: [[4,1],[1,3]] x = [1,2]; one sweep from 0: x0 = 1/4, x1 = (2 - 1/4)/3
    A = oracle.CSR.from_arrays([0, 2, 4], [0, 1, 0, 1], [4.0, 1.0, 1.0, 3.0])
    x = np.zeros(2)
    A.gs_sweep([1.0, 2.0], x)
    assert x[0] == 0.25 and x[1] == (2.0 - 0.25) / 3.0
    for _ in range(60):
        A.gs_sweep([1.0, 2.0], x)
    assert np.allclose(x, np.linalg.solve([[4, 1], [1, 3]], [1, 2]), atol=1e-14)
    # ordered sweep: the identity order is the lexicographic sweep
    y = np.zeros(2)
    A.gs_sweep_order([1.0, 2.0], y, np.array([0, 1]))
    assert y[0] == 0.25
    y = np.zeros(2)
    A.gs_sweep_order([1.0, 2.0], y, np.array([1, 0]))
    assert y[1] == 2.0 / 3.0


@pytest.mark.parametrize("D", [3, 4])
def test_cg_solves_the_assembled_system(D):
    pts = unit_points(800, seed=D)
    g2i, i2o = oracle.tree(D)
    A = oracle.assemble(D, g2i, i2o, pts, 4.0)
    nf, _ = oracle.splat(D, g2i, pts)
    b = oracle.rhs(D, g2i, i2o, nf, pts, 4.0)
    xd = sla.spsolve(A.scipy().tocsc(), b)
    x = np.zeros(8 ** D)
    it, rel = A.cg(b, x, 10000, 1e-12)
    assert rel <= 1e-12
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    # the multicolour ordering converges to the same solution
    x2 = np.zeros(8 ** D)
    order = oracle.colour_order(D)
    for _ in range(3000):
        A.gs_sweep_order(b, x2, order)
    assert np.abs(x2 - xd).max() <= 1e-6 * np.abs(xd).max()


def test_cascadic_reaches_tolerance():
    D = 5
    pts = unit_points(3000, seed=7)
    g2i, i2o = oracle.tree(D)
    nf, _ = oracle.splat(D, g2i, pts)
    x, rep = oracle.solve_cascadic(D, g2i, i2o, nf, pts, 4.0, cg_iters=40000, eps=1e-10)
    A = oracle.assemble(D, g2i, i2o, pts, 4.0)
    b = oracle.rhs(D, g2i, i2o, nf, pts, 4.0)
    xd = sla.spsolve(A.scipy().tocsc(), b)
    assert rep["fine_rel"] <= 1e-10
    assert np.abs(x - xd).max() <= 1e-7 * (xd.max() - xd.min())
    # default reference parameters: fine CG capped at cg_iters/4 = 125 (Q13)
    _, rep2 = oracle.solve_cascadic(D, g2i, i2o, nf, pts, 4.0, cg_iters=500, eps=1e-12)
    assert rep2["fine_iters"] <= 125


# ----------------------------------------------------------------- evaluate (a17)
def test_evaluate_closed_forms():
    D = 4
    res = 1 << D
    h = 1.0 / res
    g2i, i2o = oracle.tree(D)
    rng = np.random.default_rng(0)
    sol = rng.normal(size=8 ** D)
    # at s = o + 1 (interior), F = x_o exactly (Q15: node o peaks at s = o+1)
    for o in [(3, 4, 5), (0, 0, 0), (14, 2, 7)]:
        s = np.array(o) + 1.0
        assert oracle.evaluate(D, g2i, sol, (s * h)[None])[0] == sol[g2i[o[2] * res * res + o[1] * res + o[0]]]
    # linear along x between s = 3 and s = 4 (nodes o = 2, 3)
    a, b = sol[g2i[5 * 256 + 5 * 16 + 2]], sol[g2i[5 * 256 + 5 * 16 + 3]]
    q = np.array([[3.25 * h, 6 * h, 6 * h]])
    assert oracle.evaluate(D, g2i, sol, q)[0] == pytest.approx(0.75 * a + 0.25 * b, rel=1e-14)
    # s < 0 -> 0; s in [0,1) double counts node 0; s >= res double counts node res-1
    assert oracle.evaluate(D, g2i, sol, np.array([[-0.01, 0.5, 0.5]]))[0] == 0.0
    x0 = sol[g2i[0]]
    assert oracle.evaluate(D, g2i, sol, np.array([[0.5 * h, 1.0 * h, 1.0 * h]]))[0] == pytest.approx(2 * 0.5 * x0)
    xl = sol[g2i[(res - 1) * res * res + (res - 1) * res + res - 1]]
    assert oracle.evaluate(D, g2i, sol, np.array([[1.0, 1.0, 1.0]]))[0] == pytest.approx(8 * xl)


def test_iso_is_mean_of_evaluate():
    D = 4
    pts = unit_points(500, seed=8)
    g2i, _ = oracle.tree(D)
    sol = np.random.default_rng(1).normal(size=8 ** D)
    q = np.stack([pts["px"], pts["py"], pts["pz"]], axis=1)
    assert oracle.iso(D, g2i, sol, pts) == pytest.approx(oracle.evaluate(D, g2i, sol, q).mean(), rel=1e-13)


# ----------------------------------------------------------------- geometry (a2)
def test_geometry_kats_and_unit_transform():
    # src/geometry.rs:438-446 bbox; unit transform maps the bbox centre to 0.5, max extent to `scale`
    pts = oracle.as_points([[-1.0, -2.0, -3.0], [1.0, 2.0, 3.0]], [[1.0, 0, 0], [0, 3.0, 4.0]])
    m, mi = oracle.unit_transform(pts, 1.1)
    up = oracle.apply_transform(pts, m)
    assert up["pz"][0] == pytest.approx(0.5 - 0.55) and up["pz"][1] == pytest.approx(0.5 + 0.55)
    assert up["px"][0] == pytest.approx(0.5 - 1.1 / 6)
    assert np.allclose([up["nx"][1], up["ny"][1], up["nz"][1]], [0.0, 0.6, 0.8])
    M = m.reshape(4, 4).T
    Mi = mi.reshape(4, 4).T
    assert np.allclose(M @ Mi, np.eye(4), atol=1e-14)
    # about 27% of sphere points fall outside [0,1)^3 at scale 1.1 (Q1)
    raw = synthetic.sphere(20000, seed=0)
    m, _ = oracle.unit_transform(raw, 1.1)
    up = oracle.apply_transform(raw, m)
    out = ~np.all([(up[c] >= 0) & (up[c] < 1) for c in ("px", "py", "pz")], axis=0)
    assert 0.2 < out.mean() < 0.35
