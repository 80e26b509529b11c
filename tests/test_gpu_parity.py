"""GPU parity of the solve path against the FP64 oracle, through the C ABI.

Integer/index work and the sequential-sum stages (node set, splat, RHS, CSR
pattern and values, coarse RHS, prolongation, evaluate) must match the oracle
bit for bit.  Operator applications must match to 1e-13 relative (FP64) and
2e-6 (FP32); converged solves must give the same field within 1e-5 of the
field range.
"""
import numpy as np
import pytest

import oracle
import poissonreconstruction_amd as pr
from poissonreconstruction_amd import _lib as L
from poissonreconstruction_amd import synthetic

pytestmark = pytest.mark.gpu


def unit_points(kind="sphere", n=2000, seed=0, scale=1.1):
    raw = synthetic.sphere(n, seed) if kind == "sphere" else synthetic.torus(n, seed=seed)
    m, _ = oracle.unit_transform(raw, scale)
    return oracle.apply_transform(raw, m)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def make_tree(D, pts, pw):
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    t.assemble_system(pts, pw)
    return t


@pytest.mark.parametrize("D", [0, 1, 2, 3, 4, 5])
def test_node_set_and_grid_bit_exact(gpu, D):
    g2i, i2o = oracle.tree(D)
    t = pr.FEMTree(D)
    assert t.fem_node_count == 8 ** D
    assert np.array_equal(t.idx_to_offset, i2o)
    assert np.array_equal(t.grid_to_idx, g2i)


@pytest.mark.parametrize("D,n", [(1, 50), (3, 500), (5, 4000), (6, 20000)])
def test_splat_bit_exact(gpu, D, n):
    pts = unit_points("sphere", n, seed=D)
    g2i, _ = oracle.tree(D)
    nf, w = oracle.splat(D, g2i, pts)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    assert np.array_equal(bits(t.node_weight), bits(w))
    assert np.array_equal(bits(t.normal_field), bits(nf))


@pytest.mark.parametrize("D,pw", [(2, 0.0), (2, 4.0), (4, 4.0), (5, 0.0), (6, 4.0)])
def test_rhs_bit_exact(gpu, D, pw):
    pts = unit_points("torus", 3000, seed=D)
    g2i, i2o = oracle.tree(D)
    nf, _ = oracle.splat(D, g2i, pts)
    b = oracle.rhs(D, g2i, i2o, nf, pts, pw)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    assert np.array_equal(bits(t.assemble_rhs(pts, pw)), bits(b))


def corner_cloud(n, seed, res):
    """A small sphere packed into one corner of the cube, plus points whose
    grid coordinates sit exactly on 8^3 brick faces: most bricks then hold no
    splat node, and the bricks that do border ones that do not."""
    raw = synthetic.sphere(n, seed)
    base = oracle.apply_transform(raw, oracle.unit_transform(raw, 1.1)[0])
    pts = np.zeros(n + 64, dtype=base.dtype)
    pts[:n] = base
    for c in ("px", "py", "pz"):
        pts[c][:n] = 0.05 + 0.2 * base[c]
    rng = np.random.default_rng(seed)
    k = rng.integers(1, res // 8, size=(64, 3)) * 8
    for j, c in enumerate(("px", "py", "pz")):
        pts[c][n:] = k[:, j] / res
    for c in ("nx", "ny", "nz"):
        pts[c][n:] = rng.normal(size=64)
    nrm = np.sqrt(pts["nx"][n:] ** 2 + pts["ny"][n:] ** 2 + pts["nz"][n:] ** 2)
    for c in ("nx", "ny", "nz"):
        pts[c][n:] /= nrm
    return pts


@pytest.mark.parametrize("D,pw", [(7, 4.0), (7, 0.0)])
def test_rhs_bit_exact_where_most_bricks_hold_no_splat_node(gpu, D, pw):
    # the RHS pass writes +0.0 without reading the field on bricks with no
    # splat node among their 27 neighbour bricks; the rest is the same sum
    pts = corner_cloud(4000, D, 1 << D)
    g2i, i2o = oracle.tree(D)
    nf, _ = oracle.splat(D, g2i, pts)
    b = oracle.rhs(D, g2i, i2o, nf, pts, pw)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    assert np.array_equal(bits(t.assemble_rhs(pts, pw)), bits(b))
    assert (b.reshape(-1, 512) == 0).all(axis=1).mean() > 0.5  # the skip path is exercised


@pytest.mark.parametrize("D,cd", [(6, 3), (7, 3), (7, 5)])
def test_coarse_rhs_bit_exact_over_the_splat_list(gpu, D, cd):
    # the coarse RHS sums only the splat list's nodes under each coarse cell:
    # 8^(D-cd) fine nodes per cell, below, at and above one 8^3 brick
    pts = corner_cloud(3000, D + cd, 1 << D)
    g2i, i2o = oracle.tree(D)
    nf, _ = oracle.splat(D, g2i, pts)
    _, cb = oracle.coarse_system(D, cd, i2o, nf, 4.0)
    t = make_tree(D, pts, 4.0)
    rhs, _ = t.build_system_at_depth(cd)
    assert np.array_equal(bits(rhs), bits(cb))


@pytest.mark.parametrize("D,n,pw", [(1, 20, 4.0), (2, 100, 4.0), (3, 400, 4.0), (4, 1500, 4.0), (3, 400, 0.0)])
def test_csr_pattern_and_values_bit_exact(gpu, D, n, pw):
    pts = unit_points("sphere", n, seed=D + 10)
    g2i, i2o = oracle.tree(D)
    A = oracle.assemble(D, g2i, i2o, pts, pw)
    off, col, val = A.arrays()
    t = make_tree(D, pts, pw)
    goff, gcol, gval = t.assemble_system(pts, pw).to_csr()
    assert np.array_equal(goff, off)
    assert np.array_equal(gcol, col)
    assert np.array_equal(bits(gval), bits(val))


def test_csr_explicit_zero_entries(gpu):
    # a point exactly on a node: 7 of its 8 weights are 0, but the reference
    # still stores the 64 pair entries, including face neighbours (value 0.0)
    D = 3
    h = 1.0 / 8
    pts = oracle.as_points([[3 * h, 4 * h, 5 * h]], [[0.0, 0.0, 1.0]])
    g2i, i2o = oracle.tree(D)
    A = oracle.assemble(D, g2i, i2o, pts, 4.0)
    off, col, val = A.arrays()
    t = make_tree(D, pts, 4.0)
    goff, gcol, gval = t.assemble_system(pts, 4.0).to_csr()
    assert np.array_equal(goff, off) and np.array_equal(gcol, col) and np.array_equal(bits(gval), bits(val))
    i = g2i[5 * 64 + 4 * 8 + 3]
    face = g2i[5 * 64 + 4 * 8 + 4]
    row = dict(zip(col[off[i]:off[i + 1]], val[off[i]:off[i + 1]]))
    assert face in row and row[face] == 0.0


@pytest.mark.parametrize("D,pw", [(2, 4.0), (3, 4.0), (4, 4.0), (5, 4.0), (5, 0.0), (6, 4.0)])
def test_apply_fp64_and_fp32(gpu, D, pw):
    pts = unit_points("torus", 5000, seed=3)
    g2i, i2o = oracle.tree(D)
    A = oracle.assemble(D, g2i, i2o, pts, pw)
    t = make_tree(D, pts, pw)
    op = t.assemble_system(pts, pw)
    rng = np.random.default_rng(D)
    for x in (rng.normal(size=8 ** D), np.ones(8 ** D)):
        y = A.spmv(x)
        y64 = op.multiply_vector(x, pr.PRX_PREC_FP64)
        scale = np.abs(A.scipy()).dot(np.abs(x)).max()
        assert np.abs(y64 - y).max() <= 1e-13 * scale
        y32 = op.multiply_vector(x, pr.PRX_PREC_MIXED)
        assert np.abs(y32 - y).max() <= 2e-6 * scale


def test_apply_symmetric_and_linear_large(gpu):
    # size-independent properties at depth 8 (16.7M nodes): x.Ay == y.Ax, linearity
    D = 8
    pts = unit_points("torus", 200000, seed=5)
    t = make_tree(D, pts, 4.0)
    op = t.assemble_system(pts, 4.0)
    rng = np.random.default_rng(0)
    x, y = rng.normal(size=8 ** D), rng.normal(size=8 ** D)
    Ax, Ay = op.multiply_vector(x), op.multiply_vector(y)
    assert abs(x @ Ay - y @ Ax) <= 1e-12 * np.abs(x).dot(np.abs(Ay))
    Axy = op.multiply_vector(2.0 * x - 3.0 * y)
    assert np.abs(Axy - (2.0 * Ax - 3.0 * Ay)).max() <= 1e-12 * np.abs(Axy).max()


@pytest.mark.parametrize("D,cd,pw", [(4, 2, 4.0), (5, 3, 4.0), (6, 4, 0.0), (3, 1, 4.0)])
def test_coarse_system_and_prolongation_bit_exact(gpu, D, cd, pw):
    pts = unit_points("sphere", 3000, seed=1)
    g2i, i2o = oracle.tree(D)
    nf, _ = oracle.splat(D, g2i, pts)
    C, cb = oracle.coarse_system(D, cd, i2o, nf, pw)
    t = make_tree(D, pts, pw)
    rhs, coef = t.build_system_at_depth(cd)
    assert np.array_equal(bits(rhs), bits(cb))
    # the coarse CSR is the stencil of coef: rebuild its values and compare bits
    off, col, val = C.arrays()
    r = 1 << cd
    for i in range(C.n):
        z, y, x = i // (r * r), (i // r) % r, i % r
        for k in range(off[i], off[i + 1]):
            j = col[k]
            jz, jy, jx = j // (r * r), (j // r) % r, j % r
            nzc = int(jx != x) + int(jy != y) + int(jz != z)
            expect = {0: coef[0], 2: coef[1], 3: coef[2]}[nzc]
            if nzc == 0 and coef[3] > 0:
                expect = 0.0 + coef[0] + coef[3]
            assert val[k] == expect
    rng = np.random.default_rng(cd)
    c = rng.normal(size=8 ** cd)
    c[rng.uniform(size=c.shape) < 0.2] = 0.0
    assert np.array_equal(bits(t.prolongate(c, cd)), bits(oracle.prolongate(D, cd, g2i, c)))


@pytest.mark.parametrize("D", [3, 4, 5])
def test_multicolour_gs_equals_colour_ordered_sweep(gpu, D):
    pts = unit_points("torus", 3000, seed=2)
    g2i, i2o = oracle.tree(D)
    A = oracle.assemble(D, g2i, i2o, pts, 4.0)
    nf, _ = oracle.splat(D, g2i, pts)
    b = oracle.rhs(D, g2i, i2o, nf, pts, 4.0)
    t = make_tree(D, pts, 4.0)
    rng = np.random.default_rng(1)
    x0 = rng.normal(size=8 ** D) * 1e-3
    order = oracle.colour_order(D)
    xr = x0.copy()
    for _ in range(3):
        A.gs_sweep_order(b, xr, order)
    xg = t.gauss_seidel(b, x0, sweeps=3)
    assert np.abs(xg - xr).max() <= 1e-12 * np.abs(xr).max()
    # the FP32 sweep (the mixed solve's smoother) to FP32 accuracy
    xf = t.gauss_seidel(b, x0, sweeps=3, precision=pr.PRX_PREC_MIXED)
    assert np.abs(xf - xr).max() <= 1e-4 * np.abs(xr).max()


@pytest.mark.parametrize("D", [2, 3, 4])
def test_cg_matches_direct_solve(gpu, D):
    # depth 2: two-pass CG on the per-node kernel; 3 and 4: the fused CG step on 8^3 and 16^3 tiles
    pts = unit_points("sphere", 2000, seed=4)
    g2i, i2o = oracle.tree(D)
    A = oracle.assemble(D, g2i, i2o, pts, 4.0)
    nf, _ = oracle.splat(D, g2i, pts)
    b = oracle.rhs(D, g2i, i2o, nf, pts, 4.0)
    import scipy.sparse.linalg as sla
    xd = sla.spsolve(A.scipy().tocsc(), b)
    t = make_tree(D, pts, 4.0)
    x, it, rel = t.solve_cg(b, np.zeros(8 ** D), 5000, 1e-12)
    xo = np.zeros(8 ** D)
    ito, relo = A.cg(b, xo, 5000, 1e-12)
    assert np.abs(x - xd).max() <= 1e-8 * np.abs(xd).max()
    assert abs(int(it) - ito) <= max(3, ito // 20)


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
def test_cascadic_c1_field_matches_oracle(gpu, prec):
    # Plain-method workload: unit sphere, 20,000 points, depth 6, pw 4, CG to 1e-6
    D, P = 6, 20000
    pts = unit_points("sphere", P, seed=0)
    g2i, i2o = oracle.tree(D)
    nf, _ = oracle.splat(D, g2i, pts)
    xo, rep_o = oracle.solve_cascadic(D, g2i, i2o, nf, pts, 4.0, cg_iters=40000, eps=1e-10)
    t = make_tree(D, pts, 4.0)
    rep = t.solve_cascadic(pts, 4.0, D - 2, 20, 40000, 1e-10, precision=prec, residual_mode=pr.PRX_RESID_TRUE,
                           refresh_period=10 if prec == pr.PRX_PREC_FP64 else 0)
    assert rep["status"] == 0
    assert rep["true_rel"] <= 1e-9
    x = t.solution
    q = np.stack([pts["px"], pts["py"], pts["pz"]], axis=1)
    fo = oracle.evaluate(D, g2i, xo, q)
    fg = t.evaluate_many(q)
    corners_o = xo  # field at corners is the solution itself (shifted), compare over the nodes
    rng_f = corners_o.max() - corners_o.min()
    assert np.abs(fg - fo).max() <= 1e-5 * rng_f
    assert np.abs(x - xo).max() <= 1e-5 * rng_f


@pytest.mark.parametrize("prec,mode,refresh,pw", [(pr.PRX_PREC_FP64, pr.PRX_RESID_TRUE, 10, 4.0),
                                                  (pr.PRX_PREC_FP64, pr.PRX_RESID_REFERENCE, 10, 4.0),
                                                  (pr.PRX_PREC_FP64, pr.PRX_RESID_TRUE, 0, 4.0),
                                                  (pr.PRX_PREC_MIXED, pr.PRX_RESID_TRUE, 0, 4.0),
                                                  (pr.PRX_PREC_FP64, pr.PRX_RESID_TRUE, 0, 0.0),
                                                  (pr.PRX_PREC_MIXED, pr.PRX_RESID_TRUE, 0, 0.0)])
def test_fused_cg_step_equals_two_pass(gpu, prec, mode, refresh, pw):
    # The fused CG iteration (the update r -= alpha q joined to the next operator pass, |r|^2 expanded from d.q, r.q
    # and q.q, r and q in ping-pong pairs, x in pairs) is the two-pass iteration of src/solvers.rs:36-64 with other
    # rounding: iteration counts within 2 % (10 % for the mixed solve, whose FP64 refinement points follow the FP32
    # residual estimate) and solutions within the solve's accuracy.  refresh 10 exercises the periodic r = b - A x of
    # the reference, the mixed solve the residual replacements.
    D, P = 6, 20000
    pts = unit_points("torus", P, seed=3)
    sols, reps = [], []
    for two_pass in (True, False):
        t = make_tree(D, pts, pw)
        rep = t.solve_cascadic(pts, pw, D - 2, 20, 40000, 1e-8, precision=prec, residual_mode=mode,
                               refresh_period=refresh, cg_two_pass=two_pass)
        assert rep["status"] == 0
        sols.append(t.solution)
        reps.append(rep)
    slack = 10 if prec == pr.PRX_PREC_MIXED else 50
    assert abs(int(reps[0]["fine_iters"]) - int(reps[1]["fine_iters"])) <= max(3, int(reps[0]["fine_iters"]) // slack)
    assert abs(int(reps[0]["coarse_iters"]) - int(reps[1]["coarse_iters"])) <= max(3, int(reps[0]["coarse_iters"]) // 50)
    if mode == pr.PRX_RESID_TRUE:
        assert reps[1]["true_rel"] <= 1e-8
    rng_f = sols[0].max() - sols[0].min()
    assert np.abs(sols[0] - sols[1]).max() <= 1e-5 * rng_f


@pytest.mark.parametrize("D,pw,shape,prec", [(6, 4.0, "torus", pr.PRX_PREC_FP64), (7, 4.0, "sphere", pr.PRX_PREC_FP64),
                                             (6, 4.0, "torus", pr.PRX_PREC_MIXED), (7, 4.0, "sphere", pr.PRX_PREC_MIXED),
                                             (6, 0.0, "torus", pr.PRX_PREC_MIXED)])
def test_fused_cg_face_form_equals_stored_q(gpu, D, pw, shape, prec):
    # The FP32 fused step in face form (q = A d recomputed inside each 16^3 brick, read on the brick's boundary from
    # the neighbours' faces, screening values formed from the brick's tile) against the same step with q stored as a
    # vector and the screening values from the per-point gather: the same iteration, other rounding of the screening
    # term only.  Inner-iteration counts within 2 %, the same FP64 true residual, solutions within the solve accuracy.
    pts = unit_points(shape, 20000, seed=11)
    sols, reps = [], []
    for form in (pr.PRX_CG_FUSED_STORED_Q, pr.PRX_CG_FUSED_FACES):
        t = make_tree(D, pts, pw)
        rep = t.solve_cascadic(pts, pw, D - 2, 20, 40000, 1e-8, precision=prec,
                               residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, cg_form=form)
        assert rep["status"] == 0 and rep["true_rel"] <= 1e-8, (form, rep)
        sols.append(t.solution)
        reps.append(rep)
    print(f"stored-q {reps[0]['fine_iters']} iters {reps[0]['outer_iters']} outer, face form {reps[1]['fine_iters']} "
          f"iters {reps[1]['outer_iters']} outer", flush=True)
    assert abs(int(reps[0]["fine_iters"]) - int(reps[1]["fine_iters"])) <= max(3, int(reps[0]["fine_iters"]) // 50)
    rng_f = sols[0].max() - sols[0].min()
    assert np.abs(sols[0] - sols[1]).max() <= 1e-5 * rng_f


@pytest.mark.parametrize("cap,refresh", [(37, 0), (37, 10), (45, 10), (8, 0), (40, 10)])
def test_fused_cg_stops_at_the_cap_with_x_terms_pending(gpu, cap, refresh):
    # Stopped by the iteration cap (the reference's cg_iters / 4 on the fine level, src/fem_tree.rs:458-487) after
    # `cap` fused steps, up to seven x terms are still pending and must be added exactly once: the solution then equals
    # the two-pass form's after the same number of iterations, to round-off.
    D, P = 5, 8000
    pts = unit_points("sphere", P, seed=8)
    sols, reps = [], []
    for two_pass in (True, False):
        t = make_tree(D, pts, 4.0)
        rep = t.solve_cascadic(pts, 4.0, D - 2, 20, 4 * cap, 1e-14, precision=pr.PRX_PREC_FP64,
                               residual_mode=pr.PRX_RESID_REFERENCE, refresh_period=refresh, cg_two_pass=two_pass)
        assert rep["fine_iters"] == cap
        sols.append(t.solution)
        reps.append(rep)
    assert reps[0]["coarse_iters"] == reps[1]["coarse_iters"]
    assert reps[0]["status"] == reps[1]["status"] == L.PRX_ERR_NOT_CONVERGED
    if refresh and cap % refresh == 0:
        # the cap falls on a refresh iteration: both forms report the refreshed residual
        assert reps[1]["fine_rel"] == pytest.approx(reps[0]["fine_rel"], rel=1e-6)
    scale = np.abs(sols[0]).max()
    assert np.abs(sols[0] - sols[1]).max() <= 1e-9 * scale


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
def test_fused_cg_step_on_8cubed_tiles(gpu, prec, monkeypatch):
    # The tile edge is chosen per tree (PRX_TILE=8 at creation): the 8^3-brick form of the fused CG step, the GS pass
    # and the folds solve the same system as the 16^3 form (depth 6: 512 bricks of 8^3 against 64 of 16^3).
    D, P = 6, 20000
    pts = unit_points("torus", P, seed=5)
    sols, reps = [], []
    for tile in ("16", "8"):
        monkeypatch.setenv("PRX_TILE", tile)
        t = make_tree(D, pts, 4.0)
        rep = t.solve_cascadic(pts, 4.0, D - 2, 20, 40000, 1e-8, precision=prec, residual_mode=pr.PRX_RESID_TRUE,
                               refresh_period=0)
        assert rep["status"] == 0 and rep["true_rel"] <= 1e-8
        sols.append(t.solution)
        reps.append(rep)
    slack = 10 if prec == pr.PRX_PREC_MIXED else 50
    assert abs(int(reps[0]["fine_iters"]) - int(reps[1]["fine_iters"])) <= max(3, int(reps[0]["fine_iters"]) // slack)
    rng_f = sols[0].max() - sols[0].min()
    assert np.abs(sols[0] - sols[1]).max() <= 1e-5 * rng_f


def test_evaluate_bit_exact(gpu):
    D = 5
    pts = unit_points("sphere", 3000, seed=6)
    g2i, _ = oracle.tree(D)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    t.assemble_system(pts, 0.0)
    rng = np.random.default_rng(0)
    sol = rng.normal(size=8 ** D)
    t.solution = sol
    q = np.concatenate([np.stack([pts["px"], pts["py"], pts["pz"]], axis=1), rng.uniform(-0.2, 1.2, size=(500, 3)),
                        np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.5 / 32, 1.0 / 32, 31.5 / 32]])])
    assert np.array_equal(bits(t.evaluate_many(q)), bits(oracle.evaluate(D, g2i, sol, q)))
    r1 = 33
    k = np.arange(r1) / 32.0
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    qc = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)
    assert np.array_equal(bits(t.evaluate_corners().ravel()), bits(oracle.evaluate(D, g2i, sol, qc)))
    iso = t.iso_value()
    iso_o = oracle.iso(D, g2i, sol, pts)
    assert abs(iso - iso_o) <= 1e-14 * max(1.0, abs(iso_o))


def test_reference_kats_through_abi(gpu):
    # src/sparse.rs:174-196
    m = pr.SparseMatrix(3)
    for i, s in enumerate((2, 1, 1)):
        m.set_row_size(i, s)
    m.finalize_structure()
    m.set_row(0, [(0, 2.0), (1, 1.0)])
    m.set_row(1, [(1, 3.0)])
    m.set_row(2, [(2, 4.0)])
    assert np.allclose(m.multiply_vector([1.0, 2.0, 3.0]), [4.0, 6.0, 12.0], atol=1e-10)
    # src/solvers.rs:234-268
    n = 10
    a = pr.SparseMatrix(n)
    for i in range(n):
        a.set_row_size(i, 3 if 0 < i < n - 1 else 2)
    a.finalize_structure()
    for i in range(n):
        row = ([(i - 1, -1.0)] if i > 0 else []) + [(i, 2.0)] + ([(i + 1, -1.0)] if i < n - 1 else [])
        a.set_row(i, row)
    x = np.zeros(n)
    pr.solve_cg(a, np.ones(n), x, 100, 1e-10)
    expect = np.array([(i + 1) * (n - i) / 2 for i in range(n)])
    assert np.abs(x - expect).max() < 1e-6


def test_reconstruct_end_to_end(gpu):
    raw = synthetic.sphere(20000, seed=0)
    params = pr.ReconstructParams(depth=6, point_weight=4.0, cg_accuracy=1e-10, cg_iters=40000)
    mesh, st = pr.reconstruct(raw, params)
    up, xo, iso_o, rep_o, _ = oracle.reconstruct_field(raw, depth=6, point_weight=4.0, cg_accuracy=1e-10, cg_iters=40000)
    assert st.fem_nodes == 8 ** 6
    assert abs(st.iso_value - iso_o) <= 1e-5 * (xo.max() - xo.min())
    assert st.matrix_nnz == rep_o["nnz"]
    mesh0, st0 = pr.reconstruct(np.zeros((0, 6)), params)
    assert st0.fem_nodes == 0 and len(mesh0.vertices) == 0


def test_screening_sizes_and_occupied_cells(gpu):
    # prx_screening_sizes against numpy: the touched nodes are the in-grid corners of the points' trunc cells
    # (src/fem_tree.rs:207-213), the entries one per (point, in-grid corner); the face-form screening groups the
    # points, in their spatial order (stable, by the Morton code of the trunc cell clamped to the grid), into runs of
    # one trunc cell.  Points of the unit-cube transform at scale 1.1 reach 0.05 beyond the cube, so cells outside the
    # grid share a clamped code and one such cell can form several runs; every in-grid cell forms exactly one.
    import ctypes
    D = 6
    pts = unit_points("torus", 20000, seed=13)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    out = (ctypes.c_int64 * 3)()
    L.check(L.lib().prx_screening_sizes(t._h, out))
    res = 1 << D
    c = np.trunc(np.stack([pts["px"], pts["py"], pts["pz"]], axis=1) * res).astype(np.int64)
    corners = np.concatenate([c + np.array([k & 1, (k >> 1) & 1, k >> 2]) for k in range(8)])
    inside = np.all((corners >= 0) & (corners < res), axis=1)
    assert out[0] == len(np.unique(corners[inside], axis=0))
    assert out[1] == int(inside.sum())

    def spread(v):
        r = np.zeros_like(v)
        for b in range(10):
            r = r | (((v >> b) & 1) << (3 * b))
        return r
    cl = np.clip(c, 0, res - 1)
    cs = c[np.argsort(spread(cl[:, 0]) | (spread(cl[:, 1]) << 1) | (spread(cl[:, 2]) << 2), kind="stable")]
    runs = 1 + int(np.count_nonzero(np.any(cs[1:] != cs[:-1], axis=1)))
    in_grid = np.all((c >= 0) & (c < res), axis=1)
    assert out[2] == runs
    assert runs - len(np.unique(c, axis=0)) <= len(np.unique(c[~in_grid], axis=0)) * 8


def test_release_workspace_then_solve_again(gpu):
    # freeing the solver workspace between solves changes nothing: the next solve allocates again and gives the
    # same solution bit for bit (the face-form arrays and cell lists included, depth 8)
    D = 8
    pts = unit_points("sphere", 30000, seed=14)
    t = make_tree(D, pts, 4.0)
    sols = []
    for release in (False, True):
        if release:
            t.release_workspace()
        rep = t.solve_cascadic(pts, 4.0, D - 2, 20, 800000, 1e-8, precision=pr.PRX_PREC_MIXED,
                               residual_mode=pr.PRX_RESID_TRUE, refresh_period=0)
        assert rep["status"] == 0
        sols.append(t.solution)
    assert np.array_equal(sols[0].view(np.uint64), sols[1].view(np.uint64))


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
def test_pass_timing_counts_every_fine_pass(gpu, prec):
    # prx_pass_timing records one duration per launch of the finest level's fused CG pass (the face form in the
    # FP32 inner solves at depth 8): every iteration's pass, plus those enqueued after an inner solve converged (the
    # rest of its chunk of 8 and the chunk queued behind it), which return at once; nothing is recorded once timing
    # is off
    import ctypes
    D = 8
    pts = unit_points("sphere", 30000, seed=15)
    t = make_tree(D, pts, 4.0)

    def timed_solve(on):
        if on:
            L.check(L.lib().prx_pass_timing(t._h, 1, None, 0, None))
        rep = t.solve_cascadic(pts, 4.0, D - 2, 20, 800000, 1e-6, precision=prec,
                               residual_mode=pr.PRX_RESID_TRUE, refresh_period=0)
        assert rep["status"] == 0
        n = ctypes.c_size_t(0)
        buf = np.zeros(1 << 14, dtype=np.float32)
        if on:
            L.check(L.lib().prx_pass_timing(t._h, 0, buf.ctypes.data, buf.size, ctypes.byref(n)))
        return rep, buf[:n.value]

    rep, ms = timed_solve(True)
    # the fine CG's starting residual after smoothing, relative to b (far above 1: the prolongated coarse solution
    # covers only fine offsets <= res/2, src/fem_tree.rs:310-343), and the reference's ||r||/||r0|| far below 1
    assert rep["fine_r0_rel"] > 1.0 and rep["true_rel"] < 1e-3 * rep["fine_r0_rel"]
    worked = ms[ms >= 0.1 * np.median(ms)]
    assert len(worked) == rep["fine_iters"], (len(ms), len(worked), rep["fine_iters"])
    # (at most 15 per CG run; an FP64 run that restarts after its true-residual check counts twice)
    assert len(ms) - len(worked) <= 32 * max(1, rep["outer_iters"])
    assert np.all(ms > 0) and np.all(np.isfinite(ms))
    # timing off: the next solve records nothing and the counter reads zero
    L.check(L.lib().prx_pass_timing(t._h, 1, None, 0, None))
    n = ctypes.c_size_t(7)
    L.check(L.lib().prx_pass_timing(t._h, 0, None, 0, ctypes.byref(n)))
    assert n.value == 0
    rep2, ms2 = timed_solve(False)
    assert len(ms2) == 0 and rep2["fine_iters"] == rep["fine_iters"]
    t.close()
