"""Multigrid-preconditioned CG (PRX_PRECOND_MG, an extension with no
counterpart in the reference) through the C ABI.

The V-cycle must be a symmetric positive-definite operator (checked on random
vectors), exact on the coarsest grid (where it is the dense inverse of the
stencil operator), and the preconditioned solve must reach the direct
(scipy) solution of the oracle's assembled system."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle
import poissonreconstruction_amd as pr
from poissonreconstruction_amd import synthetic

pytestmark = pytest.mark.gpu


def unit_points(kind, n, seed, scale=1.1):
    raw = synthetic.sphere(n, seed) if kind == "sphere" else synthetic.torus(n, seed=seed)
    m, _ = oracle.unit_transform(raw, scale)
    return oracle.apply_transform(raw, m)


def system(D, pts, pw):
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    return t, t.assemble_system(pts, pw)


def direct(D, pts, pw, csr=False):
    g2i, i2o = oracle.tree(D)
    C = oracle.assemble(D, g2i, i2o, pts, pw)
    nf, _ = oracle.splat(D, g2i, pts)
    b = oracle.rhs(D, g2i, i2o, nf, pts, pw)
    return (C.scipy(), b, C) if csr else (C.scipy(), b)


def test_vcycle_exact_on_coarsest_grid(gpu):
    # depth 2 is the coarsest grid itself; with pw = 0 the operator is the
    # stencil alone, so one V-cycle inverts it
    D = 2
    pts = unit_points("sphere", 300, 1)
    A, _ = direct(D, pts, 0.0)
    _, S = system(D, pts, 0.0)
    u = np.random.default_rng(0).standard_normal(8 ** D)
    z = S.mg_vcycle(A @ u)
    assert np.abs(z - u).max() <= 1e-12 * np.abs(u).max()


@pytest.mark.parametrize("prec,tol", [(pr.PRX_PREC_FP64, 1e-12), (pr.PRX_PREC_MIXED, 2e-5)])
@pytest.mark.parametrize("D", [3, 5])
def test_vcycle_symmetric_positive_definite(gpu, D, prec, tol):
    pts = unit_points("torus", 3000, D)
    _, S = system(D, pts, 4.0)
    rng = np.random.default_rng(D)
    u, v = rng.standard_normal((2, 8 ** D))
    Mu, Mv = S.mg_vcycle(u, prec), S.mg_vcycle(v, prec)
    assert abs(u @ Mv - v @ Mu) <= tol * np.linalg.norm(u) * np.linalg.norm(Mv)
    assert u @ Mu > 0 and v @ Mv > 0


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
@pytest.mark.parametrize("D", [4, 6])
def test_vcycle_reduces_smooth_error(gpu, D, prec):
    # without screening M approximates A^-1 itself: one V-cycle removes most
    # of the error of the smoothest mode (where plain smoothing does nothing)
    pts = unit_points("torus", 3000, D)
    _, S = system(D, pts, 0.0)
    A, _ = direct(D, pts, 0.0)
    g = np.linspace(0.0, np.pi, 2 ** D)
    sm = (np.sin(g)[None, None, :] * np.sin(g)[None, :, None] * np.sin(g)[:, None, None]).ravel()
    g2i, _ = oracle.tree(D)
    us = np.empty(8 ** D)
    us[g2i] = sm  # row-major grid order -> node order
    e = S.mg_vcycle(A @ us, prec) - us
    assert np.linalg.norm(e) <= 0.2 * np.linalg.norm(us)


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
@pytest.mark.parametrize("D,kind,n", [(4, "sphere", 2000), (6, "torus", 20000)])
def test_mg_solve_matches_direct(gpu, D, kind, n, prec):
    pts = unit_points(kind, n, D)
    A, b, C = direct(D, pts, 4.0, csr=True)
    if D <= 4:
        xd = sla.spsolve(A.tocsc(), b)
    else:  # the oracle's FP64 CG to 1e-13 (a direct factorisation is too slow here)
        xd = np.zeros(8 ** D)
        C.cg(b, xd, 100000, 1e-13)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    rep = t.solve_cascadic(pts, 4.0, D - 2, 20, 10000, 1e-10, precision=prec, residual_mode=pr.PRX_RESID_TRUE,
                           precond=pr.PRX_PRECOND_MG)
    assert rep["status"] == pr.PRX_OK
    assert rep["true_rel"] <= 1e-10
    assert rep["fine_iters"] <= (40 if prec == pr.PRX_PREC_FP64 else 80)
    x = t.solution
    assert np.abs(x - xd).max() <= 1e-7 * np.abs(xd).max()
    # the same residual the oracle's operator sees
    assert np.linalg.norm(b - A @ x) <= 2e-10 * np.linalg.norm(b)


def test_mg_iterations_stay_bounded_with_depth(gpu):
    # the point of the preconditioner: iterations do not grow with the grid
    iters = []
    for D in (5, 7):
        pts = unit_points("torus", 40000, 7)
        t = pr.FEMTree(D)
        t.splat_normal_field(pts)
        rep = t.solve_cascadic(pts, 4.0, D - 2, 20, 10000, 1e-6, precision=pr.PRX_PREC_FP64,
                               residual_mode=pr.PRX_RESID_TRUE, precond=pr.PRX_PRECOND_MG)
        assert rep["status"] == pr.PRX_OK and rep["true_rel"] <= 1e-6
        iters.append(rep["fine_iters"])
    assert iters[1] <= iters[0] + 6, iters


def test_reconstruct_with_mg(gpu):
    raw = synthetic.sphere(20000, 0)
    base = pr.reconstruct(raw, pr.ReconstructParams(depth=6, point_weight=4.0, cg_accuracy=1e-8, cg_iters=40000,
                                                    residual_mode=pr.PRX_RESID_TRUE))
    mg = pr.reconstruct(raw, pr.ReconstructParams(depth=6, point_weight=4.0, cg_accuracy=1e-8, cg_iters=40000,
                                                  residual_mode=pr.PRX_RESID_TRUE, precond=pr.PRX_PRECOND_MG))
    (mb, sb), (mm, sm) = base, mg
    assert sm.solve_status == pr.PRX_OK and sm.final_rel_residual <= 1e-8
    assert sm.fine_iters < sb.fine_iters / 10
    # same system to the same tolerance: the same surface
    assert abs(sm.iso_value - sb.iso_value) <= 1e-6 * abs(sb.iso_value)
    assert abs(len(mm.vertices) - len(mb.vertices)) <= 0.01 * len(mb.vertices)


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
def test_vjp_after_mg_solve_matches_oracle_gradient(gpu, prec):
    # the adjoint solve follows the forward solver: multigrid here
    D, pw = 5, 4.0
    raw = synthetic.ellipsoids(3000, seed=1)
    m, _ = oracle.unit_transform(raw, 1.1)
    pts = oracle.apply_transform(raw, m)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    rep = t.solve_cascadic(pts, pw, D - 2, 20, 10000, 1e-13, precision=pr.PRX_PREC_FP64,
                           residual_mode=pr.PRX_RESID_TRUE, precond=pr.PRX_PRECOND_MG)
    assert rep["status"] == pr.PRX_OK
    q = np.stack([pts["px"], pts["py"], pts["pz"]], axis=1)
    F = t.evaluate_many(q)
    gpos, gnrm, vrep = t.vjp_points(2.0 * (F - F.mean()), precision=prec, eps=1e-11)
    assert vrep["status"] == 0 and vrep["iters"] <= 60
    Fo, xo = oracle.field_at_points(D, pts, pw, eps=1e-13)
    gpo, gno, _ = oracle.vjp_points(D, pts, pw, xo, 2.0 * (Fo - Fo.mean()), eps=1e-13)
    assert np.abs(gpos - gpo).max() <= 1e-6 * np.abs(gpo).max()
    assert np.abs(gnrm - gno).max() <= 1e-6 * np.abs(gno).max()


def test_command_line_multigrid_flag(gpu, tmp_path):
    import os
    import subprocess
    import sys
    from poissonreconstruction_amd import ply
    raw = synthetic.sphere(3000, seed=2)
    src = tmp_path / "in.ply"
    with open(src, "w") as fh:
        fh.write(f"ply\nformat ascii 1.0\nelement vertex {len(raw)}\nproperty float x\nproperty float y\nproperty float z\n"
                 "property float nx\nproperty float ny\nproperty float nz\nend_header\n")
        for r in raw:
            fh.write(" ".join(repr(float(v)) for v in r) + "\n")
    out = tmp_path / "out.ply"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-m", "poissonreconstruction_amd", "-i", str(src), "-o", str(out), "-d", "5",
                        "--point-weight", "4", "--multigrid"], capture_output=True, text=True, cwd=root)
    assert r.returncode == 0, r.stderr
    mesh, st = pr.reconstruct(ply.read_oriented_points(str(src)),
                              pr.ReconstructParams(depth=5, point_weight=4.0, precond=pr.PRX_PRECOND_MG))
    assert st.solve_status == pr.PRX_OK
    lines = open(out).read().split("\n")
    assert (int(lines[2].split()[-1]), int(lines[6].split()[-1])) == (len(mesh.vertices), len(mesh.triangles))


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
@pytest.mark.parametrize("D,pw", [(1, 4.0), (2, 4.0), (3, 4.0), (3, 0.0), (5, 0.0)])
def test_mg_small_grids_and_no_screening(gpu, D, pw, prec):
    # depth <= 2 is the dense coarsest solve alone; depth 3 one smoothed level
    # on 8^3 bricks above it; pw = 0 leaves the operator without screening
    pts = unit_points("sphere", 1500, D)
    A, b = direct(D, pts, pw)
    xd = sla.spsolve(A.tocsc(), b)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    rep = t.solve_cascadic(pts, pw, max(D - 2, 0), 20, 10000, 1e-10, precision=prec, residual_mode=pr.PRX_RESID_TRUE,
                           precond=pr.PRX_PRECOND_MG)
    assert rep["status"] == pr.PRX_OK and rep["true_rel"] <= 1e-10
    assert np.abs(t.solution - xd).max() <= 1e-8 * max(np.abs(xd).max(), 1e-300)
