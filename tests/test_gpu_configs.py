"""The precision, batching and large workloads at their full sizes.

The oracle's CSR path cannot run at these sizes (about 140 GB of host memory at
depth 9), so these tests check what does not depend on size: the stopping
test on the FP64 true residual, operator symmetry and linearity, agreement
between the solvers and precisions, agreement of batched instances with the
single-tree path (which the smaller parity tests pin to the oracle), and the
corner field against the general evaluate() form.  The depth-10 and batch
tests run first and free their device memory before the depth-9 tree (a
module fixture shared by the last two tests) is built.

  precision  torus R=1, r=0.35, 1,000,000 points, depth 9 (134,217,728 nodes)
  batching   256 superquadrics, 2e3..2e5 points each, depth 6/7/8 by size
  large      sphere displaced by 8 spherical harmonics, 5,000,000 points,
             depth 10 (1,073,741,824 nodes), corners on the device
"""
import ctypes

import numpy as np
import pytest

import poissonreconstruction_amd as pr
from poissonreconstruction_amd import _lib as L
from poissonreconstruction_amd import synthetic
from poissonreconstruction_amd.batch import Batch, depth_for

pytestmark = pytest.mark.gpu

PW = 4.0


def _tree(D, raw):
    unit, _, _ = L.unit_transform(raw, 1.1)
    t = pr.FEMTree(D)
    t.splat_normal_field(unit)
    return t, unit


def _solve(t, D, prec, eps, precond=pr.PRX_PRECOND_NONE, cap=200000):
    return t.solve_cascadic(None, PW, D - 2, 20, 4 * cap, eps, precision=prec, residual_mode=pr.PRX_RESID_TRUE,
                            refresh_period=0, precond=precond)


def _field_err(t, q, x, x_ref):
    """max |F - F_ref| / range(F_ref) at the nodes and at the points q"""
    rng = float(x_ref.max() - x_ref.min())
    t.solution = x_ref
    f_ref = t.evaluate_many(q)
    t.solution = x
    f = t.evaluate_many(q)
    return float(np.abs(x - x_ref).max() / rng), float(np.abs(f - f_ref).max() / rng)


def test_large_workload_depth10(gpu):
    D = 10
    print("large workload: displaced sphere, 5,000,000 points, depth 10", flush=True)
    t, unit = _tree(D, synthetic.displaced_sphere(5_000_000, seed=0))
    rep = _solve(t, D, pr.PRX_PREC_MIXED, 1e-6, pr.PRX_PRECOND_MG)
    print(f"  multigrid mixed: fine_iters {rep['fine_iters']} true_rel {rep['true_rel']:.3e} {rep['ms_total']:.1f} ms",
          flush=True)
    assert rep["status"] == pr.PRX_OK and rep["true_rel"] <= 1e-6
    # every corner k/res on the device, against the general evaluate() at a sample of them (bit for bit)
    corners = t.evaluate_corners().reshape(-1)
    r1 = (1 << D) + 1
    assert corners.shape == (r1 ** 3,)
    rng = np.random.default_rng(10)
    idx = rng.integers(0, r1 ** 3, size=200000)
    idx = np.concatenate([idx, [0, r1 - 1, r1 ** 3 - 1]])
    kx, ky, kz = idx % r1, (idx // r1) % r1, idx // (r1 * r1)
    q = np.stack([kx, ky, kz], axis=1).astype(np.float64) / (1 << D)
    assert np.array_equal(t.evaluate_many(q).view(np.uint64), corners[idx].view(np.uint64))
    del corners
    # the device checksums of the two forms over all 1.08G corners agree too
    sums = []
    for general in (0, 1):
        ms, sq = ctypes.c_double(0), ctypes.c_double(0)
        L.check(L.lib().prx_time_evaluate_corners(t._h, 1, general, ctypes.byref(ms), ctypes.byref(sq)))
        sums.append(sq.value)
    assert sums[0] == sums[1] and sums[0] > 0
    t.close()


def test_batching_workload_256_superquadrics(gpu):
    sizes = synthetic.batch_sizes(256, seed=0)
    print(f"batching workload: 256 superquadrics, {int(sizes.sum())} points", flush=True)
    B = Batch()
    clouds = []
    for k, n in enumerate(sizes):
        unit, _, _ = L.unit_transform(synthetic.superquadric(int(n), seed=1000 + k), 1.1)
        B.add(depth_for(int(n)), unit, PW)
        clouds.append(unit)
    depths = [depth_for(int(n)) for n in sizes]
    assert set(depths) == {6, 7, 8}
    for prec in (pr.PRX_PREC_MIXED, pr.PRX_PREC_FP64):
        rep, per = B.solve(eps=1e-6, precond=pr.PRX_PRECOND_MG, precision=prec)
        print(f"  multigrid precision {prec}: converged {rep['converged']}/{rep['instances']} "
              f"max_true_rel {rep['max_true_rel']:.3e} {rep['ms_total']:.1f} ms", flush=True)
        assert rep["instances"] == 256 and rep["converged"] == 256 and rep["max_true_rel"] <= 1e-6
    mrep, mper = B.solve(eps=1e-6, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=True,
                         precision=pr.PRX_PREC_MIXED)
    print(f"  cascadic mixed batched: converged {mrep['converged']}/{mrep['instances']} max_true_rel "
          f"{mrep['max_true_rel']:.3e} {mrep['ms_total']:.1f} ms", flush=True)
    assert mrep["converged"] == 256 and mrep["max_true_rel"] <= 1e-6
    mixed_x = {d: B.tree(depths.index(d)).solution for d in (6, 7, 8)}
    rep, per = B.solve(eps=1e-6, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=True)
    print(f"  cascadic FP64 batched: converged {rep['converged']}/{rep['instances']} max_true_rel "
          f"{rep['max_true_rel']:.3e} {rep['ms_total']:.1f} ms, {rep['graph_launches']} graph launches", flush=True)
    assert rep["converged"] == 256 and rep["max_true_rel"] <= 1e-6
    # one instance of each depth against the same cloud solved alone
    for d in (6, 7, 8):
        k = depths.index(d)
        t = pr.FEMTree(d)
        t.splat_normal_field(clouds[k])
        rs = t.solve_cascadic(clouds[k], PW, d - 2, 20, 800000, 1e-6, precision=pr.PRX_PREC_FP64,
                              residual_mode=pr.PRX_RESID_TRUE, refresh_period=0)
        xs, xb = t.solution, B.tree(k).solution
        assert np.abs(xb - xs).max() <= 1e-9 * np.abs(xs).max(), (d, k)
        assert abs(per[k]["fine_iters"] - rs["fine_iters"]) <= 2
        # the mixed batched instance solved the same system to the same residual
        assert np.abs(mixed_x[d] - xs).max() <= 1e-4 * (xs.max() - xs.min()), (d, k)
        t.close()
    # the same solves as one persistent launch per level: every instance takes
    # the graph replays' FP64 iterations (same folds, same decisions); x
    # differs only in the rounding of its deferred updates
    graph_x = {d: B.tree(depths.index(d)).solution for d in (6, 7, 8)}
    prep, pper = B.solve(eps=1e-6, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0,
                         launch=pr.PRX_BATCH_LAUNCH_PERSISTENT)
    print(f"  cascadic FP64 persistent: converged {prep['converged']}/{prep['instances']} max_true_rel "
          f"{prep['max_true_rel']:.3e} {prep['ms_total']:.1f} ms, {prep['persistent_rounds']} rounds", flush=True)
    assert prep["converged"] == 256 and prep["max_true_rel"] <= 1e-6
    assert [p["fine_iters"] for p in pper] == [p["fine_iters"] for p in per]
    for d in (6, 7, 8):
        xp, xg = B.tree(depths.index(d)).solution, graph_x[d]
        assert np.abs(xp - xg).max() <= 1e-9 * np.abs(xg).max(), d
    mprep, _ = B.solve(eps=1e-6, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, precision=pr.PRX_PREC_MIXED,
                       launch=pr.PRX_BATCH_LAUNCH_PERSISTENT)
    print(f"  cascadic mixed persistent: converged {mprep['converged']}/{mprep['instances']} max_true_rel "
          f"{mprep['max_true_rel']:.3e} {mprep['ms_total']:.1f} ms", flush=True)
    assert mprep["converged"] == 256 and mprep["max_true_rel"] <= 1e-6
    for d in (6, 7, 8):
        xm, xg = B.tree(depths.index(d)).solution, graph_x[d]
        assert np.abs(xm - xg).max() <= 1e-4 * (xg.max() - xg.min()), d
    B.close()


@pytest.fixture(scope="module")
def torus9():
    D = 9
    print("precision workload: torus, 1,000,000 points, depth 9", flush=True)
    t, unit = _tree(D, synthetic.torus(1_000_000, seed=0))
    t.assemble_system(unit, PW)
    q = np.stack([unit["px"], unit["py"], unit["pz"]], axis=1)
    yield D, t, unit, q
    t.close()


def test_precision_workload_operator_symmetric_and_linear(gpu, torus9):
    D, t, unit, _ = torus9
    op = t.assemble_system(unit, PW)
    rng = np.random.default_rng(9)
    x, y = rng.normal(size=8 ** D), rng.normal(size=8 ** D)
    for prec, tol in ((pr.PRX_PREC_FP64, 1e-12), (pr.PRX_PREC_MIXED, 1e-6)):
        Ax, Ay = op.multiply_vector(x, prec), op.multiply_vector(y, prec)
        assert abs(x @ Ay - y @ Ax) <= tol * np.abs(x).dot(np.abs(Ay))
        Axy = op.multiply_vector(2.0 * x - 3.0 * y, prec)
        assert np.abs(Axy - (2.0 * Ax - 3.0 * Ay)).max() <= 10 * tol * np.abs(Axy).max()
    # the constant vector: the Laplacian's rows sum to zero away from the boundary, so A 1 is the screening
    # mass wg plus each touched node's screening row sum wh3 sum_p phi_pi (sum_k phi_pk = 1), which is >= 0
    # where no point has extrapolated (negative) weights, i.e. two nodes in from the faces
    one = op.multiply_vector(np.ones(8 ** D))
    h = 1.0 / (1 << D)
    wg = 0.005 * PW * h ** 3 / 1.5
    i2o = t.idx_to_offset
    inner = np.all((i2o >= 2) & (i2o <= (1 << D) - 3), axis=1)
    assert np.all(one[inner] >= wg * (1 - 1e-9))
    assert np.count_nonzero(one[inner] > wg * (1 + 1e-6)) > 0  # touched nodes carry screening


def test_precision_workload_solvers_reach_1e6(gpu, torus9):
    D, t, unit, q = torus9
    sols, reps = {}, {}
    for name, prec, precond in (("fp64", pr.PRX_PREC_FP64, pr.PRX_PRECOND_NONE),
                                ("mixed", pr.PRX_PREC_MIXED, pr.PRX_PRECOND_NONE),
                                ("mg_fp64", pr.PRX_PREC_FP64, pr.PRX_PRECOND_MG),
                                ("mg_mixed", pr.PRX_PREC_MIXED, pr.PRX_PRECOND_MG)):
        rep = _solve(t, D, prec, 1e-6, precond)
        print(f"  {name}: fine_iters {rep['fine_iters']} outer {rep['outer_iters']} true_rel {rep['true_rel']:.3e} "
              f"{rep['ms_total']:.1f} ms", flush=True)
        assert rep["status"] == pr.PRX_OK and rep["true_rel"] <= 1e-6, name
        sols[name], reps[name] = t.solution, rep
    # reference field: FP64 multigrid-preconditioned CG to 1e-11
    ref = _solve(t, D, pr.PRX_PREC_FP64, 1e-11, pr.PRX_PRECOND_MG)
    assert ref["true_rel"] <= 1e-11
    x_ref = t.solution
    errs = {k: _field_err(t, q, x, x_ref) for k, x in sols.items()}
    print(f"  field error (nodes, points) vs FP64 1e-11: {errs}", flush=True)
    # multigrid-preconditioned solves meet 1e-5 of the field range at the points
    assert errs["mg_mixed"][1] <= 1e-5 and errs["mg_fp64"][1] <= 1e-5
    # the unpreconditioned cascadic CG stopped at the same residual leaves a
    # smooth error of about 1e-5 in both precisions; the two agree
    assert errs["fp64"][1] <= 3e-5 and errs["mixed"][1] <= 3e-5
    assert _field_err(t, q, sols["mixed"], sols["fp64"])[1] <= 3e-5
    # the mixed inner solve performs the FP64 iteration: iteration counts within 5%
    assert abs(reps["mixed"]["fine_iters"] - reps["fp64"]["fine_iters"]) <= 0.05 * reps["fp64"]["fine_iters"]
