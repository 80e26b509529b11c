"""Batched solves (prx_solve_batched) against the same clouds solved one by
one with the single-tree cascadic solve (the path the other GPU tests pin to
the oracle).  Each instance's batched CG performs the single tree's
arithmetic, so solutions agree to round-off and iteration counts match."""
import numpy as np
import pytest

import oracle
import poissonreconstruction_amd as pr
from poissonreconstruction_amd import synthetic
from poissonreconstruction_amd.batch import Batch, depth_for

pytestmark = pytest.mark.gpu

CLOUDS = [(3000, 4.0, 11), (8000, 4.0, 12), (24000, 4.0, 13), (5000, 0.0, 14), (12000, 2.0, 15)]


def unit(raw):
    m, _ = oracle.unit_transform(raw, 1.1)
    return oracle.apply_transform(raw, m)


def clouds():
    return [(depth_for(n), unit(synthetic.superquadric(n, seed=sd)), pw) for n, pw, sd in CLOUDS]


def single(D, pts, pw, eps, mode, refresh, two_pass=False):
    # the single-tree solve with the same CG form is the reference for round-off identity (the fused and two-pass
    # forms are the same iteration with other rounding)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    rep = t.solve_cascadic(pts, pw, D - 2, 20, 800000, eps, precision=pr.PRX_PREC_FP64, residual_mode=mode,
                           refresh_period=refresh, cg_two_pass=two_pass)
    return t.solution, rep, t


@pytest.mark.parametrize("two_pass", [False, True])
@pytest.mark.parametrize("mode,eps,refresh", [(pr.PRX_RESID_TRUE, 1e-10, 0), (pr.PRX_RESID_REFERENCE, 1e-6, 10)])
def test_batch_matches_single_tree_solves(gpu, mode, eps, refresh, two_pass):
    cs = clouds()
    B = Batch()
    for D, pts, pw in cs:
        B.add(D, pts, pw)
    rep, per = B.solve(gs=20, cg_iters=800000, eps=eps, residual_mode=mode, refresh_period=refresh, autotune=False,
                       cg_two_pass=two_pass)
    assert rep["instances"] == len(cs) and rep["converged"] == len(cs)
    for k, (D, pts, pw) in enumerate(cs):
        xs, rs, _ = single(D, pts, pw, eps, mode, refresh, two_pass)
        xb = B.tree(k).solution
        scale = np.abs(xs).max()
        assert np.abs(xb - xs).max() <= 1e-9 * scale, k
        assert abs(per[k]["fine_iters"] - rs["fine_iters"]) <= 2, (k, per[k]["fine_iters"], rs["fine_iters"])
        assert abs(per[k]["coarse_iters"] - rs["coarse_iters"]) <= 2
        if mode == pr.PRX_RESID_TRUE:
            assert per[k]["true_rel"] <= eps * 1.0001


def test_batch_graph_replays_and_autotune(gpu):
    cs = clouds()
    B = Batch()
    for D, pts, pw in cs:
        B.add(D, pts, pw)
    ref = [single(D, pts, pw, 1e-10, pr.PRX_RESID_TRUE, 0)[0] for D, pts, pw in cs]
    for rnd in range(3):  # the captured graphs are replayed by every solve
        rep, per = B.solve(eps=1e-10, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=True)
        assert rep["tile"] in (8, 16) and rep["gs_block"] in (8, 16)
        assert rep["chunk"] % 2 == 0 and 4 <= rep["chunk"] <= 64
        assert rep["graph_launches"] > 0
        for k in range(len(cs)):
            xb = B.tree(k).solution
            assert np.abs(xb - ref[k]).max() <= 1e-9 * np.abs(ref[k]).max(), (rnd, k)


def test_batch_trees_evaluate_like_single_trees(gpu):
    cs = clouds()[:3]
    B = Batch()
    for D, pts, pw in cs:
        B.add(D, pts, pw)
    B.solve(eps=1e-10, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=False)
    for k, (D, pts, pw) in enumerate(cs):
        q = np.stack([pts["px"], pts["py"], pts["pz"]], axis=1)
        tb = B.tree(k)
        _, _, ts = single(D, pts, pw, 1e-10, pr.PRX_RESID_TRUE, 0)
        ts.solution = tb.solution
        assert np.array_equal(tb.evaluate_many(q), ts.evaluate_many(q))
        assert tb.iso_value() == ts.iso_value()


def test_batch_rejects_what_it_does_not_do(gpu):
    B = Batch()
    with pytest.raises(pr.PrxError):
        B.add(5, unit(synthetic.sphere(1000)), 4.0)  # base depth D-2 must be >= 4
    B.add(6, unit(synthetic.sphere(1000)), 4.0)
    import ctypes
    from poissonreconstruction_amd import _lib as L
    o = L.default_solve_opts()
    o.precision = L.PRX_PREC_MIXED
    o.cg_form = L.PRX_CG_TWO_PASS  # the batched mixed solve runs the fused CG step only
    rc = L.lib().prx_solve_batched(B._h, 20, 1000, 1e-6, ctypes.byref(o), 0, None, None)
    assert rc == L.PRX_ERR_INVALID


def single_mixed(D, pts, pw, eps):
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    rep = t.solve_cascadic(pts, pw, D - 2, 20, 800000, eps, precision=pr.PRX_PREC_MIXED, residual_mode=pr.PRX_RESID_TRUE,
                           refresh_period=0)
    return t.solution, rep


@pytest.mark.parametrize("autotune,form", [(False, pr.PRX_CG_FUSED), (True, pr.PRX_CG_FUSED),
                                           (False, pr.PRX_CG_FUSED_FACES)])
def test_batch_mixed_matches_single_tree_mixed_solves(gpu, autotune, form):
    # the batched mixed cascadic solve (FP32 correction CG with FP64 residual
    # replacements per instance; its FP32 step with q stored, or in face form)
    # against each cloud solved alone by the single-tree mixed solve: both
    # reach the FP64 true residual, and the solutions agree to the accuracy
    # that residual fixes
    cs = clouds()
    B = Batch()
    for D, pts, pw in cs:
        B.add(D, pts, pw)
    eps = 1e-9
    for rnd in range(2):  # the second solve replays the captured graphs
        rep, per = B.solve(eps=eps, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=autotune,
                           precision=pr.PRX_PREC_MIXED, cg_form=form)
        assert rep["instances"] == len(cs) and rep["converged"] == len(cs) and rep["status"] == pr.PRX_OK
        assert rep["max_true_rel"] <= eps * 1.0001
        for k, (D, pts, pw) in enumerate(cs):
            xs, rs = single_mixed(D, pts, pw, eps)
            xb = B.tree(k).solution
            rng_f = xs.max() - xs.min()
            print(f"  round {rnd} instance {k} depth {D}: batched {per[k]['fine_iters']} iterations "
                  f"({per[k]['outer_iters']} replacements), alone {rs['fine_iters']} ({rs['outer_iters']}); "
                  f"max diff {np.abs(xb - xs).max() / rng_f:.2e} of range", flush=True)
            assert per[k]["true_rel"] <= eps * 1.0001 and per[k]["status"] == pr.PRX_OK
            assert np.abs(xb - xs).max() <= 1e-6 * rng_f, (rnd, k)
            assert abs(per[k]["fine_iters"] - rs["fine_iters"]) <= max(4, 0.1 * rs["fine_iters"]), (rnd, k)
    # and the FP64 batched solve after the mixed one (vectors swapped back)
    rep, per = B.solve(eps=1e-10, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=False)
    assert rep["converged"] == len(cs) and rep["max_true_rel"] <= 1e-10 * 1.0001


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
def test_batch_multigrid_matches_single_tree_multigrid(gpu, prec):
    # each instance runs the single-tree multigrid solve on one of several
    # streams: the same arithmetic, so the same bits and iteration counts
    cs = clouds()
    B = Batch()
    for D, pts, pw in cs:
        B.add(D, pts, pw)
    for rnd in range(2):
        rep, per = B.solve(eps=1e-10, precond=pr.PRX_PRECOND_MG, precision=prec)
        assert rep["instances"] == len(cs) and rep["converged"] == len(cs) and rep["status"] == pr.PRX_OK
        assert rep["max_true_rel"] <= 1e-10 and rep["max_fine_iters"] <= 80
        for k, (D, pts, pw) in enumerate(cs):
            t = pr.FEMTree(D)
            t.splat_normal_field(pts)
            rs = t.solve_cascadic(pts, pw, D - 2, 20, 800000, 1e-10, precision=prec, residual_mode=pr.PRX_RESID_TRUE,
                                  precond=pr.PRX_PRECOND_MG)
            assert per[k]["fine_iters"] == rs["fine_iters"], (rnd, k)
            assert np.array_equal(B.tree(k).solution, t.solution), (rnd, k)


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
def test_batch_persistent_launch_matches_single_tree_solves(gpu, prec):
    # each level's CG of every instance in one persistent launch (iterations,
    # residual replacements or confirmations and stops decided on the device)
    # against each cloud solved alone: both reach the FP64 true residual, the
    # solutions agree to what that residual fixes, the iteration counts match
    cs = clouds()
    B = Batch()
    for D, pts, pw in cs:
        B.add(D, pts, pw)
    eps = 1e-10 if prec == pr.PRX_PREC_FP64 else 1e-9
    for rnd in range(2):
        rep, per = B.solve(eps=eps, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=False, precision=prec,
                           launch=pr.PRX_BATCH_LAUNCH_PERSISTENT)
        print(f"  round {rnd}: {rep['persistent_launches']} persistent launches, {rep['persistent_rounds']} rounds, "
              f"{rep['persistent_blocks']} workgroups, {rep['ms_total']:.1f} ms", flush=True)
        assert rep["instances"] == len(cs) and rep["converged"] == len(cs) and rep["status"] == pr.PRX_OK
        assert rep["persistent_launches"] == 2 * (rnd + 1) and rep["max_true_rel"] <= eps * 1.0001
        for k, (D, pts, pw) in enumerate(cs):
            t = pr.FEMTree(D)
            t.splat_normal_field(pts)
            rs = t.solve_cascadic(pts, pw, D - 2, 20, 800000, eps, precision=prec, residual_mode=pr.PRX_RESID_TRUE,
                                  refresh_period=0)
            xs, xb = t.solution, B.tree(k).solution
            rng_f = xs.max() - xs.min()
            print(f"    instance {k} depth {D}: {per[k]['fine_iters']} fine / {per[k]['coarse_iters']} coarse iterations "
                  f"({per[k]['outer_iters']} replacements), alone {rs['fine_iters']} / {rs['coarse_iters']}; "
                  f"max diff {np.abs(xb - xs).max() / rng_f:.2e} of range", flush=True)
            assert per[k]["status"] == pr.PRX_OK and per[k]["true_rel"] <= eps * 1.0001
            if prec == pr.PRX_PREC_FP64:
                assert np.abs(xb - xs).max() <= 1e-9 * np.abs(xs).max(), (rnd, k)
                assert abs(per[k]["fine_iters"] - rs["fine_iters"]) <= 2, (rnd, k)
                assert abs(per[k]["coarse_iters"] - rs["coarse_iters"]) <= 2, (rnd, k)
            else:
                assert np.abs(xb - xs).max() <= 1e-6 * rng_f, (rnd, k)
                assert abs(per[k]["fine_iters"] - rs["fine_iters"]) <= max(4, 0.1 * rs["fine_iters"]), (rnd, k)


def test_batch_persistent_launch_rejects_what_it_does_not_run(gpu):
    B = Batch()
    for D, pts, pw in clouds()[:2]:
        B.add(D, pts, pw)
    with pytest.raises(pr.PrxError):  # the reference's stopping test with periodic refresh stays on graph replays
        B.solve(eps=1e-6, residual_mode=pr.PRX_RESID_REFERENCE, refresh_period=10, autotune=False,
                launch=pr.PRX_BATCH_LAUNCH_PERSISTENT)


def test_batch_persistent_launch_ends_at_its_deadline(gpu, monkeypatch):
    # a launch past its deadline sets the error word after its round; every workgroup leaves at its next
    # barrier, the host raises, and the batch solves again afterwards
    B = Batch()
    for D, pts, pw in clouds()[:3]:
        B.add(D, pts, pw)
    monkeypatch.setenv("PRX_PERSIST_DEADLINE_MS", "0")
    with pytest.raises(pr.PrxError, match="deadline"):
        B.solve(eps=1e-6, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=False,
                launch=pr.PRX_BATCH_LAUNCH_PERSISTENT)
    monkeypatch.delenv("PRX_PERSIST_DEADLINE_MS")
    rep, _ = B.solve(eps=1e-6, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=False,
                     launch=pr.PRX_BATCH_LAUNCH_PERSISTENT)
    assert rep["converged"] == rep["instances"] and rep["max_true_rel"] <= 1e-6
    B.close()


@pytest.mark.parametrize("prec", [pr.PRX_PREC_FP64, pr.PRX_PREC_MIXED])
def test_batch_persistent_launch_single_instance(gpu, prec):
    # a batch of one: the launch's work lists, brick counters and grid barrier with a single instance (many
    # workgroups find no brick in most rounds); the result equals the graph replays' and reaches the residual
    D, pts, pw = clouds()[0]
    B = Batch()
    B.add(D, pts, pw)
    eps = 1e-10 if prec == pr.PRX_PREC_FP64 else 1e-9
    gr, gper = B.solve(eps=eps, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=False, precision=prec,
                       launch=pr.PRX_BATCH_LAUNCH_GRAPHS)
    xg = B.tree(0).solution
    pr_, pper = B.solve(eps=eps, residual_mode=pr.PRX_RESID_TRUE, refresh_period=0, autotune=False, precision=prec,
                        launch=pr.PRX_BATCH_LAUNCH_PERSISTENT)
    xp = B.tree(0).solution
    assert gr["converged"] == 1 and pr_["converged"] == 1 and pr_["max_true_rel"] <= eps * 1.0001
    if prec == pr.PRX_PREC_FP64:
        assert pper[0]["fine_iters"] == gper[0]["fine_iters"]
        assert np.abs(xp - xg).max() <= 1e-9 * np.abs(xg).max()
    else:
        assert np.abs(xp - xg).max() <= 1e-6 * (xg.max() - xg.min())
    B.close()
