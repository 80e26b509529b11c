"""Marching cubes on the device (prx_extract_surface) against the oracle's
restatement of FEMTree::extract_surface (src/fem_tree.rs:576-636), and the
mesh of prx_reconstruct against the oracle pipeline."""
import numpy as np
import pytest

import oracle
import poissonreconstruction_amd as pr
from poissonreconstruction_amd import synthetic

pytestmark = pytest.mark.gpu


def unit_points(raw):
    m, _ = oracle.unit_transform(raw, 1.1)
    return oracle.apply_transform(raw, m)


@pytest.mark.parametrize("D,cloud", [(5, "sphere"), (6, "torus")])
def test_extract_surface_matches_oracle_on_the_same_field(gpu, D, cloud):
    raw = synthetic.sphere(3000, seed=3) if cloud == "sphere" else synthetic.torus(8000, seed=4)
    pts = unit_points(raw)
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    t.solve_cascadic(pts, 4.0, D - 2, 20, 40000, 1e-10, precision=pr.PRX_PREC_FP64, residual_mode=pr.PRX_RESID_TRUE,
                     refresh_period=0)
    x = t.solution
    iso = t.iso_value()
    V, F = t.extract_surface(iso)
    g2i, i2o = oracle.tree(D)
    Vo, Fo = oracle.extract_surface(D, g2i, i2o, x, iso)
    assert len(V) > 100 and len(F) > 100
    # same vertices in the same (insertion) order, bit for bit, and the same triangles in the same order
    assert V.shape == Vo.shape and F.shape == Fo.shape
    assert np.array_equal(V.view(np.uint64), Vo.view(np.uint64))
    assert np.array_equal(F.astype(np.int64), Fo)


def test_reconstruct_returns_the_oracle_mesh(gpu):
    raw = synthetic.sphere(4000, seed=8)
    p = pr.ReconstructParams(depth=5, point_weight=4.0, cg_accuracy=1e-12, cg_iters=40000)
    mesh, st = pr.reconstruct(raw, p)
    up, sol, iso, rep, m = oracle.reconstruct_field(raw, depth=5, cg_accuracy=1e-12, cg_iters=40000, point_weight=4.0)
    g2i, i2o = oracle.tree(5)
    Vu, Fo = oracle.extract_surface(5, g2i, i2o, sol, iso)
    _, mi = oracle.unit_transform(raw, 1.1)
    M = np.asarray(mi).reshape(4, 4).T  # column-major m[col*4+row]
    Vo = Vu @ M[:3, :3].T + M[:3, 3]
    assert st.mesh_vertices == len(mesh.vertices) and st.mesh_triangles == len(mesh.triangles)
    assert abs(st.iso_value - iso) <= 1e-9 * abs(iso) + 1e-15
    # the GPU and oracle solutions agree to the solve tolerance, so the meshes agree to round-off
    assert mesh.vertices.shape == Vo.shape and mesh.triangles.shape == Fo.shape
    assert np.array_equal(mesh.triangles.astype(np.int64), Fo)
    assert np.abs(mesh.vertices - Vo).max() < 1e-8


def test_surface_of_a_larger_tree_is_closed(gpu):
    D = 7
    pts = unit_points(synthetic.torus(100000, seed=5))
    t = pr.FEMTree(D)
    t.splat_normal_field(pts)
    t.solve_cascadic(pts, 4.0, D - 2, 20, 40000, 1e-7, precision=pr.PRX_PREC_MIXED, residual_mode=pr.PRX_RESID_TRUE,
                     refresh_period=0)
    V, F = t.extract_surface(t.iso_value())
    F = F.astype(np.int64)
    assert F.min() >= 0 and F.max() < len(V)
    # vertices on the grid: two coordinates on grid lines
    on = np.isclose(V * (1 << D), np.round(V * (1 << D)), atol=1e-9, rtol=0)
    assert np.all(on.sum(axis=1) >= 2)
    # every mesh edge away from the domain boundary is shared by exactly two
    # triangles (the 1.1-scaled cloud crosses the unit cube's faces, Q1, and the
    # clamped boundary corners open the surface there)
    e = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
    u, cnt = np.unique(e, axis=0, return_counts=True)
    h = 1.0 / (1 << D)
    inner = np.all((V[u[:, 0]] > 2 * h) & (V[u[:, 0]] < 1 - 2 * h) & (V[u[:, 1]] > 2 * h) & (V[u[:, 1]] < 1 - 2 * h), axis=1)
    assert inner.sum() > 10000
    assert np.all(cnt[inner] == 2)


def test_command_line_writes_the_reconstruct_mesh(gpu, tmp_path):
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
    root = __import__("os").path.dirname(__import__("os").path.dirname(__import__("os").path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-m", "poissonreconstruction_amd", "-i", str(src), "-o", str(out), "-d", "5",
                        "--point-weight", "4", "-v"], capture_output=True, text=True, cwd=root)
    assert r.returncode == 0, r.stderr
    assert "Mesh:" in r.stderr and "Wrote" in r.stderr
    mesh, st = pr.reconstruct(ply.read_oriented_points(str(src)), pr.ReconstructParams(depth=5, point_weight=4.0))
    lines = open(out).read().split("\n")
    nv, nt = int(lines[2].split()[-1]), int(lines[6].split()[-1])
    assert (nv, nt) == (len(mesh.vertices), len(mesh.triangles))
    V = np.array([[float(t) for t in l.split()] for l in lines[9:9 + nv]])
    assert np.array_equal(V, mesh.vertices)
