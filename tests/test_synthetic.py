"""Synthetic workload generators: the properties their definitions fix."""
import numpy as np

from poissonreconstruction_amd import synthetic


def test_superquadric_normals_are_the_inside_outside_gradient():
    pts = synthetic.superquadric(2000, seed=0, noise=0.0)
    # the generator's first draws: exponents, then scales (same uniform stream)
    u = np.random.default_rng(0).random(5)
    e1, e2 = 0.3 + 2.7 * u[:2]
    a = 0.5 + 0.5 * u[2:]

    def F(x):
        u = np.abs(x[:, 0] / a[0]) ** (2 / e2) + np.abs(x[:, 1] / a[1]) ** (2 / e2)
        return u ** (e2 / e1) + np.abs(x[:, 2] / a[2]) ** (2 / e1)

    p = pts[:, :3]
    assert np.abs(F(p) - 1.0).max() < 1e-9  # on the surface
    keep = np.all(np.abs(p) > 1e-3, axis=1)  # the gradient is singular on the coordinate planes for e > 2
    p, n = p[keep], pts[keep, 3:]
    h = 1e-7
    g = np.stack([(F(p + h * np.eye(3)[k]) - F(p - h * np.eye(3)[k])) / (2 * h) for k in range(3)], axis=1)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    assert np.median(np.sum(g * n, axis=1)) > 1 - 1e-6
    assert np.mean(np.sum(g * n, axis=1) > 0.999) > 0.99


def test_batch_sizes_are_log_uniform_in_range():
    s = synthetic.batch_sizes(256, seed=0)
    assert s.min() >= 2000 and s.max() <= 200000
    # median of a log-uniform on [2e3, 2e5] is 2e4
    assert 1.2e4 < np.median(s) < 3.3e4


def test_ellipsoid_union_points_lie_on_exactly_one_surface():
    pts = synthetic.ellipsoids(3000, seed=1)
    assert pts.shape == (3000, 6)
    assert np.allclose(np.linalg.norm(pts[:, 3:], axis=1), 1.0)


def test_displaced_sphere_points_and_normals():
    pts = synthetic.displaced_sphere(4000, seed=2)
    S = synthetic.SHDisplacement(seed=2)
    p, n = pts[:, :3], pts[:, 3:]
    rad = np.linalg.norm(p, axis=1)
    th = np.arccos(p[:, 2] / rad)
    ph = np.arctan2(p[:, 1], p[:, 0])
    assert np.abs(rad - S.r(th, ph)).max() < 1e-12  # on the surface
    assert np.abs(S.delta(th, ph)).max() <= 0.05 + 1e-9  # amplitude bound
    # the normal is orthogonal to the surface: a step along a tangent of the
    # surface (from the parametrisation) is orthogonal to it to O(step^2)
    eps = 1e-5
    for dth, dph in ((eps, 0.0), (0.0, eps)):
        t2, p2 = th + dth, ph + dph
        q = S.r(t2, p2)[:, None] * np.stack([np.sin(t2) * np.cos(p2), np.sin(t2) * np.sin(p2), np.cos(t2)], axis=1)
        tang = (q - p) / eps
        keep = np.sin(th) > 0.05
        cos = np.abs(np.sum(tang * n, axis=1)) / np.linalg.norm(tang, axis=1)
        assert np.median(cos[keep]) < 1e-4
    assert np.mean(np.sum(p * n, axis=1) > 0) == 1.0  # outward
