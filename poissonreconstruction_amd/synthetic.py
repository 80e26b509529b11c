"""Seeded synthetic oriented point clouds for the benchmark workloads.

sphere: unit sphere, points uniform by area, exact outward normals.
torus:  R=1.0, r=0.35 about z, uniform by area, positions jittered N(0, jitter^2)
        per component, normals perturbed N(0, nnoise^2) per component then
        renormalised.
ellipsoids: union of 3 ellipsoids (semi-axes U[0.3,0.8], uniform random
        rotations, centres U[-0.3,0.3]^3), points uniform by area on the union's
        surface (points inside another ellipsoid dropped), outward normals.
Each returns an (n, 6) float64 array [px, py, pz, nx, ny, nz].
"""
from __future__ import annotations

import numpy as np


def sphere(n: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    g = rng.normal(size=(n, 3))
    p = g / np.linalg.norm(g, axis=1, keepdims=True)
    return np.ascontiguousarray(np.concatenate([p, p], axis=1))


def torus(n: int, R: float = 1.0, r: float = 0.35, jitter: float = 1e-3, nnoise: float = 0.02, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    us, vs = [], []
    need = n
    while need > 0:
        m = int(need * 1.6) + 16
        u = rng.uniform(0.0, 2 * np.pi, m)
        v = rng.uniform(0.0, 2 * np.pi, m)
        keep = rng.uniform(0.0, 1.0, m) < (R + r * np.cos(v)) / (R + r)  # area element of the torus
        us.append(u[keep][:need])
        vs.append(v[keep][:need])
        need -= len(us[-1])
    u = np.concatenate(us)[:n]
    v = np.concatenate(vs)[:n]
    cu, su, cv, sv = np.cos(u), np.sin(u), np.cos(v), np.sin(v)
    p = np.stack([(R + r * cv) * cu, (R + r * cv) * su, r * sv], axis=1)
    p = p + rng.normal(scale=jitter, size=p.shape)
    nr = np.stack([cv * cu, cv * su, sv], axis=1) + rng.normal(scale=nnoise, size=(n, 3))
    nr /= np.linalg.norm(nr, axis=1, keepdims=True)
    return np.ascontiguousarray(np.concatenate([p, nr], axis=1))


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def ellipsoids(n: int, k: int = 3, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    axes = rng.uniform(0.3, 0.8, size=(k, 3))
    rots = [_rotation(rng) for _ in range(k)]
    cents = rng.uniform(-0.3, 0.3, size=(k, 3))

    def inside(x, j):
        y = (x - cents[j]) @ rots[j]
        return np.sum((y / axes[j]) ** 2, axis=1) < 1.0

    def sample(j, m):
        # area-uniform on the ellipsoid x = R diag(a) u: the area element relative to the unit
        # sphere is a0 a1 a2 ||u / a||, so accept u with probability proportional to it
        out = []
        while sum(len(o) for o in out) < m:
            u = rng.normal(size=(4 * m + 64, 3))
            u /= np.linalg.norm(u, axis=1, keepdims=True)
            a = axes[j]
            dens = np.linalg.norm(u / a, axis=1) * a.prod()
            keep = rng.uniform(size=len(u)) * dens.max() < dens
            out.append(u[keep])
        u = np.concatenate(out)[:m]
        x = (u * axes[j]) @ rots[j].T + cents[j]
        nrm = ((u / axes[j]) @ rots[j].T)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        return x, nrm

    # apportion by the area left after removing covered parts (estimated on a pilot sample)
    pilot = [sample(j, 4000) for j in range(k)]
    area = []
    for j in range(k):
        x, _ = pilot[j]
        cov = np.zeros(len(x), dtype=bool)
        for i in range(k):
            if i != j:
                cov |= inside(x, i)
        a = axes[j]
        p = 1.6075
        full = 4 * np.pi * (((a[0] * a[1]) ** p + (a[0] * a[2]) ** p + (a[1] * a[2]) ** p) / 3) ** (1 / p)
        area.append(full * (1 - cov.mean()))
    share = np.array(area) / np.sum(area)
    counts = np.floor(share * n).astype(int)
    counts[0] += n - counts.sum()
    xs, ns = [], []
    for j in range(k):
        got_x, got_n = [], []
        need = counts[j]
        while need > 0:
            x, nr = sample(j, 2 * need + 64)
            keep = np.ones(len(x), dtype=bool)
            for i in range(k):
                if i != j:
                    keep &= ~inside(x, i)
            got_x.append(x[keep][:need])
            got_n.append(nr[keep][:need])
            need -= len(got_x[-1])
        xs.append(np.concatenate(got_x))
        ns.append(np.concatenate(got_n))
    return np.ascontiguousarray(np.concatenate([np.concatenate(xs), np.concatenate(ns)], axis=1))
