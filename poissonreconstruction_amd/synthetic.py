"""Seeded synthetic oriented point clouds for the benchmark workloads.

sphere: unit sphere, points uniform by area, exact outward normals.
torus:  R=1.0, r=0.35 about z, uniform by area, positions jittered N(0, jitter^2)
        per component, normals perturbed N(0, nnoise^2) per component then
        renormalised.
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
