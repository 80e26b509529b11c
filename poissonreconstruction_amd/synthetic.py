"""Seeded synthetic oriented point clouds for the benchmark workloads.

sphere: unit sphere, points uniform by area, exact outward normals.
torus:  R=1.0, r=0.35 about z, uniform by area, positions jittered N(0, jitter^2)
        per component, normals perturbed N(0, nnoise^2) per component then
        renormalised.
ellipsoids: union of 3 ellipsoids (semi-axes U[0.3,0.8], uniform random
        rotations, centres U[-0.3,0.3]^3), points uniform by area on the union's
        surface (points inside another ellipsoid dropped), outward normals.
superquadric: x = (a1 C(eta,e1) C(w,e2), a2 C(eta,e1) S(w,e2), a3 S(eta,e1)) with
        C(t,e) = sign(cos t)|cos t|^e, S likewise; exponents e1,e2 ~ U[0.3,3],
        scales a ~ U[0.5,1]^3; points uniform by area (a fine parameter grid's
        triangles drawn by area, uniform inside, mapped through the surface),
        position noise N(0, noise^2), exact outward normals (the gradient of the
        inside-outside function at the noise-free point).
displaced_sphere: unit sphere displaced radially, r(theta,phi) = 1 + sum_k c_k Y_k
        over 8 random real spherical harmonics with l <= 6, the coefficients
        scaled so that max |sum c_k Y_k| = amp (0.05) on a dense grid; points
        uniform by area (direction rejection on the area element
        r sqrt(r^2 + |grad_S r|^2)), normals the gradient of |x| - r(theta,phi).
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


def _spow(v, e):
    return np.sign(v) * np.abs(v) ** e


def superquadric(n: int, seed: int = 0, noise: float = 5e-3, grid: int = 256) -> np.ndarray:
    rng = np.random.default_rng(seed)
    e1, e2 = rng.uniform(0.3, 3.0, size=2)
    a = rng.uniform(0.5, 1.0, size=3)

    def surf(eta, w):
        ce, se, cw, sw = np.cos(eta), np.sin(eta), np.cos(w), np.sin(w)
        return np.stack([a[0] * _spow(ce, e1) * _spow(cw, e2), a[1] * _spow(ce, e1) * _spow(sw, e2), a[2] * _spow(se, e1)], axis=-1)

    # triangulated parameter grid; triangles drawn by their area
    ge = np.linspace(-np.pi / 2, np.pi / 2, grid + 1)
    gw = np.linspace(-np.pi, np.pi, 2 * grid + 1)
    E, W = np.meshgrid(ge, gw, indexing="ij")
    X = surf(E, W)
    p00, p10, p01, p11 = X[:-1, :-1], X[1:, :-1], X[:-1, 1:], X[1:, 1:]
    a1 = 0.5 * np.linalg.norm(np.cross(p10 - p00, p01 - p00), axis=-1)
    a2 = 0.5 * np.linalg.norm(np.cross(p10 - p11, p01 - p11), axis=-1)
    areas = np.stack([a1, a2], axis=-1).reshape(-1)
    tri = rng.choice(areas.size, size=n, p=areas / areas.sum())
    cell, half = tri // 2, tri % 2
    i, j = cell // (2 * grid), cell % (2 * grid)
    u, v = rng.uniform(size=n), rng.uniform(size=n)
    flip = u + v > 1.0
    u, v = np.where(flip, 1.0 - u, u), np.where(flip, 1.0 - v, v)
    de, dw = ge[1] - ge[0], gw[1] - gw[0]
    # lower triangle (00,10,01); upper triangle (11,01,10) -> parameters
    eta = np.where(half == 0, ge[i] + u * de, ge[i + 1] - u * de)
    w = np.where(half == 0, gw[j] + v * dw, gw[j + 1] - v * dw)
    p = surf(eta, w)
    # normal: gradient of ((|x/a1|^(2/e2) + |y/a2|^(2/e2))^(e2/e1) + |z/a3|^(2/e1)), direction
    # (C(eta,2-e1) C(w,2-e2) / a1, C(eta,2-e1) S(w,2-e2) / a2, S(eta,2-e1) / a3)
    ce = np.cos(eta)
    ce = np.where(np.abs(ce) < 1e-12, 1e-12, ce)
    nr = np.stack([_spow(ce, 2 - e1) * _spow(np.cos(w), 2 - e2) / a[0], _spow(ce, 2 - e1) * _spow(np.sin(w), 2 - e2) / a[1],
                   _spow(np.sin(eta), 2 - e1) / a[2]], axis=-1)
    nr /= np.linalg.norm(nr, axis=1, keepdims=True)
    p = p + rng.normal(scale=noise, size=p.shape)
    return np.ascontiguousarray(np.concatenate([p, nr], axis=1))


def batch_sizes(k: int = 256, lo: float = 2e3, hi: float = 2e5, seed: int = 0) -> np.ndarray:
    """Cloud sizes round(exp(U[ln lo, ln hi]))."""
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), size=k))).astype(np.int64)


class SHDisplacement:
    """r(theta, phi) = 1 + sum_k c_k Y_{l_k m_k}(theta, phi), real spherical
    harmonics (theta polar, phi azimuth), with its angular derivatives."""

    def __init__(self, k: int = 8, lmax: int = 6, amp: float = 0.05, seed: int = 0):
        rng = np.random.default_rng(seed)
        self.terms = []
        for _ in range(k):
            l = int(rng.integers(1, lmax + 1))
            m = int(rng.integers(-l, l + 1))
            self.terms.append((l, m, float(rng.normal())))
        th, ph = np.meshgrid(np.linspace(0, np.pi, 721), np.linspace(-np.pi, np.pi, 1441), indexing="ij")
        peak = np.abs(self.delta(th, ph)).max()
        self.terms = [(l, m, c * amp / peak) for l, m, c in self.terms]

    @staticmethod
    def _real_sh(l, m, th, ph, grad=False):
        """Real Y_lm and, with grad, (dY/dtheta, dY/dphi)."""
        from scipy.special import sph_harm_y
        if grad:
            y, g = sph_harm_y(l, abs(m), th, ph, diff_n=1)
            parts = (y, g[..., 0], g[..., 1])
        else:
            parts = (sph_harm_y(l, abs(m), th, ph),)
        if m == 0:
            out = tuple(q.real for q in parts)
        else:
            f = np.sqrt(2.0) * (-1) ** m
            out = tuple(f * (q.real if m > 0 else q.imag) for q in parts)
        return out if grad else out[0]

    def delta(self, th, ph):
        return sum(c * self._real_sh(l, m, th, ph) for l, m, c in self.terms)

    def r(self, th, ph):
        return 1.0 + self.delta(th, ph)

    def r_and_grad(self, th, ph):
        """(r, dr/dtheta, dr/dphi)."""
        r = np.ones_like(th)
        dth = np.zeros_like(th)
        dph = np.zeros_like(th)
        for l, m, c in self.terms:
            y, gt, gp = self._real_sh(l, m, th, ph, grad=True)
            r += c * y
            dth += c * gt
            dph += c * gp
        return r, dth, dph


def displaced_sphere(n: int, seed: int = 0, k: int = 8, lmax: int = 6, amp: float = 0.05) -> np.ndarray:
    S = SHDisplacement(k, lmax, amp, seed)
    rng = np.random.default_rng(seed + 1)

    def density(r, dth, dph, th):
        return r * np.sqrt(r * r + dth * dth + (dph / np.maximum(np.sin(th), 1e-12)) ** 2)

    # bound of the area element from a grid (x1.25 margin against the grid's gaps)
    gt, gp = np.meshgrid(np.linspace(1e-3, np.pi - 1e-3, 361), np.linspace(-np.pi, np.pi, 721), indexing="ij")
    dmax = 1.25 * density(*S.r_and_grad(gt, gp), gt).max()
    keep_th, keep_ph, keep_r, keep_dt, keep_dp = [], [], [], [], []
    need = n
    while need > 0:
        u = rng.normal(size=(int(need * 1.5) + 1024, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        th = np.arccos(np.clip(u[:, 2], -1.0, 1.0))
        ph = np.arctan2(u[:, 1], u[:, 0])
        r, dth, dph = S.r_and_grad(th, ph)
        keep = np.flatnonzero(rng.uniform(size=len(u)) * dmax < density(r, dth, dph, th))[:need]
        for lst, a in zip((keep_th, keep_ph, keep_r, keep_dt, keep_dp), (th, ph, r, dth, dph)):
            lst.append(a[keep])
        need -= len(keep)
    th, ph, r, dth, dph = (np.concatenate(a) for a in (keep_th, keep_ph, keep_r, keep_dt, keep_dp))
    st, ct, sp_, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    radial = np.stack([st * cp, st * sp_, ct], axis=1)
    p = r[:, None] * radial
    # grad(|x| - r(theta, phi)) = x/|x| - (dr/dtheta e_theta + dr/dphi / sin(theta) e_phi) / |x|
    e_th = np.stack([ct * cp, ct * sp_, -st], axis=1)
    e_ph = np.stack([-sp_, cp, np.zeros_like(th)], axis=1)
    g = radial - (dth[:, None] * e_th + (dph / np.maximum(st, 1e-12))[:, None] * e_ph) / r[:, None]
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    return np.ascontiguousarray(np.concatenate([p, g], axis=1))
