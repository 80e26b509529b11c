"""reconstruct() and its parameter/statistics types (src/reconstructor.rs).

Mirrors the reference's public API: ReconstructParams (:8-15) with the same
defaults, ReconstructStats (:17-23) and reconstruct(points, params) ->
(Mesh, ReconstructStats) (:25-70), running on the GPU through the C ABI.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, fields

import numpy as np

from . import _lib as L
from ._lib import check, ptr


@dataclass
class ReconstructParams:
    depth: int = 8
    samples_per_node: float = 1.5
    scale: float = 1.1
    cg_accuracy: float = 1e-3
    cg_iters: int = 500
    point_weight: float = 0.0
    verbose: bool = False
    # extensions (not in the reference)
    precision: int = L.PRX_PREC_FP64
    residual_mode: int = L.PRX_RESID_REFERENCE
    device: int = 0
    gs_iters: int = 20
    base_depth: int = -1
    refresh_period: int = 10
    inner_tol: float = 0.05
    max_outer: int = 64
    precond: int = L.PRX_PRECOND_NONE

    def to_c(self) -> L.Params:
        p = L.Params()
        for f in fields(self):
            setattr(p, f.name, int(getattr(self, f.name)) if f.name == "verbose" else getattr(self, f.name))
        return p


@dataclass
class ReconstructStats:
    time_transform: float = 0.0
    time_tree: float = 0.0
    time_system: float = 0.0
    time_solve: float = 0.0
    time_extract: float = 0.0
    time_total: float = 0.0
    fem_nodes: int = 0
    matrix_nnz: int = 0
    mesh_vertices: int = 0
    mesh_triangles: int = 0
    iso_value: float = 0.0
    coarse_iters: int = 0
    fine_iters: int = 0
    outer_iters: int = 0
    final_rel_residual: float = 0.0
    solve_ms: float = 0.0
    solve_status: int = 0

    @staticmethod
    def from_c(s: L.Stats) -> "ReconstructStats":
        return ReconstructStats(**{f.name: getattr(s, f.name) for f in fields(ReconstructStats)})


@dataclass
class Mesh:
    """Mesh (src/marching_cubes.rs:8-11): vertices (n,3) f64, triangles (m,3) u64."""
    vertices: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    triangles: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), dtype=np.uint64))


def reconstruct(points, params: ReconstructParams | None = None):
    """reconstruct (src/reconstructor.rs:25-70).

    `points`: structured array of OrientedPoint, an (n,6) array
    [px,py,pz,nx,ny,nz], or a (positions, normals) pair.  Empty input gives an
    empty mesh and zeroed stats, as in the reference.
    """
    params = params or ReconstructParams()
    pts = L.as_points(points) if len(points) else np.zeros(0, dtype=L.POINT_DTYPE)
    cp = params.to_c()
    mesh = L.Mesh()
    st = L.Stats()
    check(L.lib().prx_reconstruct(ptr(pts) if len(pts) else None, len(pts), ctypes.byref(cp), ctypes.byref(mesh),
                                  ctypes.byref(st)))
    try:
        nv, nt = int(mesh.n_vertices), int(mesh.n_triangles)
        verts = np.ctypeslib.as_array(mesh.vertices, shape=(nv * 3,)).reshape(nv, 3).copy() if nv else np.zeros((0, 3))
        tris = (np.ctypeslib.as_array(mesh.triangles, shape=(nt * 3,)).reshape(nt, 3).copy() if nt
                else np.zeros((0, 3), dtype=np.uint64))
    finally:
        L.lib().prx_mesh_free(ctypes.byref(mesh))
    return Mesh(verts, tris), ReconstructStats.from_c(st)
