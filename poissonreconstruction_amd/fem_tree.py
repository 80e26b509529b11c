"""FEMTree: the reference's FEM system object (src/fem_tree.rs:52-513) on the GPU.

Every method keeps the reference's name and argument meaning and runs on the
device through libprx's C ABI.  Vectors use the reference's FEM node order
(Morton code of the node offset, x in bit 0).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from ._lib import check, ptr


class MatrixFreeSystem:
    """The assembled fine operator (what assemble_system returns in the
    reference, src/fem_tree.rs:180-249), kept matrix-free on the device."""

    def __init__(self, tree: "FEMTree"):
        self._tree = tree

    def num_rows(self) -> int:
        return self._tree.fem_node_count

    def multiply_vector(self, x, precision: int = L.PRX_PREC_FP64) -> np.ndarray:
        """y = A x (SparseMatrix::multiply_vector, src/sparse.rs:105-117)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.num_rows(),):
            raise ValueError("x has the wrong length")
        y = np.empty_like(x)
        check(L.lib().prx_apply(self._tree._h, int(precision), ptr(x), ptr(y)))
        return y

    def mg_vcycle(self, b, precision: int = L.PRX_PREC_FP64, smooth: int = 0, omega: float = 0.0) -> np.ndarray:
        """z = M b, one V-cycle of the multigrid preconditioner (M ~ A^-1
        without the per-point screening; smooth 0 = 2 sweeps, omega 0 =
        Chebyshev weights, > 0 = damped Jacobi with that weight)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.shape != (self.num_rows(),):
            raise ValueError("b has the wrong length")
        z = np.empty_like(b)
        check(L.lib().prx_mg_vcycle(self._tree._h, int(precision), int(smooth), float(omega), ptr(b), ptr(z)))
        return z

    def nnz(self) -> int:
        v = ctypes.c_int64(0)
        check(L.lib().prx_csr_nnz(self._tree._h, ctypes.byref(v)))
        return v.value

    def to_csr(self):
        """(offsets, columns, values) of the CSR the reference would build,
        reproduced on the device (verification mode)."""
        n = self.num_rows()
        nnz = self.nnz()
        off = np.empty(n + 1, dtype=np.int64)
        col = np.empty(nnz, dtype=np.int64)
        val = np.empty(nnz, dtype=np.float64)
        check(L.lib().prx_export_csr(self._tree._h, ptr(off), ptr(col), ptr(val)))
        return off, col, val


class FEMTree:
    """FEMTree::new(max_depth) (src/fem_tree.rs:67-72) on device `device`."""

    def __init__(self, max_depth: int, device: int = 0):
        if not 0 <= int(max_depth) <= L.PRX_MAX_DEPTH:
            raise ValueError(f"depth must be in [0, {L.PRX_MAX_DEPTH}]")
        self.max_depth = int(max_depth)
        h = ctypes.c_void_p()
        check(L.lib().prx_tree_create(self.max_depth, int(device), ctypes.byref(h)))
        self._h = h
        self._points = None
        self._n_points = 0
        self._pw = None

    def close(self):
        if getattr(self, "_h", None):
            L.lib().prx_tree_destroy(self._h)
            self._h = None

    def release_workspace(self) -> None:
        """Free the solver workspace kept between solves (the next solve allocates again)."""
        check(L.lib().prx_tree_release_workspace(self._h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ tree
    @property
    def fem_node_count(self) -> int:
        return int(L.lib().prx_tree_node_count(self._h))

    def initialize(self, points=None, samples_per_node: float = 1.5) -> None:
        """Complete refinement to max_depth (src/fem_tree.rs:83-93); the
        points and samples_per_node are ignored, as in the reference."""

    def finalize(self) -> None:
        """Node numbering (src/fem_tree.rs:96-126) is fixed at creation."""

    @property
    def idx_to_offset(self) -> np.ndarray:
        out = np.empty((self.fem_node_count, 3), dtype=np.uint32)
        check(L.lib().prx_tree_export_nodes(self._h, ptr(out), None))
        return out

    @property
    def grid_to_idx(self) -> np.ndarray:
        out = np.empty(self.fem_node_count, dtype=np.int64)
        check(L.lib().prx_tree_export_nodes(self._h, None, ptr(out)))
        return out

    # ------------------------------------------------------------------ splat / assembly
    def splat_normal_field(self, points) -> None:
        """src/fem_tree.rs:130-178; `points` are in the unit cube."""
        pts = L.as_points(points)
        self._points = pts
        self._n_points = len(pts)
        self._pw = None
        check(L.lib().prx_splat_normal_field(self._h, ptr(pts), len(pts)))

    def splat_normal_field_device(self, pos_ptr: int, nrm_ptr: int, n: int) -> None:
        """splat_normal_field with positions and normals already in device
        memory, each [n][3] float64 (e.g. torch tensors' data_ptr()); the
        caller's writes to them must be complete (stream synchronized)."""
        self._points = None
        self._n_points = int(n)
        self._pw = None
        check(L.lib().prx_splat_normal_field_device(self._h, ctypes.c_void_p(pos_ptr), ctypes.c_void_p(nrm_ptr),
                                                    int(n)))

    def splat_times(self) -> dict:
        """ms of the phases of the last splat_normal_field (measurement)."""
        ms = np.zeros(5)
        check(L.lib().prx_splat_times(self._h, ptr(ms)))
        return dict(zip(("upload", "point_sort", "splat_lists", "screening_lists", "splat"), ms.tolist()))

    @property
    def normal_field(self) -> np.ndarray:
        nf = np.empty((self.fem_node_count, 3))
        check(L.lib().prx_export_normal_field(self._h, ptr(nf), None))
        return nf

    @property
    def node_weight(self) -> np.ndarray:
        w = np.empty(self.fem_node_count)
        check(L.lib().prx_export_normal_field(self._h, None, ptr(w)))
        return w

    def _assemble(self, points, pw: float) -> None:
        if points is not None and self._points is not None and len(L.as_points(points)) != len(self._points):
            raise ValueError("assembly uses the points given to splat_normal_field")
        if self._pw != float(pw):
            check(L.lib().prx_assemble(self._h, float(pw)))
            self._pw = float(pw)

    def assemble_system(self, points, pw: float) -> MatrixFreeSystem:
        """src/fem_tree.rs:180-249 (matrix-free on the device)."""
        self._assemble(points, pw)
        return MatrixFreeSystem(self)

    def assemble_rhs(self, points, pw: float) -> np.ndarray:
        """src/fem_tree.rs:251-299."""
        self._assemble(points, pw)
        b = np.empty(self.fem_node_count)
        check(L.lib().prx_export_rhs(self._h, ptr(b)))
        return b

    def build_system_at_depth(self, depth: int):
        """Coarse RHS (row-major) and stencil constants of
        build_system_at_depth (src/fem_tree.rs:346-404)."""
        rhs = np.empty(1 << (3 * int(depth)))
        coef = np.empty(4)
        check(L.lib().prx_coarse_system(self._h, int(depth), ptr(rhs), ptr(coef)))
        return rhs, coef

    def prolongate(self, coarse, coarse_depth: int) -> np.ndarray:
        """src/fem_tree.rs:310-343 (coarse vector in row-major order)."""
        c = np.ascontiguousarray(coarse, dtype=np.float64)
        out = np.empty(self.fem_node_count)
        check(L.lib().prx_prolongate(self._h, int(coarse_depth), ptr(c), ptr(out)))
        return out

    # ------------------------------------------------------------------ solve
    def solve_cascadic(self, points, pw: float, base_depth: int, gs: int, cg: int, eps: float,
                       precision: int = L.PRX_PREC_FP64, residual_mode: int = L.PRX_RESID_REFERENCE,
                       refresh_period: int = 10, inner_tol: float = 0.05, max_outer: int = 64,
                       precond: int = L.PRX_PRECOND_NONE, mg_smooth: int = 2, mg_omega: float = 0.0,
                       mg_inner_tol: float = 1e-4, cg_two_pass: bool = False, cg_form: int = None) -> dict:
        """src/fem_tree.rs:458-487; returns the solve report.  precond
        PRX_PRECOND_MG solves the same system with multigrid-preconditioned
        CG from x = 0 instead (stopping at ||b - A x|| <= eps ||b||).
        cg_two_pass selects the two-pass CG iteration instead of the fused
        one (the same iteration, other rounding); cg_form (PRX_CG_*) selects
        any form and overrides it."""
        self._assemble(points, pw)
        o = L.default_solve_opts()
        o.precision = int(precision)
        o.residual_mode = int(residual_mode)
        o.refresh_period = int(refresh_period)
        o.gs_iters = int(gs)
        o.inner_tol = float(inner_tol)
        o.max_outer = int(max_outer)
        o.precond = int(precond)
        o.mg_smooth = int(mg_smooth)
        o.mg_omega = float(mg_omega)
        o.mg_inner_tol = float(mg_inner_tol)
        o.cg_form = int(cg_form) if cg_form is not None else (L.PRX_CG_TWO_PASS if cg_two_pass else L.PRX_CG_FUSED)
        rep = L.SolveReport()
        check(L.lib().prx_solve_cascadic(self._h, int(base_depth), int(gs), int(cg), float(eps), ctypes.byref(o),
                                         ctypes.byref(rep)))
        return rep.as_dict()

    @property
    def solution(self) -> np.ndarray:
        x = np.empty(self.fem_node_count)
        check(L.lib().prx_get_solution(self._h, ptr(x)))
        return x

    @solution.setter
    def solution(self, x) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(L.lib().prx_set_solution(self._h, ptr(x)))

    def gauss_seidel(self, b, x, sweeps: int = 1, precision: int = L.PRX_PREC_FP64) -> np.ndarray:
        """Multicolour Gauss-Seidel sweeps on the fine system."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        check(L.lib().prx_gs(self._h, int(precision), int(sweeps), ptr(b), ptr(x)))
        return x

    def solve_cg(self, b, x, max_iters: int, eps: float, residual_mode: int = L.PRX_RESID_REFERENCE,
                 refresh_period: int = 10):
        """solve_cg (src/solvers.rs:20-69) on the fine system; returns (x, iters, rel)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        it = ctypes.c_uint64(0)
        rel = ctypes.c_double(0.0)
        check(L.lib().prx_cg(self._h, L.PRX_PREC_FP64, ptr(b), ptr(x), int(max_iters), float(eps), int(residual_mode),
                             int(refresh_period), ctypes.byref(it), ctypes.byref(rel)))
        return x, it.value, rel.value

    # ------------------------------------------------------------------ evaluate
    def evaluate(self, p) -> float:
        """src/fem_tree.rs:490-513 at one point."""
        return float(self.evaluate_many(np.asarray(p, dtype=np.float64).reshape(1, 3))[0])

    def evaluate_many(self, q) -> np.ndarray:
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        out = np.empty(q.shape[0])
        check(L.lib().prx_evaluate(self._h, ptr(q), q.shape[0], ptr(out)))
        return out

    def evaluate_device(self, q_ptr: int, m: int, out_ptr: int) -> None:
        """evaluate_many with q [m][3] float64 and out [m] float64 in device
        memory; returns once out is written."""
        check(L.lib().prx_evaluate_device(self._h, ctypes.c_void_p(q_ptr), int(m), ctypes.c_void_p(out_ptr)))

    def evaluate_corners(self) -> np.ndarray:
        """Field at every marching-cubes corner, shape (res+1,)*3 indexed [z,y,x]."""
        r1 = (1 << self.max_depth) + 1
        out = np.empty(r1 ** 3)
        check(L.lib().prx_evaluate_corners(self._h, ptr(out)))
        return out.reshape(r1, r1, r1)

    def vjp_points(self, dL_dF, precision: int = L.PRX_PREC_MIXED, eps: float = 1e-10):
        """Gradient of a loss on F = the field at the splatted points (the
        current solution must be converged): returns (dL/dpositions [P,3],
        dL/dnormals [P,3], report) in unit-cube coordinates."""
        g = np.ascontiguousarray(dL_dF, dtype=np.float64).reshape(-1)
        P = self._n_points
        if g.shape[0] != P:
            raise ValueError("dL_dF must have one entry per splatted point")
        gpos = np.empty((P, 3))
        gnrm = np.empty((P, 3))
        rep = L.VjpReport()
        check(L.lib().prx_vjp_points(self._h, ptr(g), ptr(gpos), ptr(gnrm), int(precision), float(eps), ctypes.byref(rep)))
        return gpos, gnrm, {k: getattr(rep, k) for k, _ in rep._fields_}

    def vjp_points_device(self, g_ptr: int, gpos_ptr: int, gnrm_ptr: int, precision: int = L.PRX_PREC_MIXED,
                          eps: float = 1e-10) -> dict:
        """vjp_points with dL/dF [P], dL/dpositions [P][3] and dL/dnormals
        [P][3] float64 in device memory; returns the report once they are written."""
        rep = L.VjpReport()
        check(L.lib().prx_vjp_points_device(self._h, ctypes.c_void_p(g_ptr), ctypes.c_void_p(gpos_ptr),
                                            ctypes.c_void_p(gnrm_ptr), int(precision), float(eps), ctypes.byref(rep)))
        return {k: getattr(rep, k) for k, _ in rep._fields_}

    def iso_value(self) -> float:
        """Mean field over the splatted points (src/reconstructor.rs:56)."""
        v = ctypes.c_double(0.0)
        check(L.lib().prx_iso_value(self._h, ctypes.byref(v)))
        return v.value

    def extract_surface(self, iso: float):
        """FEMTree::extract_surface (src/fem_tree.rs:576-636) of the current
        solution: (vertices (n,3) unit cube, triangles (m,3) uint64)."""
        mesh = L.Mesh()
        check(L.lib().prx_extract_surface(self._h, float(iso), ctypes.byref(mesh)))
        try:
            nv, nt = int(mesh.n_vertices), int(mesh.n_triangles)
            verts = np.ctypeslib.as_array(mesh.vertices, shape=(nv * 3,)).reshape(nv, 3).copy() if nv else np.zeros((0, 3))
            tris = (np.ctypeslib.as_array(mesh.triangles, shape=(nt * 3,)).reshape(nt, 3).copy() if nt
                    else np.zeros((0, 3), dtype=np.uint64))
        finally:
            L.lib().prx_mesh_free(ctypes.byref(mesh))
        return verts, tris
