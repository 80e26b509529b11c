"""FP64 CPU oracle for the screened-Poisson solve path -- TEST INFRASTRUCTURE.

Only the tests, ``__graft_entry__.smoke()`` and the benchmark's ``cpu_baseline``
leg import this package, and only as the checker or the timed CPU baseline.
The product library (``poissonreconstruction_amd``) never imports it.

The C source ``prx_oracle.c`` restates the reference's algorithm function by
function (citations in its comments).  The reference is a Rust crate and there
is no Rust toolchain, so it cannot be built into ``oracle/_ref``; the
restatement is pinned instead by the checks listed below.

Parity pins (test file that holds each check):
  tree / Morton index     DFS restatement vs closed-form Morton code, node counts
                          1+8+64 (src/octree.rs:436-450)        tests/test_oracle.py
  stencil constants       exact rational hat-function integrals; B-spline KATs
                          (src/bspline.rs:399-437)              tests/test_oracle.py
  CSR assembly            symmetry, 21 nonzeros, zero faces, screening sum identity,
                          rank-1 one-point block, dense brute force at D<=2
                                                                tests/test_oracle.py
  splat / RHS             one-point weight orientation (node c gets (1-t)^3 products),
                          two-point negated weighted average, total weight as a
                          product of in-range 1-D weights; constant-normal and
                          closed-form RHS invariants             tests/test_oracle.py
  coarse system/prolong   impulse placement at the row-major coarse cell of
                          f >> (D-d); prolongation impulse response and exact
                          reproduction of a linear field with distinct slopes
                          per axis; screening sum identities    tests/test_oracle.py
  marching-cubes tables   SHA-256 digests of the reference's TRI_TABLE,
                          EDGE_TABLE and corner/edge topology (written from
                          src/tri_table.inl and src/fem_tree.rs:13-30,587-588
                          by tests/golden/make_mc_digest.py)    tests/test_mc_oracle.py
  GS                      hand-computed sweep; fixed point = direct solve
                                                                tests/test_oracle.py
  CG                      KAT src/solvers.rs:234-268; scipy direct solve
                                                                tests/test_oracle.py
  SpMV                    KAT src/sparse.rs:174-196            tests/test_oracle.py
  evaluate                closed form at integer s, clamping special cases
                                                                tests/test_oracle.py
  iso mean, CG iterates   tolerance only: the reference sums with rayon, whose
                          reduction order is not fixed, so the last bits of these
                          are unpinnable ("parity unpinned" at ulp level).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

POINT_DTYPE = np.dtype([("px", "f8"), ("py", "f8"), ("pz", "f8"), ("nx", "f8"), ("ny", "f8"), ("nz", "f8")])

_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class _SolveReport(ctypes.Structure):
    _fields_ = [("t_assemble", ctypes.c_double), ("t_coarse", ctypes.c_double), ("t_fine_gs", ctypes.c_double),
                ("t_fine_cg", ctypes.c_double), ("coarse_iters", ctypes.c_int64), ("fine_iters", ctypes.c_int64),
                ("nnz", ctypes.c_int64), ("coarse_rel", ctypes.c_double), ("fine_rel", ctypes.c_double)]


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_tree.restype = ctypes.c_int64
        L.orc_tree.argtypes = [ctypes.c_uint32, _vp, _vp]
        L.orc_unit_transform.argtypes = [_vp, ctypes.c_int64, ctypes.c_double, _vp, _vp]
        L.orc_apply_transform.argtypes = [_vp, ctypes.c_int64, _vp, _vp]
        L.orc_splat.argtypes = [ctypes.c_uint32, _vp, _vp, ctypes.c_int64, _vp, _vp]
        L.orc_assemble.restype = _vp
        L.orc_assemble.argtypes = [ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_double]
        L.orc_csr_rows.restype = ctypes.c_int64
        L.orc_csr_rows.argtypes = [_vp]
        L.orc_csr_nnz.restype = ctypes.c_int64
        L.orc_csr_nnz.argtypes = [_vp]
        L.orc_csr_export.argtypes = [_vp, _vp, _vp, _vp]
        L.orc_csr_import.restype = _vp
        L.orc_csr_import.argtypes = [ctypes.c_int64, _vp, _vp, _vp]
        L.orc_csr_free.argtypes = [_vp]
        L.orc_rhs.argtypes = [ctypes.c_uint32, _vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_double, _vp]
        L.orc_coarse_system.restype = _vp
        L.orc_coarse_system.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, ctypes.c_double, _vp]
        L.orc_prolongate.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp]
        L.orc_spmv.argtypes = [_vp, _vp, _vp]
        L.orc_dot.restype = ctypes.c_double
        L.orc_dot.argtypes = [_vp, _vp, ctypes.c_int64]
        L.orc_gs_sweep.argtypes = [_vp, _vp, _vp]
        L.orc_gs_sweep_order.argtypes = [_vp, _vp, _vp, _vp]
        L.orc_stencil_1d.argtypes = [_vp]
        L.orc_cg.restype = ctypes.c_int64
        L.orc_cg.argtypes = [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_double, _vp]
        L.orc_evaluate1.restype = ctypes.c_double
        L.orc_evaluate1.argtypes = [ctypes.c_uint32, _vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_evaluate.argtypes = [ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_int64, _vp]
        L.orc_extract_surface.restype = ctypes.c_int64
        L.orc_extract_surface.argtypes = [ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp]
        L.orc_free.argtypes = [_vp]
        L.orc_iso.restype = ctypes.c_double
        L.orc_iso.argtypes = [ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_int64]
        L.orc_solve_cascadic.argtypes = [ctypes.c_uint32, _vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_double,
                                         ctypes.c_uint32, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, _vp, _vp]
        L.orc_vjp_points.restype = ctypes.c_int64
        L.orc_vjp_points.argtypes = [ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_double, _vp, _vp,
                                     ctypes.c_double, _vp, _vp, _vp]
        L.orc_field_at_points.argtypes = [ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_double,
                                          ctypes.c_double, _vp, _vp]
        L.orc_stiffness_entry.restype = ctypes.c_double
        L.orc_stiffness_entry.argtypes = [ctypes.c_double, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def as_points(pos, nrm) -> np.ndarray:
    """Pack (P,3) positions and normals into the reference's OrientedPoint layout."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    nrm = np.asarray(nrm, dtype=np.float64).reshape(-1, 3)
    out = np.empty(pos.shape[0], dtype=POINT_DTYPE)
    out["px"], out["py"], out["pz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    out["nx"], out["ny"], out["nz"] = nrm[:, 0], nrm[:, 1], nrm[:, 2]
    return out


def points_array(pts) -> np.ndarray:
    if pts is None:
        return np.zeros(0, dtype=POINT_DTYPE)
    if pts.dtype == POINT_DTYPE:
        return np.ascontiguousarray(pts)
    a = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 6)
    return as_points(a[:, :3], a[:, 3:])


def stencil_1d():
    """{'M','S','G'}: 1-D integrals at d = -1, 0, 1 (src/fem_tree.rs:39-41)."""
    out = np.zeros(9)
    lib().orc_stencil_1d(_p(out))
    return {"M": out[0:3], "S": out[3:6], "G": out[6:9]}


def colour_order(depth):
    """Fine rows sorted by coordinate-parity colour (x&1)+2(y&1)+4(z&1)."""
    _, i2o = tree(depth)
    col = (i2o[:, 0] & 1) + 2 * (i2o[:, 1] & 1) + 4 * (i2o[:, 2] & 1)
    return np.argsort(col, kind="stable").astype(np.int64)


def set_threads(t: int) -> None:
    lib().orc_set_threads(int(t))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def unit_transform(points, scale=1.1):
    """(model_to_unit, unit_to_model) as column-major 4x4 (m[col*4+row])."""
    pts = points_array(points)
    m = np.zeros(16)
    mi = np.zeros(16)
    lib().orc_unit_transform(_p(pts), len(pts), float(scale), _p(m), _p(mi))
    return m, mi


def apply_transform(points, m):
    pts = points_array(points)
    out = np.empty_like(pts)
    lib().orc_apply_transform(_p(pts), len(pts), _p(np.ascontiguousarray(m, dtype=np.float64)), _p(out))
    return out


def tree(depth: int):
    """(grid_to_idx[res^3] int64, idx_to_offset[N,3] uint32) via the DFS walk."""
    res = 1 << depth
    g2i = np.empty(res ** 3, dtype=np.int64)
    i2o = np.empty((res ** 3, 3), dtype=np.uint32)
    n = lib().orc_tree(depth, _p(g2i), _p(i2o))
    assert n == res ** 3
    return g2i, i2o


def splat(depth, g2i, points):
    pts = points_array(points)
    n = 1 << (3 * depth)
    nf = np.empty((n, 3))
    w = np.empty(n)
    lib().orc_splat(depth, _p(g2i), _p(pts), len(pts), _p(nf), _p(w))
    return nf, w


class CSR:
    """Owning handle of an oracle CSR matrix (src/sparse.rs:24-33 layout)."""

    def __init__(self, handle):
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_csr_free(self.h)
            self.h = None

    @property
    def n(self):
        return int(lib().orc_csr_rows(self.h))

    @property
    def nnz(self):
        return int(lib().orc_csr_nnz(self.h))

    def arrays(self):
        off = np.empty(self.n + 1, dtype=np.int64)
        col = np.empty(self.nnz, dtype=np.int64)
        val = np.empty(self.nnz)
        lib().orc_csr_export(self.h, _p(off), _p(col), _p(val))
        return off, col, val

    def scipy(self):
        import scipy.sparse as sp
        off, col, val = self.arrays()
        return sp.csr_matrix((val, col, off), shape=(self.n, self.n))

    @staticmethod
    def from_arrays(off, col, val):
        off = np.ascontiguousarray(off, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int64)
        val = np.ascontiguousarray(val, dtype=np.float64)
        return CSR(lib().orc_csr_import(len(off) - 1, _p(off), _p(col), _p(val)))

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        lib().orc_spmv(self.h, _p(x), _p(y))
        return y

    def gs_sweep(self, b, x):
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert x.dtype == np.float64 and x.flags.c_contiguous
        lib().orc_gs_sweep(self.h, _p(b), _p(x))
        return x

    def gs_sweep_order(self, b, x, order):
        b = np.ascontiguousarray(b, dtype=np.float64)
        order = np.ascontiguousarray(order, dtype=np.int64)
        assert x.dtype == np.float64 and x.flags.c_contiguous and len(order) == self.n
        lib().orc_gs_sweep_order(self.h, _p(b), _p(x), _p(order))
        return x

    def cg(self, b, x, max_iters, eps):
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert x.dtype == np.float64 and x.flags.c_contiguous
        rel = np.zeros(1)
        it = lib().orc_cg(self.h, _p(b), _p(x), int(max_iters), float(eps), _p(rel))
        return int(it), float(rel[0])


def assemble(depth, g2i, i2o, points, pw):
    pts = points_array(points)
    return CSR(lib().orc_assemble(depth, _p(g2i), _p(i2o), _p(pts), len(pts), float(pw)))


def rhs(depth, g2i, i2o, nf, points, pw):
    pts = points_array(points)
    b = np.empty(1 << (3 * depth))
    lib().orc_rhs(depth, _p(g2i), _p(i2o), _p(np.ascontiguousarray(nf)), _p(pts), len(pts), float(pw), _p(b))
    return b


def coarse_system(depth, cdepth, i2o, nf, pw):
    b = np.empty(1 << (3 * cdepth))
    h = lib().orc_coarse_system(depth, cdepth, _p(i2o), _p(np.ascontiguousarray(nf)), float(pw), _p(b))
    return CSR(h), b


def prolongate(depth, cdepth, g2i, coarse):
    fine = np.empty(1 << (3 * depth))
    lib().orc_prolongate(depth, cdepth, _p(g2i), _p(np.ascontiguousarray(coarse, dtype=np.float64)), _p(fine))
    return fine


def evaluate(depth, g2i, sol, q):
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    out = np.empty(q.shape[0])
    lib().orc_evaluate(depth, _p(g2i), _p(np.ascontiguousarray(sol)), _p(q), q.shape[0], _p(out))
    return out


def iso(depth, g2i, sol, points):
    pts = points_array(points)
    return float(lib().orc_iso(depth, _p(g2i), _p(np.ascontiguousarray(sol)), _p(pts), len(pts)))


def solve_cascadic(depth, g2i, i2o, nf, points, pw, base=None, gs=20, cg_iters=500, eps=1e-3):
    """FEMTree::solve_cascadic with base = D-2 (src/reconstructor.rs:51-52)."""
    pts = points_array(points)
    if base is None:
        base = depth - 2 if depth > 2 else 0
    sol = np.empty(1 << (3 * depth))
    rep = _SolveReport()
    lib().orc_solve_cascadic(depth, _p(g2i), _p(i2o), _p(np.ascontiguousarray(nf)), _p(pts), len(pts), float(pw),
                             int(base), int(gs), int(cg_iters), float(eps), _p(sol), ctypes.addressof(rep))
    return sol, {k: getattr(rep, k) for k, _ in _SolveReport._fields_}


def field_at_points(depth, points, pw, eps=1e-13):
    """(F at the points, solution) after an FP64 CG solve of the assembled
    system from zero to ||r|| <= eps ||b||."""
    pts = points_array(points)
    g2i, i2o = tree(depth)
    x = np.empty(1 << (3 * depth))
    F = np.empty(len(pts))
    lib().orc_field_at_points(depth, _p(g2i), _p(i2o), _p(pts), len(pts), float(pw), float(eps), _p(x), _p(F))
    return F, x


def vjp_points(depth, points, pw, x, gF, eps=1e-13):
    """(d L/d positions [P,3], d L/d normals [P,3], lambda) for dL/dF = gF
    at the points, with x the converged solution (unit-cube coordinates)."""
    pts = points_array(points)
    g2i, i2o = tree(depth)
    gpos = np.empty((len(pts), 3))
    gnrm = np.empty((len(pts), 3))
    lam = np.empty(1 << (3 * depth))
    lib().orc_vjp_points(depth, _p(g2i), _p(i2o), _p(pts), len(pts), float(pw), _p(np.ascontiguousarray(x)),
                         _p(np.ascontiguousarray(gF, dtype=np.float64)), float(eps), _p(lam), _p(gpos), _p(gnrm))
    return gpos, gnrm, lam


def extract_surface(depth, g2i, i2o, sol, iso):
    """FEMTree::extract_surface (src/fem_tree.rs:576-636) in unit-cube
    coordinates: (vertices [nv,3], triangles [nt,3] int64)."""
    V = ctypes.POINTER(ctypes.c_double)()
    T = ctypes.POINTER(ctypes.c_int64)()
    nt = ctypes.c_int64(0)
    L = lib()
    nv = L.orc_extract_surface(depth, _p(g2i), _p(np.ascontiguousarray(i2o, dtype=np.uint32)), _p(np.ascontiguousarray(sol)),
                               float(iso), ctypes.byref(V), ctypes.byref(T), ctypes.byref(nt))
    verts = np.ctypeslib.as_array(V, shape=(max(nv, 1) * 3,))[:3 * nv].reshape(-1, 3).copy()
    tris = np.ctypeslib.as_array(T, shape=(max(nt.value, 1) * 3,))[:3 * nt.value].reshape(-1, 3).copy()
    L.orc_free(ctypes.cast(V, ctypes.c_void_p))
    L.orc_free(ctypes.cast(T, ctypes.c_void_p))
    return verts, tris


def reconstruct_field(points, depth=8, scale=1.1, cg_accuracy=1e-3, cg_iters=500, point_weight=0.0):
    """reconstruct() up to the iso value (src/reconstructor.rs:25-56): returns
    (unit points, solution, iso, report, model_to_unit)."""
    pts = points_array(points)
    m, _ = unit_transform(pts, scale)
    up = apply_transform(pts, m)
    g2i, i2o = tree(depth)
    nf, _w = splat(depth, g2i, up)
    sol, rep = solve_cascadic(depth, g2i, i2o, nf, up, point_weight, gs=20, cg_iters=cg_iters, eps=cg_accuracy)
    isov = iso(depth, g2i, sol, up)
    return up, sol, isov, rep, m
