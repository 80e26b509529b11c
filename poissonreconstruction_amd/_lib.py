"""ctypes binding of libprx.so (the C ABI in include/prx.h).

The library is loaded from this package directory.  There is no fallback:
if libprx.so is missing or no GPU is visible, calls fail with PrxError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprx.so")

PRX_OK = 0
PRX_ERR_INVALID = 1
PRX_ERR_NONFINITE = 2
PRX_ERR_NOT_CONVERGED = 3
PRX_ERR_BREAKDOWN = 4
PRX_ERR_DEVICE = 5
PRX_ERR_OOM = 6
PRX_ERR_STATE = 7
PRX_CSR_CG_NOALLOC_MAX_ROWS = 262144  # include/prx.h
PRX_PREC_FP64 = 0
PRX_PREC_MIXED = 1
PRX_RESID_REFERENCE = 0
PRX_RESID_TRUE = 1
PRX_PRECOND_NONE = 0
PRX_PRECOND_MG = 1
PRX_CG_FUSED = 0
PRX_CG_TWO_PASS = 1
PRX_CG_FUSED_STORED_Q = 2
PRX_CG_FUSED_FACES = 3
PRX_BATCH_LAUNCH_AUTO = 0
PRX_BATCH_LAUNCH_GRAPHS = 1
PRX_BATCH_LAUNCH_PERSISTENT = 2
PRX_MAX_DEPTH = 10

STATUS_NAMES = {0: "ok", 1: "invalid", 2: "nonfinite", 3: "not_converged", 4: "breakdown", 5: "device", 6: "oom", 7: "state"}

POINT_DTYPE = np.dtype([("px", "f8"), ("py", "f8"), ("pz", "f8"), ("nx", "f8"), ("ny", "f8"), ("nz", "f8")])


class PrxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libprx error {code} ({STATUS_NAMES.get(code, '?')}): {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_uint32), ("samples_per_node", ctypes.c_double), ("scale", ctypes.c_double),
                ("cg_accuracy", ctypes.c_double), ("cg_iters", ctypes.c_uint64), ("point_weight", ctypes.c_double),
                ("verbose", ctypes.c_int32), ("precision", ctypes.c_int32), ("residual_mode", ctypes.c_int32),
                ("device", ctypes.c_int32), ("gs_iters", ctypes.c_int32), ("base_depth", ctypes.c_int32),
                ("refresh_period", ctypes.c_int32), ("inner_tol", ctypes.c_double), ("max_outer", ctypes.c_uint64),
                ("precond", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("time_transform", ctypes.c_double), ("time_tree", ctypes.c_double), ("time_system", ctypes.c_double),
                ("time_solve", ctypes.c_double), ("time_extract", ctypes.c_double), ("time_total", ctypes.c_double),
                ("fem_nodes", ctypes.c_uint64), ("matrix_nnz", ctypes.c_uint64), ("mesh_vertices", ctypes.c_uint64),
                ("mesh_triangles", ctypes.c_uint64), ("iso_value", ctypes.c_double), ("coarse_iters", ctypes.c_uint64),
                ("fine_iters", ctypes.c_uint64), ("outer_iters", ctypes.c_uint64), ("final_rel_residual", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("solve_status", ctypes.c_int32)]


class Mesh(ctypes.Structure):
    _fields_ = [("vertices", ctypes.POINTER(ctypes.c_double)), ("n_vertices", ctypes.c_uint64),
                ("triangles", ctypes.POINTER(ctypes.c_uint64)), ("n_triangles", ctypes.c_uint64)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("residual_mode", ctypes.c_int32), ("refresh_period", ctypes.c_int32),
                ("gs_iters", ctypes.c_int32), ("inner_tol", ctypes.c_double), ("max_outer", ctypes.c_uint64),
                ("coarse_eps", ctypes.c_double), ("precond", ctypes.c_int32), ("mg_smooth", ctypes.c_int32),
                ("mg_omega", ctypes.c_double), ("mg_inner_tol", ctypes.c_double), ("cg_form", ctypes.c_int32),
                ("batch_launch", ctypes.c_int32)]


class SolveReport(ctypes.Structure):
    _fields_ = [("coarse_iters", ctypes.c_uint64), ("fine_iters", ctypes.c_uint64), ("outer_iters", ctypes.c_uint64),
                ("applies", ctypes.c_uint64), ("coarse_rel", ctypes.c_double), ("fine_rel", ctypes.c_double),
                ("true_rel", ctypes.c_double), ("ms_assemble", ctypes.c_double), ("ms_coarse", ctypes.c_double),
                ("ms_fine_gs", ctypes.c_double), ("ms_fine_cg", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("status", ctypes.c_int32), ("fine_r0_rel", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class VjpReport(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_uint64), ("rel", ctypes.c_double), ("ms", ctypes.c_double), ("status", ctypes.c_int32)]


class BatchReport(ctypes.Structure):
    _fields_ = [("instances", ctypes.c_uint32), ("converged", ctypes.c_uint32), ("total_nodes", ctypes.c_uint64),
                ("total_points", ctypes.c_uint64), ("max_fine_iters", ctypes.c_uint64), ("sum_fine_iters", ctypes.c_uint64),
                ("max_coarse_iters", ctypes.c_uint64), ("graph_launches", ctypes.c_uint64), ("ms_total", ctypes.c_double),
                ("ms_coarse", ctypes.c_double), ("ms_fine_gs", ctypes.c_double), ("ms_fine_cg", ctypes.c_double),
                ("ms_tune", ctypes.c_double), ("max_true_rel", ctypes.c_double), ("tile", ctypes.c_int32),
                ("chunk", ctypes.c_int32), ("gs_block", ctypes.c_int32), ("status", ctypes.c_int32),
                ("persistent_launches", ctypes.c_uint64), ("persistent_rounds", ctypes.c_uint64),
                ("persistent_blocks", ctypes.c_int32), ("persistent_xcd_match", ctypes.c_int32),
                ("persistent_phase_ms", ctypes.c_double * 4)]

    def as_dict(self):
        return {k: (list(getattr(self, k)) if isinstance(getattr(self, k), ctypes.Array) else getattr(self, k))
                for k, _ in self._fields_}


_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_sz = ctypes.c_size_t

# name -> (restype, argtypes); every function of include/prx.h
SIGNATURES = {
    "prx_last_error": (ctypes.c_char_p, []),
    "prx_device_count": (_i32, [_vp]),
    "prx_params_default": (None, [_vp]),
    "prx_solve_opts_default": (None, [_vp]),
    "prx_reconstruct": (_i32, [_vp, _sz, _vp, _vp, _vp]),
    "prx_mesh_free": (None, [_vp]),
    "prx_tree_create": (_i32, [_u32, _i32, _vp]),
    "prx_tree_destroy": (None, [_vp]),
    "prx_tree_node_count": (_i64, [_vp]),
    "prx_tree_export_nodes": (_i32, [_vp, _vp, _vp]),
    "prx_splat_normal_field": (_i32, [_vp, _vp, _sz]),
    "prx_export_normal_field": (_i32, [_vp, _vp, _vp]),
    "prx_assemble": (_i32, [_vp, _f64]),
    "prx_export_rhs": (_i32, [_vp, _vp]),
    "prx_csr_nnz": (_i32, [_vp, _vp]),
    "prx_export_csr": (_i32, [_vp, _vp, _vp, _vp]),
    "prx_apply": (_i32, [_vp, _i32, _vp, _vp]),
    "prx_mg_vcycle": (_i32, [_vp, _i32, _i32, _f64, _vp, _vp]),
    "prx_coarse_system": (_i32, [_vp, _u32, _vp, _vp]),
    "prx_prolongate": (_i32, [_vp, _u32, _vp, _vp]),
    "prx_gs": (_i32, [_vp, _i32, _u32, _vp, _vp]),
    "prx_cg": (_i32, [_vp, _i32, _vp, _vp, _u64, _f64, _i32, _i32, _vp, _vp]),
    "prx_solve_cascadic": (_i32, [_vp, _u32, _u32, _u64, _f64, _vp, _vp]),
    "prx_get_solution": (_i32, [_vp, _vp]),
    "prx_set_solution": (_i32, [_vp, _vp]),
    "prx_evaluate": (_i32, [_vp, _vp, _sz, _vp]),
    "prx_evaluate_corners": (_i32, [_vp, _vp]),
    "prx_iso_value": (_i32, [_vp, _vp]),
    "prx_unit_transform": (_i32, [_vp, _sz, _f64, _vp, _vp, _vp]),
    "prx_time_operator": (_i32, [_vp, _i32, _i32, _vp, _vp]),
    "prx_time_stencil_op": (_i32, [_vp, _i32, _i32, _i32, _vp, _vp]),
    "prx_copy_bandwidth": (_i32, [_sz, _i32, _i32, _vp]),
    "prx_splat_times": (_i32, [_vp, _vp]),
    "prx_pass_timing": (_i32, [_vp, _i32, _vp, ctypes.c_size_t, _vp]),
    "prx_tree_release_workspace": (_i32, [_vp]),
    "prx_screening_sizes": (_i32, [_vp, _vp]),
    "prx_splat_normal_field_device": (_i32, [_vp, _vp, _vp, _sz]),
    "prx_evaluate_device": (_i32, [_vp, _vp, _sz, _vp]),
    "prx_vjp_points_device": (_i32, [_vp, _vp, _vp, _vp, _i32, _f64, _vp]),
    "prx_brick_copy_bandwidth": (_i32, [_i64, _i32, _i32, _i32, _i32, _vp]),
    "prx_time_evaluate_corners": (_i32, [_vp, _i32, _i32, _vp, _vp]),
    "prx_extract_surface": (_i32, [_vp, _f64, _vp]),
    "prx_unit_transform_device": (_i32, [_vp, _sz, _f64, _i32, _vp, _vp, _vp]),
    "prx_vjp_points": (_i32, [_vp, _vp, _vp, _vp, _i32, _f64, _vp]),
    "prx_batch_create": (_i32, [_i32, _vp]),
    "prx_batch_destroy": (None, [_vp]),
    "prx_batch_add": (_i32, [_vp, _u32, _vp, _sz, _f64, _vp]),
    "prx_batch_size": (_i32, [_vp, _vp]),
    "prx_batch_tree": (_i32, [_vp, _u32, _vp]),
    "prx_solve_batched": (_i32, [_vp, _u32, _u64, _f64, _vp, _i32, _vp, _vp]),
    "prx_csr_spmv": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "prx_csr_cg": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _u64, _f64, _vp]),
    "prx_csr_cg_noalloc_device": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _u64, _f64, _vp, _vp]),
}

_lib = None


def lib():
    """Load libprx.so (built in-tree by build.py); raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PrxError(PRX_ERR_DEVICE, f"{LIB_PATH} is not built; run poissonreconstruction_amd.build.build_lib()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().prx_last_error().decode(errors="replace")


def check(rc: int) -> None:
    if rc != PRX_OK:
        raise PrxError(rc, last_error())


def ptr(a):
    if a is None:
        return None
    if isinstance(a, ctypes.Structure):
        return ctypes.addressof(a)
    return a.ctypes.data_as(_vp)


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().prx_device_count(ctypes.byref(n)))
    return n.value


def default_params() -> Params:
    p = Params()
    lib().prx_params_default(ctypes.byref(p))
    return p


def default_solve_opts() -> SolveOpts:
    o = SolveOpts()
    lib().prx_solve_opts_default(ctypes.byref(o))
    return o


def unit_transform(points, scale: float = 1.1):
    """(unit points, model_to_unit, unit_to_model): the point transform at
    src/reconstructor.rs:31-40 (matrices column-major, m[col*4+row])."""
    pts = as_points(points)
    m = np.zeros(16)
    mi = np.zeros(16)
    out = np.empty_like(pts)
    check(lib().prx_unit_transform(ptr(pts), len(pts), float(scale), ptr(m), ptr(mi), ptr(out)))
    return out, m, mi


def as_points(points) -> np.ndarray:
    """Pack points into the prx_point layout (px,py,pz,nx,ny,nz doubles)."""
    if isinstance(points, np.ndarray) and points.dtype == POINT_DTYPE:
        return np.ascontiguousarray(points)
    if isinstance(points, (tuple, list)) and len(points) == 2 and not np.isscalar(points[0]):
        pos = np.asarray(points[0], dtype=np.float64).reshape(-1, 3)
        nrm = np.asarray(points[1], dtype=np.float64).reshape(-1, 3)
        a = np.concatenate([pos, nrm], axis=1)
    else:
        a = np.asarray(points, dtype=np.float64).reshape(-1, 6)
    return np.ascontiguousarray(a).view(POINT_DTYPE).reshape(-1)
