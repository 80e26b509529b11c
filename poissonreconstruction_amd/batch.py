"""Many point clouds solved together on one device (prx_batch_* / prx_solve_batched).

Each instance is a full FEMTree of its own depth (6..10) over unit-cube points;
solve() runs the cascadic solve of every instance (src/fem_tree.rs:458-487,
base depth D-2) over one concatenated node array with graph-replayed CG
iterations.  tree(k) returns instance k as a FEMTree that the batch owns
(evaluate, solution, iso_value and vjp_points work on it).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from ._lib import check, ptr
from .fem_tree import FEMTree


class _BatchTree(FEMTree):
    """Instance k of a Batch; the handle belongs to the batch."""

    def __init__(self, batch: "Batch", handle, depth: int, points, pw: float):  # noqa: D401 - no prx_tree_create
        self._batch = batch
        self.max_depth = int(depth)
        self._h = handle
        self._points = points
        self._n_points = len(points)
        self._pw = float(pw)

    def close(self):
        self._h = None


class Batch:
    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(L.lib().prx_batch_create(int(device), ctypes.byref(h)))
        self._h = h
        self._meta = []

    def close(self):
        if getattr(self, "_h", None):
            L.lib().prx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return len(self._meta)

    def add(self, depth: int, unit_points, point_weight: float) -> int:
        """Splat and assemble one instance (points already in the unit cube)."""
        pts = L.as_points(unit_points)
        idx = ctypes.c_uint32(0)
        check(L.lib().prx_batch_add(self._h, int(depth), ptr(pts), len(pts), float(point_weight), ctypes.byref(idx)))
        self._meta.append((int(depth), pts, float(point_weight)))
        return int(idx.value)

    def solve(self, gs: int = 20, cg_iters: int = 800000, eps: float = 1e-6, residual_mode: int = L.PRX_RESID_TRUE,
              refresh_period: int = 10, autotune: bool = True, precond: int = L.PRX_PRECOND_NONE,
              precision: int = L.PRX_PREC_FP64, cg_two_pass: bool = False, cg_form: int = None,
              launch: int = L.PRX_BATCH_LAUNCH_AUTO):
        """Returns (batch report dict, list of per-instance report dicts).
        The batched cascadic solve runs in `precision` (FP64, or mixed with the
        fused CG step); precond PRX_PRECOND_MG solves every instance with
        multigrid-preconditioned CG (mixed or FP64) instead; cg_two_pass
        selects the two-pass CG iteration instead of the fused one (FP64).
        launch: PRX_BATCH_LAUNCH_PERSISTENT runs each level's CG of all
        instances in one persistent launch (true-residual stopping, the fused
        CG step with q stored; anything else raises), PRX_BATCH_LAUNCH_GRAPHS
        replays captured chunks of iterations, AUTO picks the graph replays
        (the faster of the two on many-cloud batches)."""
        o = L.default_solve_opts()
        o.precision = int(precision)
        o.residual_mode = int(residual_mode)
        o.refresh_period = int(refresh_period)
        o.precond = int(precond)
        o.cg_form = int(cg_form) if cg_form is not None else (L.PRX_CG_TWO_PASS if cg_two_pass else L.PRX_CG_FUSED)
        o.batch_launch = int(launch)
        rep = L.BatchReport()
        per = (L.SolveReport * max(len(self._meta), 1))()
        check(L.lib().prx_solve_batched(self._h, int(gs), int(cg_iters), float(eps), ctypes.byref(o), 1 if autotune else 0,
                                        ctypes.byref(rep), per))
        return rep.as_dict(), [per[k].as_dict() for k in range(len(self._meta))]

    def tree(self, k: int) -> FEMTree:
        h = ctypes.c_void_p()
        check(L.lib().prx_batch_tree(self._h, int(k), ctypes.byref(h)))
        d, pts, pw = self._meta[k]
        return _BatchTree(self, h, d, pts, pw)


def depth_for(n_points: int) -> int:
    """Depth rule of the batching workload: 6 below 20,000 points, 7 below
    100,000, else 8."""
    return 6 if n_points < 20000 else (7 if n_points < 100000 else 8)
