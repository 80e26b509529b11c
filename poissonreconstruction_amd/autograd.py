"""torch.autograd.Function over libprx's C ABI: the screened-Poisson field at
the input points as a differentiable function of the points' positions and
normals (unit-cube coordinates).

forward: splat -> assemble -> cascadic solve to a tight FP64 true residual ->
F_p = evaluate(x, q_p) (src/fem_tree.rs:130-513).  backward: one adjoint solve
with the same (symmetric) operator plus the point-local terms
(prx_vjp_points).  The forward solution stays on the device in the FEMTree
kept by the autograd context (the checkpoint); backward reuses it.

Tensors on a GPU are handed to libprx as device pointers (the *_device entry
points): no host copies of points, field values or gradients.  libprx then
shares torch's HIP runtime, which requires torch to be imported before libprx
is first loaded (importing this module first does that).  CPU tensors go
through the host entry points.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .fem_tree import FEMTree


def _device_f64(t):
    """t as a contiguous float64 tensor on its GPU, its producer's work done."""
    t = t.detach().to(torch.float64).contiguous()
    torch.cuda.current_stream(t.device).synchronize()
    return t


class FieldAtPoints(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, normals, depth: int, point_weight: float, precision: int = L.PRX_PREC_MIXED,
                eps: float = 1e-12, adjoint_eps: float = 1e-10):
        if positions.shape != normals.shape or positions.dim() != 2 or positions.shape[1] != 3:
            raise ValueError("positions and normals must both be (n, 3)")
        on_gpu = positions.is_cuda
        if on_gpu and (not normals.is_cuda or normals.device != positions.device):
            raise ValueError("positions and normals must be on the same device")
        base = depth - 2 if depth > 2 else 0
        n = positions.shape[0]
        if on_gpu:
            tree = FEMTree(depth, device=positions.device.index or 0)
            pos, nrm = _device_f64(positions), _device_f64(normals)
            tree.splat_normal_field_device(pos.data_ptr(), nrm.data_ptr(), n)
            rep = tree.solve_cascadic(None, point_weight, base, 20, 4 * 200000, eps, precision=precision,
                                      residual_mode=L.PRX_RESID_TRUE, refresh_period=0)
            F = torch.empty(n, dtype=torch.float64, device=positions.device)
            tree.evaluate_device(pos.data_ptr(), n, F.data_ptr())
        else:
            pos = positions.detach().to(torch.float64).numpy()
            nrm = normals.detach().to(torch.float64).numpy()
            pts = np.ascontiguousarray(np.concatenate([pos, nrm], axis=1))
            tree = FEMTree(depth)
            tree.splat_normal_field(pts)
            rep = tree.solve_cascadic(pts, point_weight, base, 20, 4 * 200000, eps, precision=precision,
                                      residual_mode=L.PRX_RESID_TRUE, refresh_period=0)
            F = torch.from_numpy(tree.evaluate_many(pos))
        ctx.tree = tree
        ctx.opts = (precision, adjoint_eps, positions.dtype, on_gpu)
        ctx.solve_report = rep
        return F.to(positions.dtype)

    @staticmethod
    def backward(ctx, grad_F):
        precision, adjoint_eps, dt, on_gpu = ctx.opts
        if on_gpu:
            g = _device_f64(grad_F)
            P = g.shape[0]
            gpos = torch.empty((P, 3), dtype=torch.float64, device=g.device)
            gnrm = torch.empty((P, 3), dtype=torch.float64, device=g.device)
            ctx.adjoint_report = ctx.tree.vjp_points_device(g.data_ptr(), gpos.data_ptr(), gnrm.data_ptr(),
                                                            precision=precision, eps=adjoint_eps)
        else:
            gp, gn, ctx.adjoint_report = ctx.tree.vjp_points(grad_F.detach().to("cpu", torch.float64).numpy(),
                                                             precision=precision, eps=adjoint_eps)
            gpos, gnrm = torch.from_numpy(gp), torch.from_numpy(gn)
        return gpos.to(dt), gnrm.to(dt), None, None, None, None, None


def field_at_points(positions, normals, depth: int, point_weight: float, precision: int = L.PRX_PREC_MIXED,
                    eps: float = 1e-12, adjoint_eps: float = 1e-10):
    """F at the points (unit-cube coordinates), differentiable in positions and normals."""
    return FieldAtPoints.apply(positions, normals, depth, point_weight, precision, eps, adjoint_eps)


def unit_cube_frozen(points_model, scale: float = 1.1):
    """(scale s, centre c) of unit_cube_transform (src/geometry.rs:384-402) for a
    cloud, to hold the transform fixed while differentiating."""
    p = np.asarray(points_model, dtype=np.float64).reshape(-1, 6)[:, :3]
    mn, mx = p.min(axis=0), p.max(axis=0)
    ext = float((mx - mn).max())
    return (scale / ext if ext > 0 else 1.0), (mn + mx) * 0.5


def iso_loss(positions_model, normals_model, frozen, depth: int, point_weight: float, **kw):
    """L = sum_p (F(p) - iso)^2 with iso the mean of F (src/reconstructor.rs:56),
    for model-space points under a frozen unit-cube transform; the transform
    and the normal normalisation are torch ops, so autograd chains them."""
    s, c = frozen
    c = torch.as_tensor(c, dtype=positions_model.dtype, device=positions_model.device)
    q = 0.5 + s * (positions_model - c)
    n = normals_model / torch.linalg.norm(normals_model, dim=1, keepdim=True)
    F = field_at_points(q, n, depth, point_weight, **kw)
    return ((F - F.mean()) ** 2).sum()
