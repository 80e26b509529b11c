"""Command line of the reference (src/main.rs:6-50) on the GPU path:

    python -m poissonreconstruction_amd -i cloud.ply -o mesh.ply [-d 8] [--samples-per-node 1.5]
        [--cg-accuracy 0.001] [--cg-iters 500] [--point-weight 0.0] [--scale 1.1] [-v]

plus --precision {fp64,mixed}, --residual {reference,true} and --device N.
Errors print "Error: ..." and exit 1; an empty cloud prints "No points found."
and exits 1, as the reference does.
"""
from __future__ import annotations

import argparse
import sys


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="poisson-recon", description="Screened Poisson Surface Reconstruction")
    ap.add_argument("-i", "--input", required=True)
    ap.add_argument("-o", "--output", default="output.ply")
    ap.add_argument("-d", "--depth", type=int, default=8)
    ap.add_argument("--samples-per-node", type=float, default=1.5)
    ap.add_argument("--cg-accuracy", type=float, default=0.001)
    ap.add_argument("--cg-iters", type=int, default=500)
    ap.add_argument("--point-weight", type=float, default=0.0)
    ap.add_argument("--scale", type=float, default=1.1)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--residual", choices=("reference", "true"), default="reference")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--multigrid", action="store_true",
                    help="solve with multigrid-preconditioned CG instead of the cascadic solve (an extension)")
    a = ap.parse_args(argv)

    from . import ply
    try:
        pts = ply.read_oriented_points(a.input)
    except ValueError as e:
        print(f"Error: {e}", file=sys.stderr)
        return 1
    if len(pts) == 0:
        print("No points found.", file=sys.stderr)
        return 1
    if a.verbose:
        print(f"Read {len(pts)} oriented points", file=sys.stderr)

    from . import _lib as L
    from .reconstructor import ReconstructParams, reconstruct
    p = ReconstructParams(depth=a.depth, samples_per_node=a.samples_per_node, scale=a.scale, cg_accuracy=a.cg_accuracy,
                          cg_iters=a.cg_iters, point_weight=a.point_weight, verbose=a.verbose,
                          precision=L.PRX_PREC_MIXED if a.precision == "mixed" else L.PRX_PREC_FP64,
                          residual_mode=L.PRX_RESID_TRUE if a.residual == "true" else L.PRX_RESID_REFERENCE,
                          device=a.device, precond=L.PRX_PRECOND_MG if a.multigrid else L.PRX_PRECOND_NONE)
    mesh, s = reconstruct(pts, p)
    if a.verbose:
        print(f"Timing [s]: xform={s.time_transform:.3f} tree={s.time_tree:.3f} sys={s.time_system:.3f} "
              f"solve={s.time_solve:.3f} extract={s.time_extract:.3f} total={s.time_total:.3f}", file=sys.stderr)
        print(f"FEM nodes={s.fem_nodes}, matrix NNZ={s.matrix_nnz}", file=sys.stderr)
        print(f"Mesh: {s.mesh_vertices} verts, {s.mesh_triangles} tris, iso={s.iso_value:.6f}", file=sys.stderr)
    try:
        ply.write_mesh(a.output, mesh.vertices, mesh.triangles)
    except ValueError as e:
        print(f"Error writing: {e}", file=sys.stderr)
        return 1
    if a.verbose:
        print(f"Wrote {a.output}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
