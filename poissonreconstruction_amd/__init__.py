"""MI355X-native screened-Poisson solve path (drop-in for the reference's
reconstruct() API, src/reconstructor.rs), built on libprx.so: hand-written
gfx950 HIP kernels behind the C ABI in include/prx.h."""
from ._lib import (PRX_BATCH_LAUNCH_AUTO, PRX_BATCH_LAUNCH_GRAPHS, PRX_BATCH_LAUNCH_PERSISTENT, PRX_CG_FUSED, PRX_CG_FUSED_FACES,
                   PRX_CG_FUSED_STORED_Q, PRX_CG_TWO_PASS, PRX_OK, PRX_PRECOND_MG, PRX_PRECOND_NONE, PRX_PREC_FP64, PRX_PREC_MIXED,
                   PRX_RESID_REFERENCE, PRX_RESID_TRUE, PrxError, as_points, device_count)
from .batch import Batch
from .fem_tree import FEMTree, MatrixFreeSystem
from .reconstructor import Mesh, ReconstructParams, ReconstructStats, reconstruct
from .sparse import SparseMatrix, solve_cg

__all__ = [
    "Batch", "FEMTree", "MatrixFreeSystem", "Mesh", "ReconstructParams", "ReconstructStats", "reconstruct", "SparseMatrix",
    "solve_cg", "PrxError", "as_points", "device_count", "PRX_PREC_FP64", "PRX_PREC_MIXED", "PRX_RESID_REFERENCE",
    "PRX_RESID_TRUE", "PRX_OK", "PRX_PRECOND_NONE", "PRX_PRECOND_MG", "PRX_CG_FUSED", "PRX_CG_TWO_PASS",
    "PRX_CG_FUSED_STORED_Q", "PRX_CG_FUSED_FACES", "PRX_BATCH_LAUNCH_AUTO", "PRX_BATCH_LAUNCH_GRAPHS",
    "PRX_BATCH_LAUNCH_PERSISTENT",
]
