"""SparseMatrix (src/sparse.rs:24-117) and solve_cg (src/solvers.rs:20-69)
for caller-built CSR systems, executed on the GPU.

Row layout follows the reference: set_row_size per row, finalize_structure,
then fill rows with (col, value) entries.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from ._lib import check, ptr


class SparseMatrix:
    def __init__(self, rows: int):
        self.rows = int(rows)
        self.row_sizes = np.zeros(self.rows, dtype=np.int64)
        self.offsets = None
        self.cols = None
        self.vals = None

    def set_row_size(self, row: int, size: int) -> None:
        self.row_sizes[row] = size

    def finalize_structure(self) -> None:
        self.offsets = np.zeros(self.rows + 1, dtype=np.int64)
        np.cumsum(self.row_sizes, out=self.offsets[1:])
        nnz = int(self.offsets[-1])
        self.cols = np.zeros(nnz, dtype=np.int64)
        self.vals = np.zeros(nnz, dtype=np.float64)

    def set_row(self, row: int, entries) -> None:
        a, b = self.offsets[row], self.offsets[row + 1]
        if len(entries) != b - a:
            raise ValueError("row size mismatch")
        for k, (c, v) in enumerate(entries):
            self.cols[a + k] = c
            self.vals[a + k] = v

    def num_rows(self) -> int:
        return self.rows

    def nnz(self) -> int:
        return 0 if self.cols is None else len(self.cols)

    def multiply_vector(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.rows)
        check(L.lib().prx_csr_spmv(self.rows, ptr(self.offsets), ptr(self.cols), ptr(self.vals), ptr(x), ptr(y)))
        return y


def solve_cg(a: SparseMatrix, b, x, max_iters: int, accuracy: float) -> int:
    """Conjugate gradient in place on x; returns the iteration count."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
        raise TypeError("x must be a contiguous float64 array (updated in place)")
    it = ctypes.c_uint64(0)
    check(L.lib().prx_csr_cg(a.rows, ptr(a.offsets), ptr(a.cols), ptr(a.vals), ptr(b), ptr(x), int(max_iters),
                             float(accuracy), ctypes.byref(it)))
    return int(it.value)
