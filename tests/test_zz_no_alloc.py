"""prx_csr_cg_noalloc_device, the entry point that allocates nothing, against
prx_csr_cg (solve_cg, src/solvers.rs:20-69, on a caller CSR) through the C
ABI: the same kernels in the same order, so the same x bit for bit and the
same iteration count, on two calls in a row (the second smaller, so it runs on
scratch the first call left behind)."""
import ctypes

import numpy as np
import pytest

import oracle
from poissonreconstruction_amd import _lib as L
from poissonreconstruction_amd import synthetic

pytestmark = pytest.mark.gpu


def _system(D, seed):
    raw = synthetic.sphere(3000, seed=seed)
    m, _ = oracle.unit_transform(raw, 1.1)
    pts = oracle.apply_transform(raw, m)
    g2i, i2o = oracle.tree(D)
    nf, _ = oracle.splat(D, g2i, pts)
    A = oracle.assemble(D, g2i, i2o, pts, 4.0)
    off, col, val = A.arrays()
    b = oracle.rhs(D, g2i, i2o, nf, pts, 4.0)
    return (np.ascontiguousarray(off, dtype=np.int64), np.ascontiguousarray(col, dtype=np.int64),
            np.ascontiguousarray(val, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64), A)


def _with_host_buffers(off, col, val, b, max_iters, eps):
    n = len(b)
    x = np.zeros(n)
    it = ctypes.c_uint64(0)
    L.check(L.lib().prx_csr_cg(n, L.ptr(off), L.ptr(col), L.ptr(val), L.ptr(b), L.ptr(x), max_iters, eps,
                               ctypes.byref(it)))
    return x, it.value


class _Hip:
    """The test's own device buffers, through the HIP runtime libprx is linked
    against (hipMalloc here is the caller's; the entry point under test
    allocates nothing)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.live = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed: hipError {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self._ok(self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)), "hipMalloc")
        self.live.append(p)
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self._ok(self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def sync(self):
        self._ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _without_allocation(hip, off, col, val, b, max_iters, eps):
    doff, dcol, dval, db = (hip.upload(a) for a in (off, col, val, b))
    dx = hip.upload(np.zeros(len(b)))
    it = ctypes.c_uint64(0)
    hip.sync()
    print(f"prx_csr_cg_noalloc_device n={len(b)} nnz={len(col)}", flush=True)
    rc = L.lib().prx_csr_cg_noalloc_device(len(b), doff, dcol, dval, db, dx, max_iters, eps, ctypes.byref(it), None)
    hip.sync()
    L.check(rc)
    x = hip.download(dx, np.zeros(len(b)))
    hip.free()
    return x, it.value


def test_no_alloc_cg_equals_the_allocating_entry_point(gpu):
    eps, cap = 1e-10, 4000
    hip = _Hip()
    for D, seed in ((4, 1), (3, 2)):  # the second call is smaller
        off, col, val, b, A = _system(D, seed)
        x_ref, it_ref = _with_host_buffers(off, col, val, b, cap, eps)
        x, it = _without_allocation(hip, off, col, val, b, cap, eps)
        assert it == it_ref and 0 < it < cap
        assert np.array_equal(x.view(np.uint64), x_ref.view(np.uint64))
        # and it solves the system: the reference's stopping test is on ||r||/||r0||
        # with r0 = b (x0 = 0), so the true relative residual is near eps
        r = b - A.spmv(x)
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b)


def test_no_alloc_cg_rejects_sizes_beyond_its_scratch(gpu):
    hip = _Hip()
    d = hip.upload(np.zeros(4))
    n = L.PRX_CSR_CG_NOALLOC_MAX_ROWS + 1
    rc = L.lib().prx_csr_cg_noalloc_device(n, d, d, d, d, d, 10, 1e-6, None, None)
    hip.free()
    assert rc == L.PRX_ERR_INVALID
    assert "PRX_CSR_CG_NOALLOC_MAX_ROWS" in L.last_error()
