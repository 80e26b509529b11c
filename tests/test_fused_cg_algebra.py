"""The fused CG iteration of the GPU solver, restated in FP64 numpy on an
assembled system (the oracle's CSR), against the reference's two-reduction
iteration (src/solvers.rs:36-64).

The fused form merges the iteration's two reductions: one pass forms the
previous update r' = r - alpha q, the direction d' = r' + beta d and q' = A d',
and sums d'.q', r'.q', q'.q' and r'.r'; one scalar step takes delta = r'.r',
alpha = delta / d'.q' and the next |r|^2 from the expansion
delta - 2 alpha r'.q' + alpha^2 q'.q' (for beta and the stopping test).  In
exact arithmetic it is the reference's iteration.  These tests pin that
algebra on the host: the same iterates and iteration counts as the reference
form, and the reason every pass sums r'.r' again: carried from iteration to
iteration, the expansion's rounding accumulates and the solve cannot reach a
tight tolerance."""
import numpy as np
import pytest

import oracle
from poissonreconstruction_amd import synthetic


@pytest.fixture(scope="module")
def system():
    D = 4
    raw = synthetic.torus(4000, seed=2)
    m, _ = oracle.unit_transform(raw, 1.1)
    pts = oracle.apply_transform(raw, m)
    g2i, i2o = oracle.tree(D)
    A = oracle.assemble(D, g2i, i2o, pts, 4.0).scipy().tocsr()
    nf, _ = oracle.splat(D, g2i, pts)
    b = oracle.rhs(D, g2i, i2o, nf, pts, 4.0)
    return A, b


def cg_reference(A, b, eps, max_iters):
    """src/solvers.rs:36-64 without the periodic refresh: x0 = 0, stop when
    |r|^2 <= |r0|^2 eps^2."""
    x = np.zeros_like(b)
    r = b.copy()
    d = r.copy()
    delta = r @ r
    thr = delta * eps * eps
    xs = []
    for it in range(max_iters):
        q = A @ d
        alpha = delta / (d @ q)
        x = x + alpha * d
        r = r - alpha * q
        new = r @ r
        xs.append(x.copy())
        if new <= thr:
            return x, it + 1, xs
        beta = new / delta
        delta = new
        d = r + beta * d
    return x, max_iters, xs


def cg_fused(A, b, eps, max_iters, reset=True):
    """The GPU's fused iteration: pass k forms r_k (the pending update), d_k and
    q_k and sums d.q, r.q, q.q, r.r; the fold sets delta = r.r (reset) or keeps
    the carried expansion (reset=False), alpha, the expanded next |r|^2 and beta."""
    x = np.zeros_like(b)
    r = b.copy()
    q = np.zeros_like(b)
    d = np.zeros_like(b)
    alpha = beta = 0.0
    pending = False
    delta = r @ r
    thr = delta * eps * eps
    xs = []
    for it in range(max_iters):
        if pending:
            r = r - alpha * q
        d = r + beta * d
        q = A @ d
        dq, rq, qq, rr = d @ q, r @ q, q @ q, r @ r
        if reset:
            delta = rr
        alpha = delta / dq
        x = x + alpha * d
        pending = True
        g = (delta - 2.0 * alpha * rq) + alpha * alpha * qq
        xs.append(x.copy())
        if g <= thr:
            return x, it + 1, xs
        beta = g / delta
        delta = g
    return x, max_iters, xs


def test_fused_iterates_equal_the_reference_iterates(system):
    A, b = system
    xr, itr, xsr = cg_reference(A, b, 1e-10, 5000)
    xf, itf, xsf = cg_fused(A, b, 1e-10, 5000)
    assert abs(itr - itf) <= 1
    scale = np.abs(xr).max()
    # every iterate agrees to round-off (the two forms differ only in how |r|^2 is rounded)
    for k in range(min(len(xsr), len(xsf))):
        assert np.abs(xsr[k] - xsf[k]).max() <= 1e-9 * scale, k
    assert np.linalg.norm(b - A @ xf) <= 1.01e-10 * np.linalg.norm(b)


def test_fused_matches_a_direct_solve(system):
    import scipy.sparse.linalg as sla
    A, b = system
    xd = sla.spsolve(A.tocsc(), b)
    xf, _, _ = cg_fused(A, b, 1e-12, 5000)
    assert np.abs(xf - xd).max() <= 1e-8 * np.abs(xd).max()


def test_carried_expansion_without_the_reset_fails_a_tight_tolerance(system):
    # the design point the fused pass's r'.r' sum exists for: without it the expansion's absolute error
    # (about eps_machine times the early |r|^2) swamps |r|^2 late in the solve
    A, b = system
    _, it_ok, _ = cg_fused(A, b, 1e-13, 5000, reset=True)
    with np.errstate(over="ignore", invalid="ignore"):
        xb, it_bad, _ = cg_fused(A, b, 1e-13, 5000, reset=False)
        true_bad = np.linalg.norm(b - A @ xb) / np.linalg.norm(b)
    assert it_ok < 5000
    assert it_bad == 5000 or true_bad > 1e-11 or not np.isfinite(true_bad)
