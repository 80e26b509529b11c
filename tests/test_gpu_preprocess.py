"""Point preprocessing on the device (bounding box, unit-cube map, normal
normalisation; src/reconstructor.rs:31-40, src/geometry.rs:256-269,349-402)
against the oracle's restatement and the library's host path, bit for bit."""
import ctypes

import numpy as np
import pytest

import oracle
import poissonreconstruction_amd as pr
from poissonreconstruction_amd import _lib as L
from poissonreconstruction_amd import synthetic

pytestmark = pytest.mark.gpu


def both(raw, scale=1.1):
    pts = L.as_points(raw)
    n = len(pts)
    out_h, out_d = np.empty_like(pts), np.empty_like(pts)
    mh, mih, md, mid = (np.zeros(16) for _ in range(4))
    L.check(L.lib().prx_unit_transform(L.ptr(pts), n, scale, L.ptr(mh), L.ptr(mih), L.ptr(out_h)))
    L.check(L.lib().prx_unit_transform_device(L.ptr(pts), n, scale, 0, L.ptr(md), L.ptr(mid), L.ptr(out_d)))
    return (out_h, mh, mih), (out_d, md, mid)


@pytest.mark.parametrize("cloud", ["torus", "ellipsoids", "flat"])
def test_device_transform_is_bit_identical(gpu, cloud):
    if cloud == "torus":
        raw = synthetic.torus(300000, seed=1)
    elif cloud == "ellipsoids":
        raw = synthetic.ellipsoids(50000, seed=2) * np.array([1e3, 1e3, 1e3, 1, 1, 1]) + np.array([5e4, -3, 7, 0, 0, 0])
    else:  # a degenerate axis (zero extent in z) and unnormalised normals
        raw = synthetic.sphere(20000, seed=3)
        raw[:, 2] = 0.25
        raw[:, 3:] *= 3.0
    (oh, mh, mih), (od, md, mid) = both(raw)
    # the oracle (unit_cube_transform, Mat4::inverse, transform_point/vector, normalize)
    mo, mio = oracle.unit_transform(raw, 1.1)
    po = oracle.apply_transform(raw, mo)
    assert np.array_equal(np.asarray(mo).view(np.uint64), md.view(np.uint64))
    assert np.array_equal(np.asarray(mio).view(np.uint64), mid.view(np.uint64))
    for f in ("px", "py", "pz", "nx", "ny", "nz"):
        assert np.array_equal(np.ascontiguousarray(po[f]).view(np.uint64), np.ascontiguousarray(od[f]).view(np.uint64)), f
    assert np.array_equal(mh.view(np.uint64), md.view(np.uint64))
    assert np.array_equal(mih.view(np.uint64), mid.view(np.uint64))
    assert np.array_equal(oh.view(np.uint64), od.view(np.uint64))
