"""The torch.autograd.Function with tensors on the GPU: points, field values
and gradients handed to libprx as device pointers (prx_splat_normal_field_device,
prx_evaluate_device, prx_vjp_points_device), against the same function on CPU
tensors (the host entry points, which the adjoint tests pin to the oracle and
to finite differences).

libprx shares torch's HIP runtime when torch is imported first, so the check
runs in a child process that imports torch before libprx; the rest of the
suite keeps the runtime libprx is linked against.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _child():
    import time

    import numpy as np
    import torch

    if not torch.cuda.is_available():
        raise SystemExit("torch sees no GPU")
    import poissonreconstruction_amd as pr
    from poissonreconstruction_amd import _lib as L
    from poissonreconstruction_amd import synthetic
    from poissonreconstruction_amd.autograd import field_at_points

    out = {}
    for name, D, n, seed in (("sphere", 5, 2000, 5), ("ellipsoids", 8, 200000, 7)):
        raw = synthetic.sphere(n, seed=seed) if name == "sphere" else synthetic.ellipsoids(n, seed=seed)
        unit, _, _ = L.unit_transform(raw, 1.1)
        q = np.stack([unit["px"], unit["py"], unit["pz"]], axis=1)
        nm = np.stack([unit["nx"], unit["ny"], unit["nz"]], axis=1)
        res = {}
        for dev in ("cpu", "cuda"):
            qt = torch.tensor(q, device=dev, requires_grad=True)
            nt = torch.tensor(nm, device=dev, requires_grad=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            F = field_at_points(qt, nt, D, 4.0, precision=pr.PRX_PREC_FP64)
            loss = ((F - F.mean()) ** 2).sum()
            loss.backward()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            assert F.device.type == dev and qt.grad.device.type == dev and nt.grad.device.type == dev
            res[dev] = (F.detach().cpu().numpy(), qt.grad.cpu().numpy(), nt.grad.cpu().numpy(), loss.item(), ms)
        (Fc, gqc, gnc, Lc, msc), (Fg, gqg, gng, Lg, msg) = res["cpu"], res["cuda"]
        rng = float(Fc.max() - Fc.min())
        out[name] = {
            "depth": D, "points": n,
            "field_rel": float(np.abs(Fg - Fc).max() / rng),
            "loss_rel": abs(Lg - Lc) / Lc,
            "grad_pos_rel": float(np.abs(gqg - gqc).max() / np.abs(gqc).max()),
            "grad_nrm_rel": float(np.abs(gng - gnc).max() / np.abs(gnc).max()),
            "ms_host_tensors": msc, "ms_device_tensors": msg,
        }
    print(json.dumps(out), flush=True)


def test_autograd_function_on_device_tensors(gpu):
    pytest.importorskip("torch")
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__)], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=600)
    print(r.stdout[-4000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    for name, c in res.items():
        # the same solve from the same inputs; only the order of atomic sums differs
        assert c["field_rel"] <= 1e-9 and c["loss_rel"] <= 1e-9, (name, c)
        assert c["grad_pos_rel"] <= 1e-7 and c["grad_nrm_rel"] <= 1e-7, (name, c)


if __name__ == "__main__":
    _child()
