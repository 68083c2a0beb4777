"""The largest configuration, one map at nside 2048 (50.3 M pixels), J = 10, L = 4, checked
through the same size-independent properties as nside 1024 (tests/test_gpu_nside1024.py):

- FP64 engine: the VJP against a central difference of the loss sum |S(x) - S(t)|^2 along a
  large-scale direction (eps 1e-4 of the map's scale): relative error at most 1e-6.
- Mixed engine against the FP64 engine: every statistic family within 1e-6 (max |dS| / max |S|
  per family), the gradient within 1e-5.
- Mixed engine: a step along the negative gradient sized to lower the loss by 0.1% to first
  order lowers it by that amount within 5%.

One engine is alive at a time (the FP64 one holds about 150 GB of HBM).
"""
import numpy as np
import pytest

from test_gpu_nside1024 import family_errors, lognormal, loss_and_grad

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NSIDE, J = 2048, 10


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def maps():
    x = torch.from_numpy(lognormal(1, nside=NSIDE)).cuda().reshape(1, -1)
    t = torch.from_numpy(lognormal(2, nside=NSIDE)).cuda().reshape(1, -1)
    return x, t


def test_nside2048_fp64_central_difference_and_mixed_against_fp64(maps):
    from stl_port import NativeEngine
    x, t = maps
    print(f"[gpu] nside={NSIDE} J={J} fp64 central difference", flush=True)
    with NativeEngine(NSIDE, J, precision="fp64") as e64:
        target, _ = e64.forward_device(t, generation=1, objective="target")
        _, g64, v64 = loss_and_grad(e64, x, target, 2)
        rng = np.random.default_rng(5)
        u = np.repeat(rng.standard_normal(12 * 32 * 32), (NSIDE // 32) ** 2)
        u = torch.from_numpy(u).cuda().reshape(1, -1)
        u *= x.abs().mean() / u.abs().mean()
        eps = 1e-4
        lp, _, _ = loss_and_grad(e64, x + eps * u, target, 3)
        lm, _, _ = loss_and_grad(e64, x - eps * u, target, 4)
        fd = (lp - lm) / (2 * eps)
        an = float((g64 * u).sum())
        print(f"[gpu] directional derivative fd {fd:.12e} vjp {an:.12e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an)
        plan = e64.plan
        g64 = g64.cpu().numpy()
    torch.cuda.empty_cache()
    print(f"[gpu] nside={NSIDE} J={J} mixed vs fp64", flush=True)
    with NativeEngine(NSIDE, J, precision="mixed") as em:
        _, gm, vm = loss_and_grad(em, x, target, 2)
        errs = family_errors(vm, v64, plan)
        print(f"[gpu] mixed vs fp64 {errs}", flush=True)
        assert max(errs.values()) < 1e-6
        assert float(np.abs(gm.cpu().numpy() - g64).max() / np.abs(g64).max()) < 1e-5
        gg = float((gm * gm).sum())
        l0 = float((np.abs(vm - target) ** 2).sum())
        step = 1e-3 * l0 / gg
        l1, _, _ = loss_and_grad(em, x - step * gm, target, 3)
        slope = -step * gg
        print(f"[gpu] loss {l0:.6e} -> {l1:.6e}, first-order change {slope:.6e}", flush=True)
        assert l1 < l0
        assert abs((l1 - l0) - slope) <= 0.05 * abs(slope)
