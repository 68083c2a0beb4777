"""Single map at nside 256 (786432 pixels), J = 7, L = 4: the lognormal recipe
(g ~ N(0, 1) at nside 32 copied to the 64 NESTED children, plus N(0, 0.1^2)
per pixel, x = exp(g / 2) - mean).

- FP64 engine vs the FP64 oracle: statistics 1e-13 per family, gradient 1e-12 (the reference's
  device bars, crates/stl-gpu/tests/stats.rs:248-275).
- Mixed engine (FP32 bulk, FP64 sums) vs the oracle: every statistic family
  within 1e-6 (max |dS| / max |S| per family), gradient within 1e-5.
- The VJP through torch.autograd.Function (stl_port.torch_fn.scattering) on
  the FP64 engine against central finite differences, eps 1e-4, along 16
  random unit directions, of the pairing Re sum(conj(c) S(x)): error at most
  1e-6 of the gradient norm."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NSIDE, J = 256, 7
FAMILY_NAMES = ["mean", "var", "S1", "S2", "S3", "S4"]


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def lognormal(seed, nside=NSIDE, base=32, sigma=0.1):
    rng = np.random.default_rng(seed)
    g = np.repeat(rng.standard_normal(12 * base * base), (nside // base) ** 2)
    x = np.exp(0.5 * (g + sigma * rng.standard_normal(12 * nside * nside)))
    return x - x.mean()


def family_errors(values, ref, plan):
    fam = np.array([FAMILY_NAMES.index(c["family"]) for c in plan])
    out = {}
    for f, name in enumerate(FAMILY_NAMES):
        sel = fam == f
        scale = np.abs(ref[..., sel]).max() if name != "mean" else np.sqrt(np.abs(ref[..., fam == 1]).max())
        out[name] = float(np.abs(values[..., sel] - ref[..., sel]).max() / scale)
    return out


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def reference():
    x = lognormal(11)
    print(f"[gpu] oracle nside={NSIDE} J={J}", flush=True)
    orc = O.Engine(NSIDE, J)
    orc.load(x)
    s = orc.forward()
    cot = 2.0 * (s - np.conj(s[:, ::-1]) * 0.5)  # a generic cotangent with the statistics' scales
    g = orc.gradient(cot)
    return x, s, cot, g


@pytest.mark.parametrize("precision,stat_tol,grad_tol", [("fp64", 1e-13, 1e-12), ("mixed", 1e-6, 1e-5)])
def test_engine_vs_oracle(reference, precision, stat_tol, grad_tol):
    from stl_port import NativeEngine
    x, s, cot, g = reference
    print(f"[gpu] NativeEngine nside={NSIDE} J={J} precision={precision}", flush=True)
    with NativeEngine(NSIDE, J, backend="gpu", precision=precision) as eng:
        values, _ = eng.forward(x[None])
        errs = family_errors(values, s, eng.plan)
        grad = eng.vjp(cot)
    gerr = rel(grad, g)
    print(precision, "statistics", errs, "gradient", gerr)
    assert max(errs.values()) <= stat_tol, errs
    assert gerr <= grad_tol


def test_autograd_central_differences():
    from stl_port import NativeEngine
    from stl_port.torch_fn import scattering
    rng = np.random.default_rng(12)
    print(f"[gpu] autograd + finite differences nside={NSIDE}", flush=True)
    with NativeEngine(NSIDE, J, backend="gpu") as eng:
        x0 = torch.from_numpy(lognormal(13)[None]).cuda()
        c = torch.from_numpy(rng.standard_normal(eng.ncoeff) + 1j * rng.standard_normal(eng.ncoeff)).cuda()

        def pairing(x):
            return (torch.conj(c) * scattering(eng, x)).real.sum()

        x = x0.clone().requires_grad_(True)
        pairing(x).backward()
        g = x.grad.detach()
        # A unit direction among 786432 pixels has a derivative ~1e-3 |g|_2: the
        # pairing's rounding (~1e-14 of it) over 2 eps is then ~1e-5 of that
        # derivative.  The error is bounded by 1e-6 |g|_2 (the largest
        # directional derivative over unit directions); a relative gradient
        # error of 1e-3 would exceed it.
        eps = 1e-4
        gnorm = float(torch.linalg.norm(g))
        worst, worst_rel = 0.0, 0.0
        for _ in range(16):
            v = torch.from_numpy(rng.standard_normal(x0.shape)).cuda()
            v /= torch.linalg.norm(v)
            with torch.no_grad():
                fd = (pairing(x0 + eps * v) - pairing(x0 - eps * v)).item() / (2 * eps)
            exact = float((g * v).sum())
            worst = max(worst, abs(fd - exact) / gnorm)
            worst_rel = max(worst_rel, abs(fd - exact) / abs(exact))
        print("central differences: worst |fd - exact| / |g|", worst, "worst relative", worst_rel)
        assert worst <= 1e-6
