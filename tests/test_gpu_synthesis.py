"""All-GPU synthesis (stlx_synthesize: L-BFGS on device vectors with the HIP
engine as objective, reference crates/stl-mpi/src/synthesis.rs:207-268 with
crates/stl-core/src/lbfgs.rs:356-465) against the host L-BFGS (stlp_lbfgs,
pinned by torch and by the reference's synthesis fixture in test_lbfgs.py)
driving the same engine, and against the host L-BFGS driving the FP64 oracle.

The device reductions add 256-element tree partials where the host adds
sequentially: the trajectories agree to rounding growth, held to 1e-8
relative on the losses here."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NSIDE, J = 32, 4


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def maps(seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(12 * NSIDE * NSIDE)
    target = np.exp(0.6 * g) - np.exp(0.18)
    start = rng.standard_normal(g.size) * target.std() + target.mean()
    return target[None], start[None]


_ENGINES = {}


def engine(precision="fp64"):
    from stl_port import NativeEngine
    if precision not in _ENGINES:
        print(f"[gpu] NativeEngine nside={NSIDE} J={J} precision={precision}", flush=True)
        _ENGINES[precision] = NativeEngine(NSIDE, J, backend="gpu", precision=precision)
    return _ENGINES[precision]


def host_history(objective_values, gradient, start, target, weights, opts):
    from stl_port import native_lbfgs

    def objective(x):
        r = objective_values(x) - target
        loss = float(np.sum(weights * np.abs(r) ** 2))
        return loss, gradient(2.0 * weights * r).reshape(x.shape)

    x, report = native_lbfgs(objective, start.copy(), opts)
    return np.array(report["loss_history"]), x


def test_device_synthesis_matches_host_lbfgs():
    eng = engine()
    target_map, start = maps(1)
    target, _ = eng.forward(target_map)
    w = np.ones(eng.ncoeff)
    opts = dict(max_iter=8, history_size=5)
    print("[gpu] host L-BFGS over the engine", flush=True)

    def values(x):
        return eng.forward(x)[0]

    ref, x_host = host_history(values, lambda c: eng.vjp(c), start, target, w, opts)
    print("[gpu] device synthesis", flush=True)
    x = torch.tensor(start, device="cuda:0", dtype=torch.float64)
    h = eng.synthesize(x, target, **opts)
    assert len(h) == len(ref)
    np.testing.assert_allclose(h, ref, rtol=1e-8)
    assert h[-1] < 0.5 * h[0]
    np.testing.assert_allclose(x.cpu().numpy(), x_host, atol=1e-8 * np.abs(x_host).max())


def test_device_synthesis_matches_oracle_with_ergodic_weights():
    from stl_port import ergodic_weights
    eng = engine()
    target_map, start = maps(2)
    target, var = eng.forward(target_map, moments=True)
    w = ergodic_weights(eng.native_pixels(), var[0])
    assert np.all(np.isfinite(w)) and np.all(w > 0)
    opts = dict(max_iter=5)
    orc = O.Engine(NSIDE, J)

    def values(x):
        orc.load(x)
        return orc.forward()

    print("[gpu] host L-BFGS over the oracle", flush=True)
    ref, _ = host_history(values, lambda c: orc.gradient(c), start, target, w, opts)
    print("[gpu] device synthesis, ergodic weights", flush=True)
    x = torch.tensor(start, device="cuda:0", dtype=torch.float64)
    h = eng.synthesize(x, target, weights=w, **opts)
    assert len(h) == len(ref)
    np.testing.assert_allclose(h, ref, rtol=1e-8)


def test_mixed_precision_synthesis_descends():
    eng64, eng32 = engine(), engine("mixed")
    target_map, start = maps(3)
    target, _ = eng64.forward(target_map)
    x64 = torch.tensor(start, device="cuda:0", dtype=torch.float64)
    x32 = x64.clone()
    h64 = eng64.synthesize(x64, target, max_iter=10)
    print("[gpu] mixed synthesis", flush=True)
    h32 = eng32.synthesize(x32, target, max_iter=10)
    assert abs(h32[0] - h64[0]) <= 1e-6 * h64[0]
    assert h32[-1] < 0.2 * h32[0]


def test_synthesis_errors():
    eng = engine()
    x = torch.zeros((1, 12 * NSIDE * NSIDE - 1), device="cuda:0", dtype=torch.float64)
    target = np.zeros(eng.ncoeff, np.complex128)
    with pytest.raises((RuntimeError, ValueError)):
        eng.synthesize(x, target)
    x = torch.zeros((1, 12 * NSIDE * NSIDE), device="cuda:0", dtype=torch.float64)
    with pytest.raises(RuntimeError, match="L-BFGS"):
        eng.synthesize(x, target, history_size=0)
    with pytest.raises(ValueError):
        eng.synthesize(x, target, weights=np.ones(3))
