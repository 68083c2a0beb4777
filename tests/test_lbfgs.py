"""L-BFGS (stlp_lbfgs) on the host: torch.optim.LBFGS with strong Wolfe
(reference crates/stl-core/src/lbfgs.rs:356-465), pinned by torch's own
optimizer on small problems and by the reference's synthesis fixture
(crates/stl-mpi/tests/synthesis.rs:59-87: 20 iterations of STL-Dev on the
websky convergence map at nside 64, J = 5, L = 2)."""
import numpy as np
import pytest
import torch

import oracle
from fixtures import cases
from stl_port import native


def _torch_history(fun, x0, **opts):
    # The reference's defaults (torch's are history 100, tolerances 1e-7 and 1e-9).
    opts = dict(dict(history_size=10, tolerance_grad=1e-12, tolerance_change=1e-15), **opts)
    x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([x], line_search_fn="strong_wolfe", **opts)
    history = []

    def closure():
        opt.zero_grad()
        loss, grad = fun(x.detach().numpy().copy())
        x.grad = torch.tensor(grad, dtype=torch.float64)
        history.append(float(loss))
        return torch.tensor(loss, dtype=torch.float64)

    opt.step(closure)
    return np.array(history), x.detach().numpy()


def rosenbrock(x):
    a, b = x[:-1], x[1:]
    loss = np.sum(100.0 * (b - a * a) ** 2 + (1.0 - a) ** 2)
    g = np.zeros_like(x)
    g[:-1] += -400.0 * a * (b - a * a) - 2.0 * (1.0 - a)
    g[1:] += 200.0 * (b - a * a)
    return loss, g


def quadratic(n, seed):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, n))
    a = m @ m.T / n + np.eye(n)
    c = rng.standard_normal(n)
    return lambda x: (0.5 * x @ a @ x - c @ x, a @ x - c), np.linalg.solve(a, c)


@pytest.mark.parametrize("x0,opts", [
    (np.array([-1.2, 1.0]), dict(max_iter=20)),
    (np.array([-1.2, 1.0]), dict(max_iter=40, history_size=3)),
    (np.array([2.0, 2.0]), dict(max_iter=15, lr=0.5)),
])
def test_rosenbrock_matches_torch(x0, opts):
    x, report = native.native_lbfgs(rosenbrock, x0.copy(), opts)
    ref, xr = _torch_history(rosenbrock, x0, **opts)
    h = np.array(report["loss_history"])
    assert len(h) == len(ref) == report["evaluations"]
    # Two-element dot products and fused updates round as torch's do: the same bits.
    np.testing.assert_array_equal(h, ref)
    np.testing.assert_array_equal(x, xr)


def test_rosenbrock_4d_converges_like_torch():
    """Four-element dot products round differently from torch's BLAS: the
    trajectories agree closely, then both converge to (1, 1, 1, 1)."""
    x0, opts = np.array([-1.2, 1.0, -0.5, 0.3]), dict(max_iter=80, history_size=4)
    x, report = native.native_lbfgs(rosenbrock, x0.copy(), opts)
    ref, xr = _torch_history(rosenbrock, x0, **opts)
    np.testing.assert_allclose(report["loss_history"][:30], ref[:30], rtol=1e-8)
    np.testing.assert_allclose(x, np.ones(4), atol=1e-5)
    np.testing.assert_allclose(xr, np.ones(4), atol=1e-5)


def test_quadratic_converges_and_matches_torch():
    f, xs = quadratic(40, 3)
    x0 = np.zeros(40)
    x, report = native.native_lbfgs(f, x0.copy(), dict(max_iter=100, max_eval=200))
    ref, xr = _torch_history(f, x0, max_iter=100, max_eval=200)
    assert report["evaluations"] == len(ref)
    np.testing.assert_allclose(report["loss_history"], ref, rtol=1e-9)
    np.testing.assert_allclose(x, xs, atol=1e-8)


def test_stationary_start_evaluates_once():
    x, report = native.native_lbfgs(lambda x: (1.0, np.zeros_like(x)), np.ones(5))
    assert report["evaluations"] == 1
    np.testing.assert_array_equal(x, np.ones(5))


def test_max_eval_bounds_evaluations():
    _, report = native.native_lbfgs(rosenbrock, np.array([-1.2, 1.0]), dict(max_iter=100, max_eval=7))
    assert report["evaluations"] <= 7 + 25  # the line search of the last iteration may finish
    ref, _ = _torch_history(rosenbrock, np.array([-1.2, 1.0]), max_iter=100, max_eval=7)
    assert report["evaluations"] == len(ref)


def test_errors():
    with pytest.raises(ValueError):
        native.native_lbfgs(rosenbrock, np.zeros(2), dict(max_iter=0))
    with pytest.raises(ValueError):
        native.native_lbfgs(rosenbrock, np.zeros(2), dict(bogus=1))
    with pytest.raises(ValueError):
        native.native_lbfgs(rosenbrock, np.array([np.nan, 0.0]))
    with pytest.raises(ZeroDivisionError):
        native.native_lbfgs(lambda x: (1 / 0, x), np.zeros(2))
    with pytest.raises(RuntimeError, match="non-finite"):
        native.native_lbfgs(lambda x: (np.inf, x), np.ones(2))
    with pytest.raises(ValueError, match="gradient shape"):
        native.native_lbfgs(lambda x: (1.0, np.ones(3)), np.ones(2))


def test_synthesis_trajectory_matches_fixture():
    """The 25 evaluations of STL-Dev's synthesis, with the oracle's
    statistics as objective (unit weights; reference synthesis.rs:207-224)."""
    c = cases("synthesis", "synthesis")[0]
    eng = oracle.Engine(c["nside"], c["J"], L=c["L"], block_pixels=c["block_pixels"], j_to_dg=c["j_to_dg"])
    eng.load(c.arr("target_map"))
    target = eng.forward()
    ref = c.complex("target_values").reshape(target.shape)
    assert np.max(np.abs(target - ref) / (1e-11 + 1e-10 * np.abs(ref))) <= 1.0

    def objective(x):
        eng.load(x)
        r = eng.forward() - target
        return float(np.sum(np.abs(r) ** 2)), eng.gradient(2.0 * r).reshape(x.shape)

    x, report = native.native_lbfgs(objective, c.arr("initial_map").copy(),
                                    dict(max_iter=c["niter"], history_size=c["history_size"], lr=c["lr"],
                                         tolerance_grad=c["tolerance_grad"],
                                         tolerance_change=c["tolerance_change"]))
    h, r = np.array(report["loss_history"]), c.arr("loss_history")
    assert len(h) == len(r)
    assert all(f"{a:.6e}" == f"{b:.6e}" for a, b in zip(h, r))
    np.testing.assert_allclose(h, r, rtol=1e-8)
    np.testing.assert_allclose(x, c.arr("result_map"), atol=1e-8)
