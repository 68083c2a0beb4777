"""The recompute field store (the reference's FieldStore::Recompute, crates/stl-gpu/src/
engine.rs:172-186, docs/gpu.md:241-256): the forward keeps the maps and moduli only, and the
reverse recomputes each level's z and branches.  The reference holds Keep and Recompute to the
same bits (crates/stl-gpu/tests/stats.rs:185-217); so does this library, since the recomputation
runs the forward's own kernels on the same inputs.

- keep vs recompute: values, variances and gradients equal bit for bit with every level launched
  on its own (batch_nside=0: the recompute store always launches per level), FP64 and mixed,
  graph replays and eager launches; with the keep store's coarse levels batched (the default)
  they agree to rounding (the batched levels' split gathers add a row's slots in another order);
- the recompute engine against the oracle at the reference's device bars;
- checkpoint / restore of a recompute engine (its state is the maps and moduli);
- nside 2048, J = 10, FP64, 2 maps with the recompute store: fits one GPU, equals the keep
  store's result bit for bit, and reports its peak device memory.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def run(eng, x, cot_seed, moments=True):
    v, s = eng.forward(x, generation=1, moments=moments)
    rng = np.random.default_rng(cot_seed)
    cot = rng.standard_normal(v.shape) + 1j * rng.standard_normal(v.shape)
    return v, s, eng.vjp(cot, generation=1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("timing", [False, True])
def test_recompute_equals_keep_bitwise(precision, timing):
    from stl_port import NativeEngine
    nside, J, B = 64, 5, 2
    x = np.random.default_rng(1).standard_normal((B, 12 * nside * nside))
    print(f"[gpu] fields keep vs recompute nside={nside} J={J} B={B} {precision} timing={timing}", flush=True)
    out = {}
    for fields, bn in (("keep", 0), ("recompute", 0), ("keep", 32)):
        with NativeEngine(nside, J, batch=B, precision=precision, fields=fields, timing=timing,
                          batch_nside=bn) as e:
            info = e.info()
            assert info["fields"] == fields
            key = fields + ("_batched" if bn else "")
            out[key] = run(e, x, 2) + (info["device_memory"]["fields_bytes"],)
            # a second evaluation on other maps (graph replays, buffers reused across levels)
            out[key + "2"] = run(e, x[::-1].copy(), 3)
    for a, b in zip(out["keep"][:3], out["recompute"][:3]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(out["keep2"], out["recompute2"]):
        np.testing.assert_array_equal(a, b)
    tol = 1e-11 if precision == "fp64" else 1e-5
    for a, b in zip(out["keep_batched"][:3], out["recompute"][:3]):
        assert np.abs(a - b).max() <= tol * np.abs(b).max()
    assert out["recompute"][3] < 0.75 * out["keep"][3]


def test_recompute_matches_oracle_and_restores():
    from stl_port import NativeEngine
    nside, J = 32, 4
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal((2, 12 * nside * nside))
    with NativeEngine(nside, J, precision="fp64", fields="recompute") as e:
        vx, _ = e.forward(x, generation=1, objective="x")
        snap = e.checkpoint()
        vy, _ = e.forward(y, generation=1, objective="y")
        e.vjp(vy, generation=1, objective="y")
        e.restore(snap, generation=1, objective="x")
        gx = e.vjp(vx, generation=1, objective="x")
        snap.close()
    ref = O.Engine(nside, J)
    ref.load(x)
    sx = ref.forward()
    assert np.abs(vx - sx).max() <= 1e-13 * np.abs(sx).max()
    gr = ref.gradient(vx)
    assert np.abs(gx - gr).max() <= 1e-12 * np.abs(gr).max()


def test_auto_keeps_small_pyramids():
    from stl_port import NativeEngine
    with NativeEngine(16, 3, precision="fp64") as e:
        assert e.info()["fields"] == "keep"


def lognormal(nside, seed, base=32, sigma=0.1):
    rng = np.random.default_rng(seed)
    g = np.repeat(rng.standard_normal(12 * base * base), (nside // base) ** 2)
    x = np.exp(0.5 * (g + sigma * rng.standard_normal(12 * nside * nside)))
    return x - x.mean()


def test_fp64_pair_at_nside_2048_recompute_equals_keep():
    """Two FP64 maps at nside 2048 (J = 10, 15042 coefficients each): the recompute engine's
    statistics and gradients, then (after it is freed) the keep engine's (every level launched
    on its own, as the recompute store does), bit for bit; the recompute store's peak device
    memory is printed."""
    from stl_port import NativeEngine
    nside, J = 2048, 10
    xd = torch.from_numpy(np.stack([lognormal(nside, 3), lognormal(nside, 4)])).cuda()
    grad = torch.empty_like(xd)
    res = {}
    for fields in ("recompute", "keep"):
        print(f"[gpu] nside {nside} J {J} fp64 batch 2 fields={fields}", flush=True)
        with NativeEngine(nside, J, batch=2, precision="fp64", fields=fields, batch_nside=0) as e:
            v, _ = e.forward_device(xd, generation=1)
            rng = np.random.default_rng(5)
            cot = rng.standard_normal(v.shape) + 1j * rng.standard_normal(v.shape)
            e.vjp_device(cot, grad, generation=1)
            torch.cuda.synchronize()
            mem = e.info()["device_memory"]
            print(f"[gpu] fields={fields}: total {mem['total_bytes'] / 1e9:.1f} GB, peak "
                  f"{mem['peak_bytes'] / 1e9:.1f} GB, fields {mem['fields_bytes'] / 1e9:.1f} GB", flush=True)
            res[fields] = (v, grad.cpu().numpy())
        torch.cuda.empty_cache()
    np.testing.assert_array_equal(res["recompute"][0], res["keep"][0])
    np.testing.assert_array_equal(res["recompute"][1], res["keep"][1])
