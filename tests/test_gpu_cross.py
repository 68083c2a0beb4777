"""Cross statistics of map pairs (NativeEngine(cross=True)) on the GPU against the FP64 oracle.

The oracle computes the two-channel plan (its cross terms pinned by the reference's C = 2
stats fixtures); the cross plan is its c1 != c2 part in the same order (the reference's
plan_of(cross = true), crates/stl-bridge/src/lib.rs:238-245).  Bars: the reference's own
GPU-vs-CPU bars (python/stl_ensemble/tests/test_masks.py:66-85): values 1e-13, variances
1e-12, gradient 1e-11 relative to the largest entry (FP64); the mixed path 1e-5 per family
and 1e-4 on the gradient.
"""
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NSIDE, J = 32, 4
NPIX = 12 * NSIDE ** 2
FAMILY_NAMES = ["mean", "var", "S1", "S2", "S3", "S4"]
KEYS = ("family", "native_dg", "c1", "c2", "j1", "j2", "j3", "l1", "l2", "l3")


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel(a, b):
    return float(np.abs(np.ravel(a) - np.ravel(b)).max() / np.abs(b).max())


def worst_family(values, ref, plan):
    fam = np.array([FAMILY_NAMES.index(c["family"]) for c in plan])
    return max(float(np.abs(values[..., fam == f] - ref[..., fam == f]).max() / np.abs(ref[..., fam == f]).max())
               for f in set(fam.tolist()))


def cross_columns():
    full = O.plan(J, 4, None, 2)
    cols = np.flatnonzero(full[:, 2] != full[:, 3])
    return full, cols


def oracle_cross(x, cot, w=None):
    """Values, variances and gradient of the cross coefficients (x [B, 2, npix])."""
    full, cols = cross_columns()
    B = x.shape[0]
    ref = O.Engine(NSIDE, J, C=2, B=B)
    if w is not None:
        ref.set_masks(w)
    ref.load(x)
    v, s2 = ref.forward(moments=True)
    var = np.maximum(s2 - np.abs(v) ** 2, 0.0)
    cfull = np.zeros(v.shape, np.complex128)
    cfull[..., cols] = cot
    ref.load(x)
    ref.forward()
    g = ref.gradient(cfull).reshape(x.shape)
    return v[..., cols], var[..., cols], g


def test_cross_plan_is_the_two_channel_cross_terms():
    from stl_port import NativeEngine
    with NativeEngine(8, 3, cross=True) as small:
        pass
    full3 = O.plan(3, 4, None, 2)
    want = [dict(zip(KEYS, (FAMILY_NAMES[r[0]],) + tuple(None if v < 0 else int(v) for v in r[1:])))
            for r in full3 if r[2] != r[3]]
    got = [{k: c[k] for k in KEYS} for c in small.plan]
    assert got == want


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cross_matches_oracle(precision):
    from stl_port import NativeEngine
    rng = np.random.default_rng(7)
    x = rng.standard_normal((2, 2, NPIX))
    x[:, 1] += 0.5 * x[:, 0]  # correlated channels
    print(f"[gpu] cross nside={NSIDE} J={J} precision={precision}", flush=True)
    with NativeEngine(NSIDE, J, batch=2, cross=True, precision=precision) as eng:
        v, var = eng.forward(x, moments=True)
        cot = rng.standard_normal(v.shape) + 1j * rng.standard_normal(v.shape)
        g = eng.vjp(cot)
        va, _ = eng.forward_cross(x[:, 0], x[:, 1], moments=True)
        assert np.array_equal(va, v)
        ga, gb = eng.vjp_cross(cot)
        assert np.array_equal(ga, g[:, 0]) and np.array_equal(gb, g[:, 1])
        plan = eng.plan
    vr, varr, gr = oracle_cross(x, cot)
    assert v.shape == vr.shape
    if precision == "fp64":
        assert rel(v, vr) < 1e-13 and rel(var, varr) < 1e-12 and rel(g, gr) < 1e-11
    else:
        assert worst_family(v, vr, plan) < 1e-5 and rel(g, gr) < 1e-4


def test_cross_with_masks_matches_oracle():
    from stl_port import NativeEngine
    rng = np.random.default_rng(8)
    x = rng.standard_normal((1, 2, NPIX))
    lat = np.array([O.centre_lonlat(p, NSIDE)[1] for p in range(NPIX)])
    w = np.stack([np.ones(NPIX), (np.abs(lat) > math.radians(30.0)).astype(float)])
    print(f"[gpu] cross + masks nside={NSIDE} J={J}", flush=True)
    with NativeEngine(NSIDE, J, batch=1, cross=True, precision="fp64", masks=w) as eng:
        v, var = eng.forward(x, moments=True)
        assert v.shape == (1, 2, eng.ncoeff)
        cot = rng.standard_normal(v.shape) + 1j * rng.standard_normal(v.shape)
        g = eng.vjp(cot)
    vr, varr, gr = oracle_cross(x, cot, w)
    assert rel(v, vr) < 1e-13 and rel(var, varr) < 1e-12 and rel(g, gr) < 1e-11
