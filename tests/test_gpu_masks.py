"""Masked averages on the GPU (stlp_set_masks) against the FP64 oracle's masks.

The reference's own GPU-vs-CPU mask test (python/stl_ensemble/tests/test_masks.py) holds its
GPU engine to 1e-13 on values, 1e-12 on variances and 1e-11 on the gradient against its CPU
engine; the same bars hold here against the oracle (FP64).  Cases: latitude cuts (the
reference's test masks), the whole-sky mask against the unmasked engine, ids against
indicators, and eight masks whose membership differs on every coarse pixel (192 classes: the
reverse reads its per-class matrices from HBM instead of LDS).  The mixed path is held to
1e-5 per family and 1e-4 on the gradient, as without masks.
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


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def latitude_masks(nside, cuts):
    lat = np.array([O.centre_lonlat(p, nside)[1] for p in range(12 * nside ** 2)])
    return [np.arange(lat.size) if c == 0 else np.flatnonzero(np.abs(lat) > math.radians(c)) for c in cuts]


def indicators(ids):
    w = np.zeros((len(ids), NPIX))
    for k, s in enumerate(ids):
        w[k, s] = 1.0
    return w


def rel(a, b):
    return float(np.abs(np.ravel(a) - np.ravel(b)).max() / np.abs(b).max())


def worst_family(values, ref, plan):
    fam = np.array([FAMILY_NAMES.index(c["family"]) for c in plan])
    out = 0.0
    for f in range(6):
        sel = fam == f
        scale = np.abs(ref[..., sel]).max()
        if f == 0:
            scale = np.sqrt(np.abs(ref[..., fam == 1]).max())
        out = max(out, float(np.abs(values[..., sel] - ref[..., sel]).max() / scale))
    return out


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, NPIX))
    cot_rng = np.random.default_rng(4)
    return x, cot_rng


def oracle_masked(x, w, cot):
    ref = O.Engine(NSIDE, J, B=2)
    ref.set_masks(w)
    ref.load(x)
    v, s2 = ref.forward(moments=True)
    var = np.maximum(s2 - np.abs(v) ** 2, 0.0)
    ref.load(x)
    ref.forward()
    g = ref.gradient(cot)
    counts = np.array([[ref.mask_count(k, n) for n in range(J)] for k in range(w.shape[0])])
    return v, var, g, counts


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_latitude_masks_match_oracle(data, precision):
    from stl_port import NativeEngine
    x, rng = data
    ids = latitude_masks(NSIDE, (0.0, 20.0, 45.0))
    w = indicators(ids)
    print(f"[gpu] masks latitude nside={NSIDE} J={J} precision={precision}", flush=True)
    # Masked sweeps launch every level on its own; the whole-sky comparison at the end is
    # with the same launches (batch_nside=0), so that only the masks differ.
    with NativeEngine(NSIDE, J, batch=2, precision=precision, masks=ids, batch_nside=0) as eng:
        v, var = eng.forward(x, moments=True)
        assert v.shape == var.shape == (2, 3, eng.ncoeff)
        cot = rng.standard_normal(v.shape) + 1j * rng.standard_normal(v.shape)
        g = eng.vjp(cot)
        vr, varr, gr, counts = oracle_masked(x, w, cot)
        pix = eng.native_pixels()
        assert pix.shape == (3, eng.ncoeff)
        assert np.array_equal(pix, counts[:, [c["native_dg"] for c in eng.plan]])
        if precision == "fp64":
            assert rel(v, vr) < 1e-13 and rel(var, varr) < 1e-12 and rel(g, gr) < 1e-11
        else:
            assert worst_family(v, vr, eng.plan) < 1e-5 and rel(g, gr) < 1e-4
        # Indicators give the same averages as ids (the same kernels: moments again, since the
        # moment sums may take another summation order); None restores the whole sky.
        eng.set_masks(w)
        v2, _ = eng.forward(x, moments=True)
        assert np.array_equal(v2, v)
        eng.set_masks(None)
        v0, _ = eng.forward(x)
        assert v0.shape == (2, eng.ncoeff)
        assert rel(v[:, 0], v0) < (1e-14 if precision == "fp64" else 1e-12)


def test_many_mask_classes_match_oracle(data):
    """Eight masks from the bits of the level-3 parent index: 192 membership classes per level,
    more per-class matrices than LDS holds, at every level."""
    from stl_port import NativeEngine
    x, rng = data
    parent = np.arange(NPIX) // (4 ** 3)
    w = np.stack([((parent >> k) & 1).astype(float) for k in range(8)])
    print(f"[gpu] masks 8 classes-heavy nside={NSIDE} J={J}", flush=True)
    with NativeEngine(NSIDE, J, batch=2, precision="fp64", masks=w) as eng:
        v, var = eng.forward(x, moments=True)
        cot = rng.standard_normal(v.shape) + 1j * rng.standard_normal(v.shape)
        g = eng.vjp(cot)
    vr, varr, gr, _ = oracle_masked(x, w, cot)
    assert rel(v, vr) < 1e-13 and rel(var, varr) < 1e-12 and rel(g, gr) < 1e-11


def test_masked_device_synthesis_loss(data):
    """The device synthesis with masks: weights and targets [K, ncoeff] per map; its first loss
    is the oracle's masked loss of the start."""
    from stl_port import NativeEngine
    x, rng = data
    ids = latitude_masks(NSIDE, (0.0, 30.0))
    w = indicators(ids)
    with NativeEngine(NSIDE, J, batch=1, precision="fp64", masks=ids) as eng:
        t, _ = eng.forward(x[1])
        xd = torch.tensor(x[:1], dtype=torch.float64, device="cuda")
        hist = eng.synthesize(xd, t, max_iter=3)
    ref = O.Engine(NSIDE, J, B=1)
    ref.set_masks(w)
    ref.load(x[:1])
    v0 = ref.forward()
    assert hist[0] == pytest.approx(float((np.abs(v0 - t) ** 2).sum()), rel=1e-10)
    assert hist[-1] < hist[0]


def test_masked_synthesis_with_weights_and_as_an_ensemble():
    """The device synthesis with masks and weights [K, ncoeff] (one per mask and coefficient,
    the layout of native_pixels), per map and as an ensemble (target [K, ncoeff] on the mean
    of the members): the first loss equals the oracle's weighted masked loss of the start to
    1e-10, and the loss falls."""
    from stl_port import NativeEngine
    x = np.random.default_rng(9).standard_normal((4, NPIX))
    ids = latitude_masks(NSIDE, (0.0, 30.0))
    w_ind = indicators(ids)
    with NativeEngine(NSIDE, J, batch=2, precision="fp64", masks=ids) as eng:
        t, _ = eng.forward(x[2:4])
        K, nc = len(ids), eng.ncoeff
        weights = np.random.default_rng(3).uniform(0.5, 2.0, (K, nc))
        ref = O.Engine(NSIDE, J, B=2)
        ref.set_masks(w_ind)
        ref.load(x[:2])
        v0 = ref.forward()  # [2, K, ncoeff]
        print("[gpu] masked synthesis with weights, per map", flush=True)
        xd = torch.tensor(x[:2], dtype=torch.float64, device="cuda")
        hist = eng.synthesize(xd, t, weights=weights, max_iter=3)
        expect = float((weights[None] * np.abs(v0 - t) ** 2).sum())
        assert hist[0] == pytest.approx(expect, rel=1e-10)
        assert hist[-1] < hist[0]
        print("[gpu] masked synthesis with weights, ensemble", flush=True)
        te = t.mean(axis=0)  # [K, ncoeff]
        xd = torch.tensor(x[:2], dtype=torch.float64, device="cuda")
        hist = eng.synthesize(xd, te, weights=weights, max_iter=3, ensemble=True)
        expect = float((weights * np.abs(v0.mean(axis=0) - te) ** 2).sum())
        assert hist[0] == pytest.approx(expect, rel=1e-10)
        assert hist[-1] < hist[0]
        with pytest.raises(ValueError, match="weights"):
            eng.synthesize(xd, t, weights=weights[0], max_iter=1)
