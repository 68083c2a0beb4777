"""Masked averages of the oracle (crates/stl-core/src/mask.rs, engine.rs:500-724).

Pinned without the reference's own engine (its Rust tests need the disc model, not restated):
  * the mask pyramid holds the parents of every level (mask.rs:43-70, tests/mask.rs:108-121);
  * the whole-sky mask is the unmasked engine (values, second moments, gradient);
  * masks made of whole base faces partition every level, so the count-weighted sum of their
    averages is the whole-sky average (values and second moments) and cotangents
    count_k / N times a common one give the whole-sky gradient;
  * the mean and var coefficients are the closed forms over the mask;
  * the gradient with latitude masks (not aligned on any level) against central differences.
"""
import math

import numpy as np
import pytest

import oracle as O

NSIDE, J = 16, 3
NPIX = 12 * NSIDE ** 2


def latitude_masks(nside, cuts):
    """Indicators of |b| > cut (degrees) about the z axis (cut 0: the whole sky)."""
    lat = np.array([O.centre_lonlat(p, nside)[1] for p in range(12 * nside ** 2)])
    return np.stack([np.ones(lat.size) if c == 0 else (np.abs(lat) > math.radians(c)).astype(float) for c in cuts])


@pytest.fixture(scope="module")
def setup():
    rng = np.random.default_rng(5)
    x = 1.0 + rng.standard_normal((2, NPIX))
    e = O.Engine(NSIDE, J, B=2)
    e.load(x)
    v0, s0 = e.forward(moments=True)
    rng2 = np.random.default_rng(6)
    cot = rng2.standard_normal(v0.shape) + 1j * rng2.standard_normal(v0.shape)
    e.load(x)
    e.forward()
    g0 = e.gradient(cot)
    return e, x, v0, s0, cot, g0


def test_mask_pyramid_holds_the_parents():
    w = latitude_masks(8, (30.0,))
    e = O.Engine(8, 3)
    e.set_masks(w)
    level = np.flatnonzero(w[0])
    for n in range(3):
        assert e.mask_count(0, n) == level.size
        level = np.unique(level // 4)
    with pytest.raises(RuntimeError, match="empty mask"):
        e.set_masks(np.zeros((1, 12 * 64)))


def test_whole_sky_mask_is_the_unmasked_engine(setup):
    e, x, v0, s0, cot, g0 = setup
    e.set_masks(np.ones((1, NPIX)))
    try:
        e.load(x)
        v, s = e.forward(moments=True)
        assert v.shape == (2, 1, v0.shape[1])
        assert np.abs(v[:, 0] - v0).max() <= 1e-14 * np.abs(v0).max()
        assert np.abs(s[:, 0] - s0).max() <= 1e-14 * np.abs(s0).max()
        g = e.gradient(cot[:, None, :])
        assert np.abs(g - g0).max() <= 1e-13 * np.abs(g0).max()
    finally:
        e.set_masks(None)


def test_face_masks_partition_every_level(setup):
    e, x, v0, s0, cot, g0 = setup
    plan = O.plan(J, 4)
    lev = plan[:, 1]
    npf = NSIDE ** 2
    groups = [(0, 1, 2, 3), (4, 5, 6, 7, 8), (9, 10, 11)]
    w = np.zeros((len(groups), NPIX))
    for k, faces in enumerate(groups):
        for f in faces:
            w[k, f * npf:(f + 1) * npf] = 1.0
    e.set_masks(w)
    try:
        counts = np.array([[e.mask_count(k, n) for n in lev] for k in range(len(groups))], dtype=float)
        total = np.array([12 * (NSIDE >> n) ** 2 for n in lev], dtype=float)
        assert np.array_equal(counts.sum(axis=0), total)
        e.load(x)
        v, s = e.forward(moments=True)
        assert v.shape == (2, len(groups), len(plan))
        vsum = (counts[None] * v).sum(axis=1) / total
        ssum = (counts[None] * s).sum(axis=1) / total
        assert np.abs(vsum - v0).max() <= 1e-13 * np.abs(v0).max()
        assert np.abs(ssum - s0).max() <= 1e-13 * np.abs(s0).max()
        g = e.gradient(cot[:, None, :] * (counts / total)[None])
        assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max()
    finally:
        e.set_masks(None)


def test_mean_and_var_are_the_mask_averages(setup):
    e, x, v0, s0, cot, g0 = setup
    w = latitude_masks(NSIDE, (20.0, 45.0))
    e.set_masks(w)
    try:
        e.load(x)
        v = e.forward()
        for k in range(2):
            sel = w[k] != 0
            for b in range(2):
                assert v[b, k, 0] == pytest.approx(x[b, sel].mean(), rel=1e-14)
                assert v[b, k, 1] == pytest.approx((x[b, sel] ** 2).mean(), rel=1e-14)
                assert v[b, k, 0].imag == 0.0 and v[b, k, 1].imag == 0.0
    finally:
        e.set_masks(None)


def test_masked_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    w = latitude_masks(NSIDE, (0.0, 20.0, 45.0))
    x = 1.0 + rng.standard_normal(NPIX)
    u = rng.standard_normal(NPIX)
    e = O.Engine(NSIDE, J)
    e.set_masks(w)
    cot = rng.standard_normal((1, 3, e.ncoef)) + 1j * rng.standard_normal((1, 3, e.ncoef))

    def f(t):
        e.load(x + t * u)
        return float(np.sum(np.real(np.conj(cot) * e.forward())))

    h = 1e-5
    fd = (f(h) - f(-h)) / (2 * h)
    e.load(x)
    e.forward()
    an = float(e.gradient(cot).ravel() @ u)
    assert abs(an - fd) <= 1e-7 * abs(fd)
