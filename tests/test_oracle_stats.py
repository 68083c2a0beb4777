"""The oracle's full-sky sweeps against the reference's stats fixtures
(mirrors crates/stl-core/tests/stats.rs:77-210)."""
import json

import numpy as np
import pytest

import oracle as O
from fixtures import cases

STATS = cases("stats", "stats")


def engine(case):
    return O.Engine(case["nside"], case["J"], case["L"], case["kernel_sz"], case["gauge"], case["channels"],
                    case["batch"], case["fewer"], case["j_to_dg"], case["block_pixels"])


def close_ratio(a, b, atol, rtol):
    return float(np.max(np.abs(a - b) / (atol + rtol * np.abs(b))))


def flat(z):
    return np.concatenate([z.real.ravel(), z.imag.ravel()])


@pytest.mark.parametrize("case", STATS, ids=lambda c: c.name)
def test_plan(case):
    fam = {n: i for i, n in enumerate(O.FAMILIES)}
    rows = json.loads((case.dir / "plan.json").read_text())
    ref = np.array([[fam[r[0]]] + [-1 if v is None else v for v in r[1:]] for r in rows])
    mine = O.plan(case["J"], case["L"], case["j_to_dg"], case["channels"])
    assert np.array_equal(mine, ref)


@pytest.mark.parametrize("case", STATS, ids=lambda c: c.name)
def test_statistics_and_gradients(case):
    e = engine(case)
    e.load(case.arr("map0"))
    for d in range(e.levels):
        assert np.abs(e.map(d).ravel() - case.arr(f"map{d}").ravel()).max() <= 1e-12
    values = e.forward()
    ref = case.complex("values").reshape(values.shape)
    # statistics: 1e-11 absolute + 1e-10 relative (stats.rs:183-186)
    assert close_ratio(flat(values), flat(ref), 1e-11, 1e-10) <= 1.0
    for d in range(e.levels):
        for n in range(d + 1, e.levels):
            assert np.abs(e.stored_modulus(d, n).ravel() - case.arr(f"m{d}_{n}").ravel()).max() <= 1e-12
    mg = e.reverse(case.complex("gradient").reshape(values.shape))
    for d in range(e.levels):
        assert close_ratio(mg[d].ravel(), case.arr(f"maps_gradient{d}").ravel(), 1e-10, 1e-9) <= 1.0
    bands = e.analysis_vjp(mg)
    for d in range(e.levels):
        assert close_ratio(bands[d].ravel(), case.arr(f"band_gradient{d}").ravel(), 1e-10, 1e-9) <= 1.0


@pytest.mark.parametrize("case", [c for c in STATS if c["nside"] == 16 and c["block_pixels"] == 64],
                         ids=lambda c: c.name)
def test_gradient_matches_finite_difference(case):
    """Central difference along a pseudo-random direction (stats.rs:77-110)."""
    e = engine(case)
    x = case.arr("map0").ravel()
    g = case.complex("gradient").reshape(case["batch"], -1)

    def pairing(xv):
        e.load(xv)
        return float(np.sum((np.conj(g) * e.forward()).real))

    e.load(x)
    e.forward()
    grad = e.gradient(g).ravel()
    state = 0x2545F4914F6CDD1D
    v = np.empty(x.size)
    mask = (1 << 64) - 1
    for i in range(x.size):
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        v[i] = (state >> 11) / float(1 << 53) - 0.5
    eps = 1e-5
    fd = (pairing(x + eps * v) - pairing(x - eps * v)) / (2 * eps)
    exact = float(grad @ v)
    assert abs(fd - exact) <= 1e-7 * abs(exact)


def test_coefficient_spatial_variance():
    """mean/var second moments are the spatial variances of x and x^2 (stats.rs:139-163)."""
    case = next(c for c in STATS if "C2_B2_full_b64" in c.name)
    e = engine(case)
    x = case.arr("map0").reshape(case["batch"], case["channels"], -1)
    e.load(x)
    values = e.forward()
    e.load(x)
    values2, second = e.forward(moments=True)
    assert np.array_equal(values, values2)
    var = second - np.abs(values) ** 2
    plan = O.plan(case["J"], case["L"], case["j_to_dg"], case["channels"])
    for i, row in enumerate(plan):
        if row[0] > 1:
            continue
        for b in range(case["batch"]):
            f = x[b, row[2]] if row[0] == 0 else x[b, row[2]] ** 2
            expect = np.mean((f - f.mean()) ** 2)
            assert abs(max(var[b, i], 0.0) - expect) <= 1e-12 * expect
