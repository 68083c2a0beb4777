"""The oracle's geometry against the reference's geometry fixtures
(mirrors crates/stl-core/tests/geometry.rs:20-235)."""
from collections import defaultdict

import numpy as np
import pytest

import oracle as O
from fixtures import cases, conv_noise

TOL = 1e-12


def weight_tol(nside):
    # geometry.rs:76-78: one ulp of acos/atan2 moves the weights by ~4e-16 nside.
    return max(TOL, 1e-14 * nside)


def as_operator(pix, w):
    m = defaultdict(float)
    for p, v in zip(pix, w):
        m[int(p)] += v
    return m


def same_operator(a, b, tol):
    return all(abs(a.get(p, 0.0) - b.get(p, 0.0)) <= tol for p in set(a) | set(b))


@pytest.mark.parametrize("case", cases("geometry", "centres"), ids=lambda c: c.name)
def test_centres(case):
    nside = case["nside"]
    order = O.order_of(nside)
    ids, ix, iy, face = case.arr("ids"), case.arr("ix"), case.arr("iy"), case.arr("face")
    theta, phi, lon, lat = case.arr("theta"), case.arr("phi"), case.arr("lon"), case.arr("lat")
    err = np.zeros(4)
    for i, pid in enumerate(ids):
        assert O.nest2xyf(pid, order) == (ix[i], iy[i], face[i])
        t, p = O.xyf2ang(ix[i], iy[i], face[i], nside)
        lo, la = O.centre_lonlat(pid, nside)
        err = np.maximum(err, np.abs([t - theta[i], p - phi[i], lo - lon[i], la - lat[i]]))
    assert (err <= 1e-14).all(), err


@pytest.mark.parametrize("case", cases("geometry", "ring2nest"), ids=lambda c: c.name)
def test_ring2nest(case):
    order = case["order"]
    for r, n in zip(case.arr("ring"), case.arr("nest")):
        assert O.ring2nest(r, order) == n


@pytest.mark.parametrize("case", cases("geometry", "interp"), ids=lambda c: c.name)
def test_interpolation(case):
    nside = case["nside"]
    order = O.order_of(nside)
    vecs, theta, phi = case.arr("vecs").reshape(-1, 3), case.arr("theta"), case.arr("phi")
    pr, wr = case.arr("pix_ring").reshape(-1, 4), case.arr("w_ring").reshape(-1, 4)
    pn, wn = case.arr("pix_nest").reshape(-1, 4), case.arr("w_nest").reshape(-1, 4)
    bad = []
    for i in range(len(theta)):
        p, w = O.get_interpol_ring(theta[i], phi[i], nside)
        if not same_operator(as_operator(p, w), as_operator(pr[i], wr[i]), TOL):
            bad.append(("ring", i))
        p, w = O.interp_weights_nested(vecs[i], order)
        if not same_operator(as_operator(p, w), as_operator(pn[i], wn[i]), weight_tol(nside)):
            bad.append(("nest", i))
    assert not bad, bad[:10]


@pytest.mark.parametrize("case", cases("geometry", "rotation"), ids=lambda c: c.name)
def test_rotations(case):
    G = case["G"]
    theta, phi, R = case.arr("theta"), case.arr("phi"), case.arr("R").reshape(-1, G, 3, 3)
    err = max(np.abs(O.rotation_matrices(theta[k], phi[k], G, case["gauge"]) - R[k]).max() for k in range(len(theta)))
    assert err <= 1e-14


@pytest.mark.parametrize("case", cases("geometry", "kernel_grid"), ids=lambda c: c.name)
def test_kernel_grid(case):
    v = O.local_kernel_grid(case["kernel_sz"], case["nside"])
    assert np.abs(v.ravel() - case.arr("vec").ravel()).max() <= 1e-15


@pytest.mark.parametrize("case", cases("geometry", "conv"), ids=lambda c: c.name)
def test_convolution_geometry_and_forward(case):
    nside, ksz, G, gauge = case["nside"], case["kernel_sz"], case["G"], case["gauge"]
    ids = case.arr("ids")
    pos, w, complete = O.conv_geometry(ids, nside, ksz, G, gauge)
    assert (complete == case.arr("complete").astype(bool)).all()
    P = ksz * ksz
    ker = np.tile(case.arr("weights").ravel().astype(np.float32).astype(np.float64), G)
    assert np.array_equal(ker, case.arr("weight_f32").ravel())
    if case.has("pos_safe"):
        rp, rw = case.arr("pos_safe").reshape(4, G, -1, P), case.arr("w_norm").reshape(4, G, -1, P)
        for k in np.nonzero(complete)[0][:64]:
            for g in range(G):
                for t in range(P):
                    assert same_operator(as_operator(pos[:, g, k, t], w[:, g, k, t]),
                                         as_operator(rp[:, g, k, t], rw[:, g, k, t]), weight_tol(nside))
    x, y = case.arr("x").ravel(), case.arr("y").reshape(G * 2, -1)
    mine = O.conv_forward(pos, w, ker, x)
    ratio = 0.0
    for k in np.nonzero(complete)[0]:
        th, _ = O.xyf2ang(*O.nest2xyf(ids[k], O.order_of(nside)), nside)
        ratio = max(ratio, np.abs(mine[:, k] - y[:, k]).max() / conv_noise(th, nside))
    assert ratio <= 1.0, ratio


@pytest.mark.parametrize("case", cases("geometry", "down"), ids=lambda c: c.name)
def test_decimation_weights(case):
    nside = case["nside"]
    coarse, rows, cols, vals = O.smooth_decimation(case.arr("ids"), nside)
    assert np.array_equal(coarse, case.arr("coarse_ids"))
    assert np.array_equal(rows, case.arr("rows")) and np.array_equal(cols, case.arr("cols"))
    assert np.abs(vals - case.arr("values")).max() <= TOL


@pytest.mark.parametrize("case", cases("geometry", "neighbours"), ids=lambda c: c.name)
def test_neighbourhood_sets(case):
    """kth_neighbourhood (cdshealpix 0.10.1, neighbours.rs:154-157) as sets: the
    reference sorts and deduplicates every use (neighbours.rs:182-184,
    decimation.rs:556-558), so only membership matters."""
    nside, k = case["nside"], case["k"]
    nb = case.arr("neighbours")
    for i, pid in enumerate(case.arr("ids")):
        expected = np.sort(nb[i][nb[i] >= 0])
        assert np.array_equal(O.kth_neighbourhood(pid, nside, k), expected), pid


def test_ext_pix_unit_offsets_are_the_neighbours():
    """dia.rs:309-323: the 3x3 window of the extended frame is the 1-ring."""
    for nside in (2, 4, 8, 16):
        order = O.order_of(nside)
        for pid in range(12 * nside * nside):
            x, y, f = O.nest2xyf(pid, order)
            ours = sorted({c for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                           if (c := O.ext_pix(f, x + dx, y + dy, nside)) >= 0})
            assert ours == list(O.kth_neighbourhood(pid, nside, 1) if nside > 1 else range(12))


def test_offsets_round_trip():
    """dia.rs:325-341."""
    nside = 16
    order = O.order_of(nside)
    for pid in range(0, 12 * nside * nside, 7):
        x, y, f = O.nest2xyf(pid, order)
        for dy in range(-6, 7):
            for dx in range(-6, 7):
                q = O.ext_pix(f, x + dx, y + dy, nside)
                if q < 0:
                    continue
                assert O.rel_offset(pid, q, nside) == (dx, dy)
                bx, by = O.rel_offset(q, pid, nside)
                assert max(abs(bx), abs(by)) == max(abs(dx), abs(dy))


def test_full_sky_decimation_rows_equal_partial_grid_rows():
    """decimation.rs:545-548: a full-sky row equals the row built on any grid
    holding its candidates, to the bit."""
    nside = 32
    ids = np.arange(12 * nside * nside)
    coarse, rows, cols, vals = O.smooth_decimation(ids, nside)
    indptr, fc, fv = O.decimation_rows(nside, 0, len(coarse))
    assert np.array_equal(fc, cols) and np.array_equal(fv, vals)
    assert np.array_equal(np.diff(indptr), np.bincount(rows, minlength=len(coarse)))
