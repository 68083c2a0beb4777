"""The oracle's full-sky wavelet and decimation against the reference's block
fixtures (mirrors crates/stl-core/tests/block.rs:42-105), plus adjoint
identities that the reference states independently of its fixtures."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from fixtures import cases, conv_noise

TOL = 1e-12


def level_input(case):
    name = case.name
    tag = name[name.find("_n"): name.find("_d")]
    want = f"level{tag}_d{case['dg']}"
    return next(c for c in cases("block", "level") if c.name == want).arr("x")


def engine_for(case):
    return O.Engine(case["nside"], case["J"], case["L"], case["kernel_sz"], case["gauge"], 1, 1)


def full_decimation(nside):
    indptr, cols, vals = O.decimation_rows(nside, 0, 3 * nside * nside)
    return sp.csr_matrix((vals, cols, indptr), shape=(3 * nside * nside, 12 * nside * nside))


@pytest.mark.parametrize("case", cases("block", "wavelet"), ids=lambda c: c.name)
def test_wavelet_kernel(case):
    w = O.wavelet_weights(case["kernel_sz"])
    assert np.abs(w - case.arr("weights").ravel()).max() <= 1e-15


@pytest.mark.parametrize("case", cases("block", "decimate"), ids=lambda c: c.name)
def test_decimate_and_adjoint(case):
    nside = case["nside"] >> case["dg"]
    D = full_decimation(nside)
    x = level_input(case).reshape(-1, 12 * nside * nside)
    y = (D @ x.T).T
    assert np.abs(y.ravel() - case.arr("y").ravel()).max() <= TOL
    g = case.arr("g").reshape(-1, 3 * nside * nside)
    gx = (D.T @ g.T).T
    assert np.abs(gx.ravel() - case.arr("gx").ravel()).max() <= TOL


@pytest.mark.parametrize("case", cases("block", "wave"), ids=lambda c: c.name)
def test_wave_and_adjoint(case):
    """z on a block is the full-sky z at the block rows; the adjoint of the
    block's cotangent (zero elsewhere) is the full-sky adjoint on the domain."""
    e = engine_for(case)
    dg = case["dg"]
    nside = case["nside"] >> dg
    N = 12 * nside * nside
    order = O.order_of(nside)
    x = level_input(case).reshape(-1, N)
    R = x.shape[0]
    block, domain = case.arr("block"), case.arr("domain")
    z = e.conv(dg, x)  # [R, L, N]
    ref = case.complex("z").reshape(R, case["L"], -1)
    ratio = 0.0
    for i, pid in enumerate(block):
        th, _ = O.xyf2ang(*O.nest2xyf(pid, order), nside)
        ratio = max(ratio, np.abs(z[:, :, pid] - ref[:, :, i]).max() / conv_noise(th, nside))
    assert ratio <= 1.0, ratio
    gz = np.zeros((R, case["L"], N), np.complex128)
    gz[:, :, block] = case.complex("gz").reshape(R, case["L"], -1)
    gx = e.conv_adjoint(dg, gz)
    ref = case.arr("gx").reshape(R, -1)
    ratio = 0.0
    for i, pid in enumerate(domain):
        th, _ = O.xyf2ang(*O.nest2xyf(pid, order), nside)
        ratio = max(ratio, np.abs(gx[:, pid] - ref[:, i]).max() / conv_noise(th, nside))
    assert ratio <= 1.0, ratio
    # nothing outside the domain
    outside = np.setdiff1d(np.arange(N), domain)
    assert not outside.size or np.abs(gx[:, outside]).max() == 0.0


@pytest.mark.parametrize("nside", [4, 8, 16])
def test_conv_adjoint_identity(nside):
    """<W x, g> = <x, W^T g> (real part of the complex pairing)."""
    e = O.Engine(nside, 1, 4, 5)
    rng = np.random.default_rng(nside)
    N = 12 * nside * nside
    x = rng.standard_normal((2, N))
    g = rng.standard_normal((2, 4, N)) + 1j * rng.standard_normal((2, 4, N))
    lhs = np.sum(e.conv(0, x).real * g.real + e.conv(0, x).imag * g.imag)
    rhs = np.sum(x * e.conv_adjoint(0, g))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_kernel_is_zero_mean_and_l1_normalised():
    w = O.wavelet_weights(5).reshape(2, 25)
    k = w[0] + 1j * w[1]
    assert abs(k.sum()) <= 1e-15
    assert abs(np.abs(k).sum() - 1.0) <= 1e-15
    # point reflection conjugates the kernel: k(-x, -y) = conj(k(x, y))
    assert np.abs(k[::-1] - np.conj(k)).max() <= 1e-16


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 32])
def test_decimation_rows_are_l1_normalised(nside):
    D = full_decimation(2 * nside)
    assert np.abs(np.asarray(D.sum(axis=1)).ravel() - 1.0).max() <= 1e-14
    assert (D.data > 0).all()
