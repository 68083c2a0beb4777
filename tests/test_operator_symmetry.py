"""Symmetries of the folded wavelet convolution that the device store relies on.

The device library stores the convolution of a level for the faces 0 and 4
only (stl-dev-mpi_amd/csrc/common.hpp, geometry.hip):

- faces 1-3 / 5-7 are rotations by pi/2 of faces 0 / 4, with a gauge
  permutation and conjugation (gauge_src, gauge_conj);
- faces 8-11 are the north-south mirrors of faces 0-3:
  A[M(p), M(c)][g] = conj(A[p, c][g]) with M(face f, x, y) = (f + 8, n-1-y, n-1-x);
- the rows of the equator ring (faces 4-7, x + y = n - 1) are their own mirrors
  but not mirror-symmetric, so the columns of the south cells they read are
  corrected by a list built as in build_mirror_fix.

Checked on the oracle's full-sky rows (the reference's convolution geometry,
conv.rs:73-135 with dia.rs:213-238) at nside 8, every row.
"""
from collections import defaultdict

import numpy as np
import pytest

import oracle as O

NSIDE = 8
ORDER = 3
NPF = NSIDE * NSIDE
N = 12 * NPF
G = 4


def spread(v):
    return sum(((v >> b) & 1) << (2 * b) for b in range(ORDER))


def compress(v):
    return sum(((v >> (2 * b)) & 1) << b for b in range(ORDER))


def nest(x, y, f):
    return f * NPF + spread(x) + (spread(y) << 1)


def rot(p, k):
    f, loc = divmod(p, NPF)
    return ((f & ~3) | ((f + k) & 3)) * NPF + loc


def mirror(p):
    f, loc = divmod(p, NPF)
    x, y = compress(loc), compress(loc >> 1)
    f2 = f + 8 if f < 4 else (f - 8 if f >= 8 else f)
    return nest(NSIDE - 1 - y, NSIDE - 1 - x, f2)


def gauge_conj(k, g):
    return (k == 1 and g < 2) or k == 2 or (k == 3 and g >= 2)


@pytest.fixture(scope="module")
def A():
    """Dense folded operator [G][row][cell]: kernel x bilinear weight summed per cell."""
    ker = O.wavelet_weights(5).astype(np.float32).astype(np.float64).reshape(2, 25)
    kc = ker[0] + 1j * ker[1]
    out = np.zeros((G, N, N), complex)
    for p in range(N):
        pos, w = O.full_row(NSIDE, 5, G, "cosmo", p)
        for g in range(G):
            np.add.at(out[g, p], pos[g].ravel(), (kc[:, None] * w[g]).ravel())
    return out


def tol(A):
    return 1e-13 * np.abs(A).max()


RING = [nest(i, NSIDE - 1 - i, f) for f in range(4, 8) for i in range(NSIDE)]
MIRROR = np.array([mirror(c) for c in range(N)])


def test_mirror_is_an_involution_onto_the_mirrored_faces():
    assert sorted(MIRROR.tolist()) == list(range(N))
    assert all(MIRROR[MIRROR[c]] == c for c in range(N))
    assert all(MIRROR[c] // NPF == c // NPF + 8 for c in range(4 * NPF))
    assert all(MIRROR[c] == c for c in RING)


def test_rotation_symmetry(A):
    for k in range(1, 4):
        perm = [rot(c, k) for c in range(N)]
        for p in [*range(0, NPF), *range(4 * NPF, 5 * NPF), *range(8 * NPF, 9 * NPF)]:
            for g in range(G):
                want = A[g ^ (2 * (k & 1)), p]
                if gauge_conj(k, g):
                    want = np.conj(want)
                got = np.zeros(N, complex)
                got[perm] = want
                assert np.abs(A[g, rot(p, k)] - got).max() <= tol(A), (k, p, g)


def test_south_rows_are_conjugated_mirrors_of_north_rows(A):
    for p in range(4 * NPF):
        for g in range(G):
            got = A[g, mirror(p)][MIRROR]
            assert np.abs(got - np.conj(A[g, p])).max() <= tol(A), (p, g)


def test_only_the_equator_ring_breaks_the_mirror(A):
    broken = []
    for p in range(4 * NPF, 8 * NPF):
        err = max(np.abs(A[g, mirror(p)][MIRROR] - np.conj(A[g, p])).max() for g in range(G))
        if err > tol(A):
            broken.append(p)
    assert sorted(broken) == sorted(RING)


def test_mirrored_columns_with_the_correction_list_are_exact(A):
    # The correction list, as build_mirror_fix makes it: for each entry (row p
    # on the ring, cell c) with c south: + A[p, c] at c; with c north:
    # - conj(A[p, c]) at M(c).
    fix = defaultdict(lambda: np.zeros(G, complex))
    for p in RING:
        for c in np.nonzero(np.abs(A[:, p, :]).sum(0))[0]:
            fc = c // NPF
            if fc >= 8:
                fix[(c, p)] += A[:, p, c]
            elif fc < 4:
                fix[(mirror(c), p)] -= np.conj(A[:, p, c])
    assert 0 < len(fix) < 4 * NSIDE * 64
    # The kernels' south column of q: the column of M(q), rows mirrored and
    # coefficients conjugated, plus the list's entries of q.
    for q in range(8 * NPF, 12 * NPF):
        col = np.zeros((G, N), complex)
        mq = mirror(q)
        for r in np.nonzero(np.abs(A[:, :, mq]).sum(0))[0]:
            col[:, mirror(r)] += np.conj(A[:, r, mq])
        for (c, p), d in fix.items():
            if c == q:
                col[:, p] += d
        assert np.abs(col - A[:, :, q]).max() <= tol(A), q
