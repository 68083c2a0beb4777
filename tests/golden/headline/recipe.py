"""Seeded inputs of the headline-size parity vectors (shared by make_vectors.py and the tests).

Maps follow the lognormal recipe of the benchmark: g ~ N(0, 1) at nside 32 copied to its
NESTED children, plus N(0, 0.1^2) per pixel, x = exp(g / 2) - mean (numpy default_rng).
"""
import numpy as np

NDIR = 8      # gradient projections per case
NSAMPLE = 4096  # gradient pixels per case

ENSEMBLE = [(128, 8), (256, 4), (512, 2)]  # (nside, maps); J = log2(nside) - 1


def _cases():
    out = [dict(name="bench_n1024", nside=1024, J=9, seed=1, moments=True)]
    for g, (nside, maps) in enumerate(ENSEMBLE):
        for b in range(maps):
            out.append(dict(name=f"ens_n{nside}_b{b}", nside=nside, J=nside.bit_length() - 2,
                            seed=100 + 10 * g + b, moments=False))
    return out


CASES = _cases()


def case(name):
    return next(c for c in CASES if c["name"] == name)


def lognormal(nside, seed, base=32, sigma=0.1):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(12 * base * base)
    g = np.repeat(g, (nside // base) ** 2) + sigma * rng.standard_normal(12 * nside * nside)
    x = np.exp(0.5 * g)
    return x - x.mean()


def cotangent(ncoef, seed):
    """Complex N(0, 1) cotangent of the statistics."""
    rng = np.random.default_rng(10_000 + seed)
    return rng.standard_normal(ncoef) + 1j * rng.standard_normal(ncoef)


def direction(nside, i):
    """Projection direction i: N(0, 1) per pixel."""
    return np.random.default_rng(20_000 + i).standard_normal(12 * nside * nside)


def sample_ids(nside):
    """Sorted pixel sample: the 12 base-face corners' first and last pixels plus random ones."""
    npix = 12 * nside * nside
    per_face = nside * nside
    fixed = np.concatenate([np.arange(12) * per_face, np.arange(1, 13) * per_face - 1])
    rng = np.random.default_rng(30_000 + nside)
    rand = rng.choice(npix, NSAMPLE - fixed.size, replace=False)
    return np.unique(np.concatenate([fixed, rand])).astype(np.int64)
