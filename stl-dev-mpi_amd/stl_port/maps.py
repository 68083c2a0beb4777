"""Map files and initial maps of the statistics and synthesis tools.

The file formats of the reference's command-line tools (crates/stl-mpi/src/io.rs): full-sky
NESTED maps as `.npy` (read: f8, f4, i8, i4 of either byte order; written: little-endian f8)
or HEALPix FITS binary tables (one `E` or `D` column; RING tables are reordered to NESTED on
reading), read whole or by pixel ranges and written whole or by slices.  The initial maps of
stl-synth (crates/stl-cli/src/lib.rs:523-540, 646-760): counter-based white noise, or a coarser
map upsampled by NESTED parent with noise for the missing scales.  The work is done by
libstl_bridge.so's host code (stlx_* map entry points); no Python fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .native import bridge

_SIZE = C.POINTER(C.c_size_t)


def _pointer(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _p(path):
    return os.fsencode(os.fspath(path))


def nside_of(npix):
    """nside of a full-sky map of npix pixels (io.rs:18-24); ValueError otherwise."""
    npix = int(npix)
    nside = int(round((npix // 12) ** 0.5))
    if nside <= 0 or 12 * nside * nside != npix or nside & (nside - 1):
        raise ValueError(f"{npix} pixels is not a full-sky HEALPix map")
    return nside


def map_len(path, library=None):
    """Pixels of a map file, from its header (io.rs:370-379)."""
    lib = bridge(library)
    out = C.c_size_t(0)
    lib.call('stlx_map_len', _p(path), C.byref(out))
    return int(out.value)


def read_map(path, library=None):
    """The whole NESTED map of a `.npy` / `.fits` file as float64 (io.rs:227-233)."""
    lib = bridge(library)
    out = np.empty(map_len(path, library))
    lib.call('stlx_read_map', _p(path), _pointer(out), out.size)
    return out


def read_map_range(path, start, end, library=None):
    """NESTED pixels [start, end) read at their offset (io.rs:387-416)."""
    lib = bridge(library)
    start, end = int(start), int(end)
    if start < 0 or end < start:
        raise ValueError(f"bad pixel range {start}..{end}")
    out = np.empty(end - start)
    lib.call('stlx_read_map_range', _p(path), start, end, _pointer(out))
    return out


def read_npy(path, library=None):
    """A `.npy` array as float64 with its shape (io.rs:27-49)."""
    lib = bridge(library)
    ndim = C.c_size_t(32)
    shape = np.zeros(32, dtype=np.uintp)
    lib.call('stlx_read_npy', _p(path), None, 0, shape.ctypes.data_as(_SIZE), C.byref(ndim))
    dims = tuple(int(d) for d in shape[:ndim.value])
    out = np.empty(dims)
    lib.call('stlx_read_npy', _p(path), _pointer(out), out.size, None, C.byref(ndim))
    return out


def write_npy(path, values, library=None):
    """Little-endian float64 `.npy` of the array's shape (io.rs:52-78)."""
    lib = bridge(library)
    v = np.ascontiguousarray(values, dtype=np.float64)
    shape = np.asarray(v.shape, dtype=np.uintp)
    lib.call('stlx_write_npy', _p(path), _pointer(v), shape.ctypes.data_as(_SIZE), v.ndim)


def write_map(path, values, column="VALUE", double=False, library=None):
    """A flat map file by extension: HEALPix FITS (float32 `E`, float64 `D` when double) for
    `.fits` / `.fit`, else `.npy` float64 (io.rs:52-55, 99-170)."""
    lib = bridge(library)
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    lib.call('stlx_write_map', _p(path), _pointer(v), v.size, column.encode(), bool(double))


class SlicedMap:
    """A map file written by slices (io.rs:431-520): `create` sizes it and writes the headers,
    `write` puts pixel ranges at their offsets.  `.npy` float64, or HEALPix FITS (`E` float32,
    `D` when double) for a `.fits` path."""

    def __init__(self, path, double=False, library=None):
        self.path = os.fspath(path)
        self.double = bool(double)
        self._lib = bridge(library)

    def create(self, n, column=""):
        self._lib.call('stlx_sliced_create', _p(self.path), int(n), column.encode(), self.double)

    def write(self, start, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        self._lib.call('stlx_sliced_write', _p(self.path), self.double, int(start), _pointer(v), v.size)


def pixel_normal(seed, start, end, library=None):
    """Standard normal values of pixels [start, end) for seed: splitmix64 of the seed and the
    pixel id, Box-Muller of two 53-bit uniforms (lib.rs:523-540); the same map whatever the
    decomposition."""
    lib = bridge(library)
    start, end = int(start), int(end)
    out = np.empty(max(end - start, 0))
    lib.call('stlx_pixel_normal', C.c_uint64(int(seed) & (2**64 - 1)), start, out.size, _pointer(out))
    return out


def moments(x):
    """Mean and standard deviation (population) of a map, as stl-synth and stl-stats
    standardise (lib.rs:636-641, 1366-1371)."""
    x = np.asarray(x, dtype=np.float64)
    mean = float(x.sum()) / x.size
    std = float(np.sqrt(((x - mean) ** 2).sum() / x.size))
    return mean, std


def initial_maps(target, members=1, seed=11, init=None, init_units="target", init_noise="auto",
                 mean=0.0, std=1.0, library=None):
    """Initial maps of stl-synth (lib.rs:646-760) in standardised units: [members, npix] and the
    report's "init" entry.

    target: the standardised target map (NESTED).  Without init, member m is the white noise of
    seed + m.  With init (a path; "{m}" in it gives one file per member, the member index on three
    digits), the init map of nside <= the target's is upsampled by NESTED parent (pixel p takes
    coarse pixel p // factor), standardised with the target's (mean, std) unless init_units is
    "standardized", plus noise_std * the white noise of seed + m: noise_std is init_noise when a
    number, else ("auto") the RMS of the target's scales below the init resolution (0 at equal
    nside).  One init shared by several members needs a noise.
    """
    target = np.asarray(target, dtype=np.float64).reshape(-1)
    npix = target.size
    nside = nside_of(npix)
    members = int(members)
    if members < 1:
        raise ValueError("members must be positive")
    if init_units not in ("target", "standardized"):
        raise ValueError("--init-units: 'target' or 'standardized'")
    if init is None:
        out = np.stack([pixel_normal(int(seed) + m, 0, npix, library) for m in range(members)])
        return out, None
    if isinstance(init_noise, str) and init_noise == "conditional":
        raise NotImplementedError("--init-noise conditional is not supported by this library")
    path = os.fspath(init)
    per_member = "{m}" in path

    def member_path(m):
        return path.replace("{m}", f"{m:03d}")

    first = member_path(0)
    n_init = map_len(first, library)
    nside_init = nside_of(n_init)
    if nside_init > nside or nside % nside_init:
        raise ValueError(f"{first}: nside {nside_init} incompatible with {nside}")
    factor = npix // n_init
    coarse = []
    for m in range(members):
        p = member_path(m)
        if map_len(p, library) != n_init:
            raise ValueError(f"{p}: not the nside of {first}")
        c = read_map_range(p, 0, n_init, library)
        if init_units == "target":
            c = (c - mean) / std
        coarse.append(c)
    if init_noise is None or (isinstance(init_noise, str) and init_noise == "auto"):
        if factor > 1:
            blocks = target.reshape(-1, factor)
            noise_std = float(np.sqrt(((blocks - blocks.mean(axis=1, keepdims=True)) ** 2).sum() / npix))
        else:
            noise_std = 0.0
    else:
        noise_std = float(init_noise)
    if members > 1 and not per_member and noise_std <= 0.0:
        raise ValueError("--ensemble from one --init needs a noise (--init-noise > 0 or auto below the target "
                         "nside), or one init per member ({m} in the path): identical members stay identical")
    out = np.empty((members, npix))
    for m in range(members):
        base = np.repeat(coarse[m], factor)
        out[m] = base + noise_std * pixel_normal(int(seed) + m, 0, npix, library) if noise_std > 0.0 else base
    info = {"file": path, "per_member": per_member, "nside": nside_init, "noise_std": noise_std}
    return out, info
