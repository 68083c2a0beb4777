"""ctypes binding of the CPU FP64 oracle (liboracle.so).

Test infrastructure only: the tests, the smoke check and the CPU-baseline
leg of the benchmark use it as the checker.  The product library never
imports, links or calls it.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int)
_U8 = C.POINTER(C.c_uint8)

GAUGES = {"cosmo": 0, "phi": 1}
FAMILIES = ["mean", "var", "S1", "S2", "S3", "S4"]

_lib = None


def build():
    """Compile the oracle with g++ (plain IEEE double, no FMA contraction)."""
    import subprocess
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.is_file():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(str(LIB_PATH))
        sig = {
            "or_last_error": ([], C.c_char_p),
            "or_set_threads": ([C.c_int], None),
            "or_max_threads": ([], C.c_int),
            "or_nest2xyf": ([C.c_int64, C.c_int, _I64], None),
            "or_xyf2nest": ([C.c_int64, C.c_int64, C.c_int64, C.c_int], C.c_int64),
            "or_xyf2ang": ([C.c_int64, C.c_int64, C.c_int64, C.c_int64, _D], None),
            "or_centre_lonlat": ([C.c_int64, C.c_int64, _D], None),
            "or_ring2nest": ([C.c_int64, C.c_int], C.c_int64),
            "or_get_interpol_ring": ([C.c_double, C.c_double, C.c_int64, _I64, _D], None),
            "or_interp_weights_nested": ([_D, C.c_int, _I64, _D], None),
            "or_rotation_matrices": ([C.c_double, C.c_double, C.c_int, C.c_int, _D], None),
            "or_local_kernel_grid": ([C.c_int, C.c_int64, _D], None),
            "or_wavelet_weights": ([C.c_int, _D], None),
            "or_ext_pix": ([C.c_int64, C.c_int64, C.c_int64, C.c_int64], C.c_int64),
            "or_rel_offset": ([C.c_int64, C.c_int64, C.c_int64, _I64], C.c_int),
            "or_kth_neighbourhood": ([C.c_int64, C.c_int64, C.c_int, _I64, C.c_int], C.c_int),
            "or_conv_geometry": ([_I64, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, _I64, _D, _U8], C.c_int),
            "or_smooth_decimation": ([_I64, C.c_int64, C.c_int64, _I64, _I64, _I64, _I64, _D, C.c_int64], C.c_int64),
            "or_decimation_rows": ([C.c_int64, C.c_int64, C.c_int64, _I64, _I64, _D, C.c_int64], C.c_int64),
            "or_plan": ([C.c_int, C.c_int, _I32, C.c_int, _I32, C.c_int], C.c_int),
            "or_engine_new": ([C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _I32,
                               C.c_int64, C.c_int64], C.c_void_p),
            "or_engine_free": ([C.c_void_p], None),
            "or_engine_levels": ([C.c_void_p], C.c_int),
            "or_engine_ncoef": ([C.c_void_p], C.c_int),
            "or_engine_load": ([C.c_void_p, _D], C.c_int),
            "or_engine_forward": ([C.c_void_p, _D, _D], C.c_int),
            "or_engine_set_masks": ([C.c_void_p, _D, C.c_int], C.c_int),
            "or_engine_mask_count": ([C.c_void_p, C.c_int, C.c_int], C.c_int64),
            "or_engine_reverse": ([C.c_void_p, _D, _D], C.c_int),
            "or_engine_analysis_vjp": ([C.c_void_p, _D, _D], C.c_int),
            "or_engine_get_map": ([C.c_void_p, C.c_int, _D], C.c_int),
            "or_engine_get_m": ([C.c_void_p, C.c_int, C.c_int, _D], C.c_int),
            "or_engine_get_z": ([C.c_void_p, C.c_int, _D], C.c_int),
            "or_engine_conv": ([C.c_void_p, C.c_int, _D, C.c_int, _D], C.c_int),
            "or_engine_conv_adjoint": ([C.c_void_p, C.c_int, _D, C.c_int, _D], C.c_int),
            "or_engine_kernel": ([C.c_void_p, _D], None),
            "or_full_row": ([C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int64, _I64, _D], C.c_int),
        }
        for name, (a, r) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = a, r
        _lib = L
    return _lib


def _p(a, t=_D):
    return a.ctypes.data_as(t)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().or_last_error().decode())


def set_threads(n):
    lib().or_set_threads(int(n))


def order_of(nside):
    return int(nside).bit_length() - 1


def nest2xyf(pix, order):
    out = np.zeros(3, np.int64)
    lib().or_nest2xyf(int(pix), int(order), _p(out, _I64))
    return tuple(int(v) for v in out)


def xyf2ang(ix, iy, f, nside):
    out = np.zeros(2)
    lib().or_xyf2ang(int(ix), int(iy), int(f), int(nside), _p(out))
    return float(out[0]), float(out[1])


def centre_lonlat(pix, nside):
    out = np.zeros(2)
    lib().or_centre_lonlat(int(pix), int(nside), _p(out))
    return float(out[0]), float(out[1])


def ring2nest(pix, order):
    return int(lib().or_ring2nest(int(pix), int(order)))


def get_interpol_ring(theta, phi, nside):
    p, w = np.zeros(4, np.int64), np.zeros(4)
    lib().or_get_interpol_ring(float(theta), float(phi), int(nside), _p(p, _I64), _p(w))
    return p, w


def interp_weights_nested(v, order):
    v = np.ascontiguousarray(v, np.float64)
    p, w = np.zeros(4, np.int64), np.zeros(4)
    lib().or_interp_weights_nested(_p(v), int(order), _p(p, _I64), _p(w))
    return p, w


def rotation_matrices(theta, phi, G, gauge):
    out = np.zeros((G, 3, 3))
    lib().or_rotation_matrices(float(theta), float(phi), int(G), GAUGES[gauge], _p(out))
    return out


def local_kernel_grid(ksz, nside):
    out = np.zeros((ksz * ksz, 3))
    lib().or_local_kernel_grid(int(ksz), int(nside), _p(out))
    return out


def wavelet_weights(ksz):
    out = np.zeros(2 * ksz * ksz)
    lib().or_wavelet_weights(int(ksz), _p(out))
    return out


def ext_pix(face, ex, ey, nside):
    return int(lib().or_ext_pix(int(face), int(ex), int(ey), int(nside)))


def rel_offset(row, cell, nside):
    out = np.zeros(2, np.int64)
    ok = lib().or_rel_offset(int(row), int(cell), int(nside), _p(out, _I64))
    return (int(out[0]), int(out[1])) if ok else None


def kth_neighbourhood(pix, nside, k):
    cap = max(12 * nside * nside, (2 * k + 1) ** 2) if k >= nside else (2 * k + 1) ** 2
    out = np.zeros(cap, np.int64)
    n = lib().or_kth_neighbourhood(int(pix), int(nside), int(k), _p(out, _I64), cap)
    return out[:n].copy()


def conv_geometry(ids, nside, ksz, G, gauge):
    """Bound stencils of the grid rows: pos/w [4, G, K, P], complete [K]."""
    ids = np.ascontiguousarray(ids, np.int64)
    K, P = len(ids), ksz * ksz
    pos = np.zeros((4, G, K, P), np.int64)
    w = np.zeros((4, G, K, P))
    complete = np.zeros(K, np.uint8)
    _check(lib().or_conv_geometry(_p(ids, _I64), K, int(nside), int(ksz), int(G), GAUGES[gauge],
                                  _p(pos, _I64), _p(w), _p(complete, _U8)))
    return pos, w, complete.astype(bool)


def conv_forward(pos, w, ker, x):
    """Reference-order forward of bound stencils: x [n_in] -> y [G*2, K]."""
    _, G, K, P = pos.shape
    gathered = np.zeros((G, K, P))
    for n in range(4):
        gathered = gathered + x[pos[n]] * w[n]
    ker = ker.reshape(G, 2, P)
    return np.einsum("gkp,gop->gok", gathered, ker).reshape(G * 2, K)


def smooth_decimation(ids, nside):
    ids = np.ascontiguousarray(ids, np.int64)
    cap = 200 * len(ids) + 1000
    coarse = np.zeros(len(ids), np.int64)
    nco = np.zeros(1, np.int64)
    rows, cols, vals = np.zeros(cap, np.int64), np.zeros(cap, np.int64), np.zeros(cap)
    e = lib().or_smooth_decimation(_p(ids, _I64), len(ids), int(nside), _p(coarse, _I64), _p(nco, _I64),
                                   _p(rows, _I64), _p(cols, _I64), _p(vals), cap)
    if e < 0:
        _check(-1)
    assert e <= cap
    return coarse[: nco[0]].copy(), rows[:e].copy(), cols[:e].copy(), vals[:e].copy()


def decimation_rows(nside, c0, c1):
    """Full-sky rows of coarse cells [c0, c1) from nside: (indptr, cols, vals)."""
    cap = 160 * (c1 - c0) + 16 * 12 * nside * nside
    indptr = np.zeros(c1 - c0 + 1, np.int64)
    cols, vals = np.zeros(cap, np.int64), np.zeros(cap)
    e = lib().or_decimation_rows(int(nside), int(c0), int(c1), _p(indptr, _I64), _p(cols, _I64), _p(vals), cap)
    assert e <= cap
    return indptr, cols[:e].copy(), vals[:e].copy()


def plan(J, L, j_to_dg=None, C=1):
    j_to_dg = np.ascontiguousarray(list(range(J)) if j_to_dg is None else j_to_dg, np.int32)
    cap = 100000
    out = np.zeros((cap, 10), np.int32)
    n = lib().or_plan(int(J), int(L), _p(j_to_dg, _I32), int(C), _p(out, _I32), cap)
    return out[:n].copy()


class Engine:
    """Full-sky FP64 sweeps (forward, reverse, analysis VJP) of one configuration."""

    def __init__(self, nside, J, L=4, ksz=5, gauge="cosmo", C=1, B=1, fewer=False, j_to_dg=None,
                 block_pixels=4096, cache_rows=300_000):
        self.nside, self.J, self.L, self.C, self.B = int(nside), int(J), int(L), int(C), int(B)
        dg = np.ascontiguousarray(list(range(J)) if j_to_dg is None else j_to_dg, np.int32)
        self.j_to_dg = dg
        self.h = lib().or_engine_new(self.nside, self.J, self.L, int(ksz), GAUGES[gauge], self.C, self.B,
                                     int(bool(fewer)), _p(dg, _I32), int(block_pixels), int(cache_rows))
        if not self.h:
            _check(-1)
        self.levels = lib().or_engine_levels(self.h)
        self.ncoef = lib().or_engine_ncoef(self.h)
        self.npix = [12 * (self.nside >> d) ** 2 for d in range(self.levels)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_engine_free(self.h)
            self.h = None

    def load(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(self.B * self.C * self.npix[0])
        _check(lib().or_engine_load(self.h, _p(x)))

    nmasks = 0

    def set_masks(self, w):
        """Masked averages (mask.rs): w [K, 12 nside^2] (nonzero: in), or None for the whole sky."""
        if w is None:
            _check(lib().or_engine_set_masks(self.h, None, 0))
            self.nmasks = 0
            return
        w = np.ascontiguousarray(np.atleast_2d(w), np.float64)
        _check(lib().or_engine_set_masks(self.h, _p(w), w.shape[0]))
        self.nmasks = w.shape[0]

    def mask_count(self, k, n):
        return int(lib().or_engine_mask_count(self.h, int(k), int(n)))

    def _vshape(self):
        return (self.B, self.nmasks, self.ncoef) if self.nmasks else (self.B, self.ncoef)

    def forward(self, moments=False):
        v = np.zeros(self._vshape(), np.complex128)
        s = np.zeros(self._vshape()) if moments else None
        _check(lib().or_engine_forward(self.h, _p(v.view(np.float64)), _p(s) if moments else None))
        return (v, s) if moments else v

    def reverse(self, cot):
        cot = np.ascontiguousarray(cot, np.complex128).reshape(self._vshape())
        total = sum(self.B * self.C * n for n in self.npix)
        mg = np.zeros(total)
        _check(lib().or_engine_reverse(self.h, _p(cot.view(np.float64)), _p(mg)))
        out, at = [], 0
        for n in self.npix:
            out.append(mg[at: at + self.B * self.C * n].reshape(self.B * self.C, n))
            at += self.B * self.C * n
        return out

    def analysis_vjp(self, mg):
        flat = np.ascontiguousarray(np.concatenate([np.asarray(g).ravel() for g in mg]))
        bands = np.zeros_like(flat)
        _check(lib().or_engine_analysis_vjp(self.h, _p(flat), _p(bands)))
        out, at = [], 0
        for n in self.npix:
            out.append(bands[at: at + self.B * self.C * n].reshape(self.B * self.C, n))
            at += self.B * self.C * n
        return out

    def gradient(self, cot):
        """Level-0 map cotangent of Re sum(conj(cot) * values)."""
        return self.analysis_vjp(self.reverse(cot))[0]

    def map(self, d):
        out = np.zeros((self.B * self.C, self.npix[d]))
        _check(lib().or_engine_get_map(self.h, int(d), _p(out)))
        return out

    def stored_modulus(self, d, n):
        out = np.zeros((self.B * self.C * self.L, self.npix[n]))
        _check(lib().or_engine_get_m(self.h, int(d), int(n), _p(out)))
        return out

    def z(self, n):
        out = np.zeros((self.B * self.C, self.L, self.npix[n]), np.complex128)
        _check(lib().or_engine_get_z(self.h, int(n), _p(out.view(np.float64))))
        return out

    def conv(self, n, x):
        x = np.ascontiguousarray(x, np.float64)
        R = x.size // self.npix[n]
        y = np.zeros((R, self.L, self.npix[n]), np.complex128)
        _check(lib().or_engine_conv(self.h, int(n), _p(x), R, _p(y.view(np.float64))))
        return y

    def conv_adjoint(self, n, gy):
        gy = np.ascontiguousarray(gy, np.complex128)
        R = gy.size // (self.L * self.npix[n])
        gx = np.zeros((R, self.npix[n]))
        _check(lib().or_engine_conv_adjoint(self.h, int(n), _p(gy.view(np.float64)), R, _p(gx)))
        return gx

    def kernel(self):
        P = 25 if True else 0
        out = np.zeros(self.L * 2 * 25 * 4)  # generous; trimmed below
        lib().or_engine_kernel(self.h, _p(out))
        return out


def full_row(nside, ksz, G, gauge, row):
    P = ksz * ksz
    pos = np.zeros((G, P, 4), np.int64)
    w = np.zeros((G, P, 4))
    _check(lib().or_full_row(int(nside), int(ksz), int(G), GAUGES[gauge], int(row), _p(pos, _I64), _p(w)))
    return pos, w


if os.environ.get("STL_ORACLE_THREADS"):
    set_threads(int(os.environ["STL_ORACLE_THREADS"]))
