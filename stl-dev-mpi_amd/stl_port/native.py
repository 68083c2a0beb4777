"""ctypes binding of libstl_bridge.so, the MI355X ScatCov engine.

Mirrors the reference's `python/stl_ensemble/stl_port/native.py` (NativeEngine,
Snapshot, `_key`): same constructor keywords, shapes, identity keys and
errors, for the GPU engine of the complete one-channel plan.  The library has
no fallback: a missing library is an error.  Extra keywords of this library:
`precision` ("fp64" | "mixed") and `timing`.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import threading
from pathlib import Path

import numpy as np

_DOUBLE = C.POINTER(C.c_double)
_FLOAT = C.POINTER(C.c_float)
_UINT32 = C.POINTER(C.c_uint32)
_INT32 = C.POINTER(C.c_int32)
_UINT64 = C.POINTER(C.c_uint64)
_INT64 = C.POINTER(C.c_int64)
_SIZE = C.POINTER(C.c_size_t)
_VOID = C.c_void_p
_CALLBACK = C.CFUNCTYPE(C.c_double, _DOUBLE, C.c_size_t, _DOUBLE, _VOID)
_LIBRARIES = {}

DEFAULT_LIBRARY = Path(__file__).resolve().parents[1] / "libstl_bridge.so"

# The reference binding's table (native.py:55-108) plus this library's device forms and test hook.
SIGNATURES = {
    "stlp_last_error": ([], C.c_char_p),
    "stlp_string_free": ([_VOID], None),
    "stlp_new": ([C.c_char_p], _VOID),
    "stlp_free": ([_VOID], None),
    "stlp_info": ([_VOID], _VOID),
    "stlp_forward": ([_VOID, _DOUBLE, C.c_size_t, C.c_uint64, C.c_uint64, C.c_uint64, C.c_bool, _DOUBLE, _DOUBLE],
                     C.c_int),
    "stlp_vjp": ([_VOID, _DOUBLE, C.c_size_t, C.c_uint64, C.c_uint64, C.c_uint64, _DOUBLE], C.c_int),
    "stlp_forward_device": ([_VOID, C.c_uint64, C.c_size_t, C.c_uint64, C.c_uint64, C.c_uint64, C.c_bool, _DOUBLE,
                             _DOUBLE], C.c_int),
    "stlp_vjp_device": ([_VOID, _DOUBLE, C.c_size_t, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64], C.c_int),
    "stlp_checkpoint": ([_VOID], _VOID),
    "stlp_checkpoint_free": ([_VOID], None),
    "stlp_checkpoint_bytes": ([_VOID], C.c_size_t),
    "stlp_restore": ([_VOID, _VOID, C.c_uint64, C.c_uint64, C.c_uint64], C.c_int),
    "stlp_set_masks": ([_VOID, _DOUBLE, C.c_size_t], C.c_int),
    "stlp_spin_fields": ([_VOID, _DOUBLE, C.c_size_t, C.c_size_t, _DOUBLE, _DOUBLE], C.c_int),
    "stlp_forward_levels": ([_VOID, _DOUBLE, C.c_size_t, _DOUBLE], C.c_int),
    "stlp_native_pixels": ([_VOID, _SIZE], C.c_int),
    "stlp_lbfgs": ([_DOUBLE, C.c_size_t, C.c_char_p, _CALLBACK, _VOID], C.c_int64),
    "stlp_patch_new": ([C.c_char_p], _VOID),
    "stlp_patch_free": ([_VOID], None),
    "stlp_patch_info": ([_VOID], _VOID),
    "stlp_pixel_vectors": ([C.c_size_t, _INT64, C.c_size_t, _DOUBLE], C.c_int),
    "stlp_patch_select": ([_VOID, _DOUBLE, C.c_size_t, C.c_double, C.c_double, _SIZE], C.c_int),
    "stlp_patch_ids": ([_VOID, C.c_size_t, _INT64], C.c_int),
    "stlp_patch_core_pixels": ([_VOID, _SIZE], C.c_int),
    "stlp_patch_forward": ([_VOID, _DOUBLE, C.c_size_t, C.c_uint64, C.c_bool, _DOUBLE, _DOUBLE], C.c_int),
    "stlp_patch_vjp": ([_VOID, _DOUBLE, C.c_size_t, C.c_uint64, _DOUBLE], C.c_int),
    "stlp_patch_forward_device": ([_VOID, C.c_uint64, C.c_size_t, C.c_uint64, C.c_bool, _DOUBLE, _DOUBLE], C.c_int),
    "stlp_patch_vjp_device": ([_VOID, _DOUBLE, C.c_size_t, C.c_uint64, C.c_uint64], C.c_int),
    "stlp_spec_new": ([C.c_char_p], _VOID),
    "stlp_spec_free": ([_VOID], None),
    "stlp_spec_info": ([_VOID], _VOID),
    "stlp_spec_set_masks": ([_VOID, _DOUBLE, C.c_size_t], C.c_int),
    "stlp_spec_forward": ([_VOID, _DOUBLE, C.c_size_t, C.c_uint64, C.c_uint64, C.c_uint64, _DOUBLE], C.c_int),
    "stlp_spec_vjp": ([_VOID, _DOUBLE, C.c_size_t, C.c_uint64, C.c_uint64, C.c_uint64, _DOUBLE], C.c_int),
    "stlp_unet_new": ([C.c_char_p], _VOID),
    "stlp_unet_free": ([_VOID], None),
    "stlp_unet_info": ([_VOID], _VOID),
    "stlp_unet_get_params": ([_VOID, _FLOAT], C.c_int),
    "stlp_unet_set_params": ([_VOID, _FLOAT, C.c_size_t], C.c_int),
    "stlp_unet_get_grads": ([_VOID, _FLOAT], C.c_int),
    "stlp_unet_zero_grad": ([_VOID], C.c_int),
    "stlp_unet_adam": ([_VOID, C.c_double], C.c_int),
    "stlp_unet_set_cond": ([_VOID, _FLOAT, C.c_size_t], C.c_int),
    "stlp_unet_apply": ([_VOID, _DOUBLE, C.c_size_t, _DOUBLE, _DOUBLE, _DOUBLE], C.c_int),
    "stlp_train_set_masks": ([_VOID, _DOUBLE, C.c_size_t], C.c_int),
    "stlp_train_set_common": ([_VOID, _DOUBLE], C.c_int),
    "stlp_train_set_field": ([_VOID, C.c_uint32, _DOUBLE, _DOUBLE, _DOUBLE, _DOUBLE, _DOUBLE, _DOUBLE,
                              C.c_size_t, C.c_double], C.c_int),
    "stlp_train_bias": ([_VOID, C.c_uint32, _UINT64, _DOUBLE, C.c_size_t], C.c_int),
    "stlp_train_loss": ([_VOID, C.c_uint32, C.c_bool, _DOUBLE], C.c_int),
    "stlp_train_step": ([_VOID, C.c_uint32, C.c_double, _DOUBLE], C.c_int),
    "stlp_train_live": ([_VOID, C.c_uint32, _UINT64, _DOUBLE, C.c_size_t, _UINT64, _DOUBLE, C.c_size_t, C.c_bool,
                         C.c_double, _DOUBLE], C.c_int),
    "stlp_train_lbfgs": ([_VOID, _UINT32, C.c_size_t, C.c_char_p], _VOID),
    "stlp_train_set_joint": ([_VOID, _DOUBLE, _DOUBLE, C.c_uint32, C.c_uint32, C.c_char_p, C.c_double], C.c_int),
    "stlp_train_front": ([_VOID, _DOUBLE, C.c_size_t, C.c_bool, _DOUBLE], C.c_int),
    "stlx_apply_operator": ([_VOID, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64], C.c_int),
    "stlx_synthesize": ([_VOID, C.c_uint64, C.c_size_t, _DOUBLE, _DOUBLE, C.c_char_p, _DOUBLE, C.c_size_t],
                        C.c_int64),
    "stlx_ensemble_new": ([_INT32, _INT32, _INT32, C.c_size_t, C.c_char_p], _VOID),
    "stlx_ensemble_free": ([_VOID], None),
    "stlx_ensemble_ncoef": ([_VOID, _SIZE], C.c_int),
    "stlx_ensemble_forward_device": ([_VOID, _UINT64, C.c_size_t, C.c_bool, _DOUBLE, _DOUBLE], C.c_int),
    "stlx_ensemble_vjp_device": ([_VOID, _DOUBLE, C.c_size_t, _UINT64, C.c_size_t], C.c_int),
    "stlx_share": ([_VOID, C.c_char_p], _VOID),
    "stlx_map_len": ([C.c_char_p, _SIZE], C.c_int),
    "stlx_read_map": ([C.c_char_p, _DOUBLE, C.c_size_t], C.c_int),
    "stlx_read_map_range": ([C.c_char_p, C.c_size_t, C.c_size_t, _DOUBLE], C.c_int),
    "stlx_read_npy": ([C.c_char_p, _DOUBLE, C.c_size_t, _SIZE, _SIZE], C.c_int),
    "stlx_write_npy": ([C.c_char_p, _DOUBLE, _SIZE, C.c_size_t], C.c_int),
    "stlx_write_map": ([C.c_char_p, _DOUBLE, C.c_size_t, C.c_char_p, C.c_bool], C.c_int),
    "stlx_sliced_create": ([C.c_char_p, C.c_size_t, C.c_char_p, C.c_bool], C.c_int),
    "stlx_sliced_write": ([C.c_char_p, C.c_bool, C.c_size_t, _DOUBLE, C.c_size_t], C.c_int),
    "stlx_pixel_normal": ([C.c_uint64, C.c_int64, C.c_size_t, _DOUBLE], C.c_int),
}


def _pointer(a):
    return a.ctypes.data_as(_DOUBLE)


def library_path(path=None):
    if path is None:
        path = os.environ.get("STL_NATIVE_LIBRARY", str(DEFAULT_LIBRARY))
    return str(Path(path).resolve())


def _library(path=None):
    path = library_path(path)
    if path in _LIBRARIES:
        return _LIBRARIES[path]
    if not Path(path).is_file():
        raise RuntimeError(f"Native STL library is missing: {path}. Build it (python -c 'import __graft_entry__ as g; "
                           f"g.build()'); no fallback scattering implementation is used.")
    lib = C.CDLL(path)
    digest = hashlib.sha256()
    with open(path, "rb") as source:
        for chunk in iter(lambda: source.read(1 << 20), b""):
            digest.update(chunk)
    lib._stlp_binary_sha256 = digest.hexdigest()
    lib._stlp_binary_path = path
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes, fn.restype = argtypes, restype
    _LIBRARIES[path] = lib
    return lib


def _error(lib):
    return RuntimeError(lib.stlp_last_error().decode("utf-8", errors="replace"))


def _key(generation, member, objective):
    """Identity of a forward state (reference native.py:122-129)."""
    if isinstance(generation, bool) or not isinstance(generation, (int, np.integer)) or generation < 0 \
            or generation >= 2**64:
        raise ValueError("generation must be an unsigned 64-bit evaluation number")
    if isinstance(member, bool) or not isinstance(member, (int, np.integer)) or member < 0 or member >= 2**64:
        raise ValueError("member must be an unsigned 64-bit member number")
    if not isinstance(objective, str) or not objective:
        raise ValueError("objective must be a nonempty name")
    return (int(generation), int(member), int.from_bytes(hashlib.sha256(objective.encode()).digest()[:8], "little"))


class Snapshot:
    """Owned complete host state; call close() when its reverse pass is done."""

    def __init__(self, lib, handle, key, owner):
        self._lib, self._handle, self.key, self.owner = lib, handle, key, owner

    def nbytes(self):
        return int(self._lib.stlp_checkpoint_bytes(self._handle)) if self._handle else 0

    def close(self):
        if self._handle:
            self._lib.stlp_checkpoint_free(self._handle)
            self._handle = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *_):
        self.close()


class NativeEngine:
    """One persistent GPU workspace with an immutable plan, resolution and batch.

    forward returns (complex values [batch, ncoeff], real spatial variances of
    the terms, empty [batch, 0] unless moments=True); vjp returns the gradient
    of Re sum(conj(cotangents) * values) with respect to the maps [batch, npix].
    forward_device / vjp_device take and return device arrays (torch tensors or
    raw device addresses) without host copies of the maps.

    cross=True: the maps are pairs [batch, 2, npix] and the plan holds the cross
    terms of the two channels only (S1, S2, S3, S4 with c1 != c2; S3/S4 in both
    orders); see forward_cross / vjp_cross (reference native.py:160-162, 431-447).
    masks (or set_masks): averages over each mask at the coefficient's level;
    values, variances and cotangents [batch, K, ncoeff].
    """

    def __init__(self, nside, j, batch=1, backend="gpu", kernel=5, orientations=4, block_pixels=4096,
                 cache_bytes=1 << 30, threads=1, cache="transposition", device=0, library=None, cross=False,
                 masks=None, spin=0, spin_kernel=5, spin_pyramid="scalar", spin_a=2, precision="fp64",
                 timing=False, conv_variant=None, families=None):
        self._lib = _library(library)
        self.library_sha256 = self._lib._stlp_binary_sha256
        self._handle = None
        self._lock = threading.RLock()
        if backend not in ("cpu", "gpu"):
            raise ValueError("backend must be cpu or gpu")
        if any(isinstance(v, bool) or int(v) != v or int(v) <= 0
               for v in (nside, j, batch, kernel, orientations, block_pixels, threads)):
            raise ValueError("engine sizes/thread count must be positive integers")
        if int(cache_bytes) < 0 or int(device) < 0:
            raise ValueError("cache size/device must be nonnegative")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be fp64 or mixed")
        self.nside, self.j, self.batch = int(nside), int(j), int(batch)
        self.npix = 12 * self.nside ** 2
        self.backend = backend
        self.cross = bool(cross)
        self.spin = int(spin)
        self.precision = precision
        self.device = int(device)
        self.config = dict(nside=self.nside, j=self.j, batch=self.batch, backend=backend, kernel=int(kernel),
                           orientations=int(orientations), block_pixels=int(block_pixels),
                           cache_bytes=int(cache_bytes), threads=int(threads), cache=cache, device=self.device,
                           cross=self.cross, spin=self.spin, precision=precision, timing=bool(timing))
        if conv_variant is not None:
            self.config["conv_variant"] = int(conv_variant)
        if families is not None:
            self.config["families"] = ",".join(families)
        self._handle = self._lib.stlp_new(json.dumps(self.config).encode())
        if not self._handle:
            raise _error(self._lib)
        self._describe()
        if masks is not None:
            self.set_masks(masks)

    def _describe(self):
        self.plan = self.info()["plan"]
        self.plan_digest = hashlib.sha256(json.dumps(self.plan, sort_keys=True).encode()).hexdigest()
        self.ncoeff = len(self.plan)
        self._current_key = None
        self.nmasks = 0

    def share(self, batch=None, cross=None, families=None):
        """Another engine on this one's level fields and operators (the reference's
        GpuEvaluator::share, crates/stl-gpu/src/engine.rs:1909-1916): its own batch (at most
        this engine's maps), cross flag, plan families and masks, no field memory of its own.
        The engines take turns: a vjp needs the last forward on the fields to be its own."""
        with self._lock:
            self._open()
            cfg = {}
            if batch is not None:
                cfg["batch"] = int(batch)
            if cross is not None:
                cfg["cross"] = bool(cross)
            if families is not None:
                cfg["families"] = ",".join(families)
            h = self._lib.stlx_share(self._handle, json.dumps(cfg).encode())
            if not h:
                raise _error(self._lib)
            other = object.__new__(NativeEngine)
            other.__dict__.update({k: v for k, v in self.__dict__.items() if k not in ("_handle", "_lock")})
            other._lock = threading.RLock()
            other._handle = h
            other.batch = int(batch) if batch is not None else self.batch
            other.cross = bool(cross) if cross is not None else self.cross
            other.config = dict(self.config, batch=other.batch, cross=other.cross)
            other._describe()
            return other

    def _open(self):
        if not self._handle:
            raise RuntimeError("native engine is closed")

    def set_masks(self, masks):
        """Average over the masks from now on (None: the whole sky).  masks: a list of NESTED id
        arrays, or [K, npix] / [npix] indicators (nonzero: in); level n + 1 holds the parents of
        the level-n mask (reference native.py:243-264, mask.rs:43-70).  Values, variances and
        cotangents are then [batch, K, ncoeff]."""
        with self._lock:
            self._open()
            if masks is None:
                w = np.zeros((0, self.npix))
            elif (isinstance(masks, np.ndarray) and masks.dtype != object and masks.shape[-1] == self.npix
                  and masks.ndim in (1, 2)):
                w = np.ascontiguousarray(np.atleast_2d(masks), dtype=np.float64)
            else:
                w = np.zeros((len(masks), self.npix))
                for k, ids in enumerate(masks):
                    ids = np.asarray(ids, dtype=np.int64)
                    if ids.size and (ids.min() < 0 or ids.max() >= self.npix):
                        raise ValueError(f"mask {k}: pixel ids out of range")
                    w[k, ids] = 1.0
            if not np.isfinite(w).all():
                raise ValueError("mask weights must be finite")
            self._current_key = None
            if self._lib.stlp_set_masks(self._handle, _pointer(w), w.shape[0]):
                raise _error(self._lib)
            self.nmasks = w.shape[0]

    def native_pixels(self):
        """Pixels each coefficient averages: [K, ncoeff] with masks, else [ncoeff]."""
        with self._lock:
            self._open()
            out = np.empty((max(self.nmasks, 1), self.ncoeff), dtype=np.uintp)
            if self._lib.stlp_native_pixels(self._handle, out.ctypes.data_as(_SIZE)):
                raise _error(self._lib)
            out = out.astype(np.int64)
            return out if self.nmasks else out[0]

    def _shape(self):
        return (self.batch, self.nmasks, self.ncoeff) if self.nmasks else (self.batch, self.ncoeff)

    def _map_shape(self):
        return (self.batch, 2, self.npix) if self.cross else (self.batch, self.npix)

    def _map_len(self):
        return self.batch * (2 if self.cross else 1) * self.npix

    def info(self):
        with self._lock:
            self._open()
            p = self._lib.stlp_info(self._handle)
            if not p:
                raise _error(self._lib)
            try:
                result = json.loads(C.string_at(p))
                result["library_sha256"] = self.library_sha256
                return result
            finally:
                self._lib.stlp_string_free(p)

    def forward(self, maps, generation=0, member=0, objective="gaussian", moments=False):
        """Statistics of a batch of maps [batch, npix] ([npix] when batch == 1), keeping the state for vjp."""
        with self._lock:
            self._open()
            key = _key(generation, member, objective)
            x = np.ascontiguousarray(maps, dtype=np.float64)
            shape = self._map_shape()
            if x.shape == shape[1:] and self.batch == 1:
                x = x[None]
            if x.shape != shape:
                raise ValueError(f"expected maps {shape}, got {x.shape}")
            if not np.isfinite(x).all():
                raise ValueError("input map contains non-finite values")
            values = np.empty(self._shape(), dtype=np.complex128)
            variance = np.empty(self._shape() if moments else (self.batch, 0), dtype=np.float64)
            self._current_key = None
            status = self._lib.stlp_forward(self._handle, _pointer(x), x.size, *key, moments,
                                           _pointer(values.view(np.float64)), _pointer(variance))
            if status:
                raise _error(self._lib)
            self._current_key = key
            return values, variance

    def _cot(self, cotangents, key):
        if key != self._current_key:
            raise RuntimeError("stale/wrong-member/objective forward state")
        c = np.ascontiguousarray(cotangents, dtype=np.complex128)
        shape = self._shape()
        if c.shape == shape[1:] and self.batch == 1:
            c = c[None]
        if c.shape != shape:
            raise ValueError(f"expected cotangents {shape}, got {c.shape}")
        if not np.isfinite(c).all():
            raise ValueError("non-finite coefficient cotangent")
        return c

    def vjp(self, cotangents, generation=0, member=0, objective="gaussian"):
        """Gradient of Re sum(conj(cotangents) * values) with respect to the maps, [batch, npix]."""
        with self._lock:
            self._open()
            key = _key(generation, member, objective)
            c = self._cot(cotangents, key)
            grad = np.empty(self._map_shape(), dtype=np.float64)
            if self._lib.stlp_vjp(self._handle, _pointer(c.view(np.float64)), c.size, *key, _pointer(grad)):
                raise _error(self._lib)
            return grad

    def forward_device(self, maps, generation=0, member=0, objective="gaussian", moments=False):
        """forward() on a float64 device tensor [batch, npix] (torch) or a raw device address."""
        with self._lock:
            self._open()
            key = _key(generation, member, objective)
            addr, size = _device_address(maps, self._map_len())
            values = np.empty(self._shape(), dtype=np.complex128)
            variance = np.empty(self._shape() if moments else (self.batch, 0), dtype=np.float64)
            self._current_key = None
            if self._lib.stlp_forward_device(self._handle, addr, size, *key, moments,
                                             _pointer(values.view(np.float64)), _pointer(variance)):
                raise _error(self._lib)
            self._current_key = key
            return values, variance

    def vjp_device(self, cotangents, out, generation=0, member=0, objective="gaussian"):
        """vjp() into a float64 device tensor `out` [batch, npix] (or a raw device address)."""
        with self._lock:
            self._open()
            key = _key(generation, member, objective)
            c = self._cot(cotangents, key)
            addr, _ = _device_address(out, self._map_len())
            if self._lib.stlp_vjp_device(self._handle, _pointer(c.view(np.float64)), c.size, *key, addr):
                raise _error(self._lib)
            return out

    def forward_levels(self, levels):
        """Statistics of maps given at every level (no decimation; for validation)."""
        with self._lock:
            self._open()
            x = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64).ravel() for v in levels]))
            values = np.empty(self._shape(), dtype=np.complex128)
            self._current_key = None
            if self._lib.stlp_forward_levels(self._handle, _pointer(x), x.size, _pointer(values.view(np.float64))):
                raise _error(self._lib)
            return values

    def forward_cross(self, a, b, generation=0, member=0, objective="gaussian", moments=False):
        """Cross statistics S(a, b) of the pairs a, b [batch, npix] (cross=True)."""
        if not self.cross:
            raise RuntimeError("forward_cross needs NativeEngine(..., cross=True)")
        a = np.asarray(a, dtype=np.float64).reshape(self.batch, self.npix)
        b = np.asarray(b, dtype=np.float64).reshape(self.batch, self.npix)
        return self.forward(np.stack([a, b], 1), generation, member, objective, moments)

    def vjp_cross(self, cotangents, generation=0, member=0, objective="gaussian"):
        """(grad_a, grad_b), each [batch, npix], after forward_cross."""
        g = self.vjp(cotangents, generation, member, objective)
        return g[:, 0], g[:, 1]

    def checkpoint(self):
        with self._lock:
            self._open()
            p = self._lib.stlp_checkpoint(self._handle)
            if not p:
                raise _error(self._lib)
            return Snapshot(self._lib, p, self._current_key, self._handle)

    def restore(self, snapshot, generation=0, member=0, objective="gaussian"):
        with self._lock:
            self._open()
            key = _key(generation, member, objective)
            if not isinstance(snapshot, Snapshot) or not snapshot._handle:
                raise ValueError("checkpoint is closed or invalid")
            if snapshot.key != key:
                raise RuntimeError("stale/wrong-member/objective checkpoint state")
            if snapshot._lib is not self._lib or snapshot.owner != self._handle:
                raise ValueError("checkpoint engine/plan owner mismatch")
            if self._lib.stlp_restore(self._handle, snapshot._handle, *key):
                raise _error(self._lib)
            self._current_key = key

    def synthesize(self, x, target, weights=None, lr=1.0, max_iter=20, max_eval=None, tolerance_grad=1e-12,
                   tolerance_change=1e-15, history_size=10, ensemble=False):
        """L-BFGS on the device of sum_k w_k |S_k(x) - target_k|^2 from the float64 device tensor x
        [batch, npix] (updated in place); returns the loss of every evaluation.  With `ensemble` the
        batch is one ensemble: the loss is on the mean statistics of the members, target [ncoef]
        (the reference's stl-synth --ensemble, crates/stl-mpi/src/synthesis.rs:362-470, no scatter term)."""
        with self._lock:
            self._open()
            addr, size = _device_address(x, self._map_len())
            shape = (self.ncoeff,) if ensemble else self._shape()
            t = np.ascontiguousarray(np.asarray(target, dtype=np.complex128).reshape(shape))
            w = None if weights is None else np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(-1))
            if w is not None and w.size != self.ncoeff:
                raise ValueError(f"expected {self.ncoeff} weights")
            opts = dict(lr=lr, max_iter=max_iter, tolerance_grad=tolerance_grad, tolerance_change=tolerance_change,
                        history_size=history_size, ensemble=bool(ensemble))
            if max_eval is not None:
                opts["max_eval"] = max_eval
            cap = (max_eval or max_iter * 5 // 4) + 64
            hist = np.zeros(cap)
            self._current_key = None
            n = self._lib.stlx_synthesize(self._handle, addr, size, _pointer(t.view(np.float64)),
                                          _pointer(w) if w is not None else None, json.dumps(opts).encode(),
                                          _pointer(hist), cap)
            if n < 0:
                raise _error(self._lib)
            return hist[:min(n, cap)].copy()

    def apply_operator(self, kind, level, x, out):
        """Test hook: one level's operator on float64 device tensors (see stlx_apply_operator)."""
        with self._lock:
            self._open()
            rows = int(x.shape[0])
            if self._lib.stlx_apply_operator(self._handle, int(kind), int(level), rows, int(x.data_ptr()),
                                             int(out.data_ptr())):
                raise _error(self._lib)
            return out

    def close(self):
        if getattr(self, "_handle", None):
            self._lib.stlp_free(self._handle)
            self._handle = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *_):
        self.close()


class NativeEnsemble:
    """Maps of uneven nside evaluated together on one GPU: groups of
    (nside, batch, j) (j defaults to log2(nside) - 1), each forward and each
    VJP one HIP graph per group, the groups on concurrent streams.  forward_device takes one float64
    device tensor [batch, 12 nside^2] per group and returns one complex
    statistics array [batch, ncoeff_g] per group; vjp_device takes the
    cotangents in the same shapes and writes the gradients into `out`."""

    def __init__(self, groups, precision="mixed", device=0, library=None):
        self._lib = _library(library)
        self._handle = None
        self._lock = threading.RLock()
        norm = []
        for g in groups:
            nside, batch = int(g[0]), int(g[1])
            j = int(g[2]) if len(g) > 2 else nside.bit_length() - 2
            norm.append((nside, batch, j))
        if not norm:
            raise ValueError("an ensemble needs at least one group")
        self.groups = norm
        arr = lambda k: np.ascontiguousarray([g[k] for g in norm], dtype=np.int32)
        self._nside, self._batch, self._j = arr(0), arr(1), arr(2)
        opts = json.dumps(dict(precision=precision, device=int(device))).encode()
        self._handle = self._lib.stlx_ensemble_new(self._nside.ctypes.data_as(_INT32), self._j.ctypes.data_as(_INT32),
                                                   self._batch.ctypes.data_as(_INT32), len(norm), opts)
        if not self._handle:
            raise _error(self._lib)
        nc = np.zeros(len(norm), dtype=np.uintp)
        if self._lib.stlx_ensemble_ncoef(self._handle, nc.ctypes.data_as(_SIZE)):
            raise _error(self._lib)
        self.ncoeff = [int(v) for v in nc]
        self.npix = [12 * g[0] ** 2 for g in norm]
        self._sizes = [g[1] * n for g, n in zip(norm, self.ncoeff)]

    def _addresses(self, arrays):
        if len(arrays) != len(self.groups):
            raise ValueError(f"expected {len(self.groups)} arrays, one per group")
        out = np.zeros(len(arrays), dtype=np.uint64)
        for k, (a, g) in enumerate(zip(arrays, self.groups)):
            out[k], _ = _device_address(a, g[1] * 12 * g[0] ** 2)
        return out

    def forward_device(self, maps, moments=False):
        with self._lock:
            if not self._handle:
                raise RuntimeError("native ensemble is closed")
            addr = self._addresses(maps)
            n = sum(self._sizes)
            values = np.empty(n, dtype=np.complex128)
            var = np.empty(n if moments else 1, dtype=np.float64)
            if self._lib.stlx_ensemble_forward_device(self._handle, addr.ctypes.data_as(_UINT64), len(addr), moments,
                                                      _pointer(values.view(np.float64)), _pointer(var)):
                raise _error(self._lib)
            split = np.cumsum(self._sizes)[:-1]
            vals = [v.reshape(g[1], -1) for v, g in zip(np.split(values, split), self.groups)]
            if not moments:
                return vals
            return vals, [v.reshape(g[1], -1) for v, g in zip(np.split(var, split), self.groups)]

    def vjp_device(self, cotangents, out):
        with self._lock:
            if not self._handle:
                raise RuntimeError("native ensemble is closed")
            c = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.complex128).ravel() for v in cotangents]))
            if c.size != sum(self._sizes):
                raise ValueError("cotangent shapes do not match the ensemble's statistics")
            addr = self._addresses(out)
            if self._lib.stlx_ensemble_vjp_device(self._handle, _pointer(c.view(np.float64)), c.size,
                                                  addr.ctypes.data_as(_UINT64), len(addr)):
                raise _error(self._lib)
            return out

    def close(self):
        if getattr(self, "_handle", None):
            self._lib.stlx_ensemble_free(self._handle)
            self._handle = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *_):
        self.close()


def ergodic_weights(native_pixels, variance):
    """Ergodic weights 1 / (n sigma_i^2), sigma_i^2 = V_i / N_i (reference
    crates/stl-mpi/src/synthesis.rs:296-314): V_i the spatial variance of the
    terms of coefficient i on the target (forward(..., moments=True)), N_i the
    pixels it averages (NativeEngine.native_pixels), n the number of
    coefficients; a zero variance gets the largest finite weight."""
    v = np.asarray(variance, dtype=np.float64).reshape(-1)
    n_i = np.asarray(native_pixels, dtype=np.float64).reshape(-1)
    if v.size != n_i.size:
        raise ValueError("one variance per coefficient expected")
    w = np.full(v.shape, np.inf)
    pos = v > 0
    w[pos] = n_i[pos] / v[pos]
    finite = np.isfinite(w)
    w[~finite] = w[finite].max() if finite.any() else 0.0
    return w / w.size


def native_lbfgs(callback, x, options=None, library=None):
    """One L-BFGS minimization (strong Wolfe) through stlp_lbfgs (reference
    native.py:1007-1052): x (float64, updated in place when C-contiguous
    float64) is the start; callback(x_view) -> (loss, gradient of x's shape)
    gets a read-only view valid during the call.  Returns (x, report)."""
    lib = _library(library)
    resolved = dict(lr=1.0, max_iter=20, max_eval=None, tolerance_grad=1e-12, tolerance_change=1e-15,
                    history_size=10)
    if options:
        unknown = set(options) - set(resolved)
        if unknown:
            raise ValueError(f"unknown native L-BFGS options: {sorted(unknown)}")
        resolved.update(options)
    for k in ("max_iter", "history_size"):
        if isinstance(resolved[k], bool) or int(resolved[k]) != resolved[k] or resolved[k] < 1:
            raise ValueError(f"{k} must be a positive integer")
    if resolved["max_eval"] is not None and (int(resolved["max_eval"]) != resolved["max_eval"]
                                             or resolved["max_eval"] < 1):
        raise ValueError("max_eval must be a positive integer or None")
    x = np.ascontiguousarray(x, dtype=np.float64)
    if x.size == 0 or not np.isfinite(x).all():
        raise ValueError("finite nonempty initial ensemble required")
    shape, history, errors = x.shape, [], []

    @_CALLBACK
    def evaluate(p, n, gp, _context):
        try:
            trial = np.ctypeslib.as_array(p, shape=(n,)).reshape(shape)
            trial.flags.writeable = False
            value, grad = callback(trial)
            grad = np.asarray(grad, dtype=np.float64)
            if grad.shape != shape:
                raise ValueError(f"gradient shape {grad.shape} != {shape}")
            np.ctypeslib.as_array(gp, shape=(n,))[:] = grad.ravel()
            history.append(float(value))
            return float(value)
        except BaseException as exc:  # reported to the library as a non-finite loss
            errors.append(exc)
            return float("nan")

    opts = {k: v for k, v in resolved.items() if v is not None}
    count = lib.stlp_lbfgs(_pointer(x), x.size, json.dumps(opts).encode(), evaluate, None)
    if errors:
        raise errors[0]
    if count < 0:
        raise _error(lib)
    return x, dict(evaluations=int(count), loss_history=history, options=resolved, optimizer="native L-BFGS",
                   line_search="strong_wolfe")


def _device_address(a, expected):
    """(address, element count) of a contiguous float64 device tensor or a raw address."""
    if isinstance(a, int):
        return a, expected
    try:
        import torch
    except ImportError as exc:  # pragma: no cover
        raise TypeError("device arrays are torch tensors or integer addresses") from exc
    if not isinstance(a, torch.Tensor) or not a.is_cuda:
        raise TypeError("expected a device (cuda) tensor")
    if a.dtype != torch.float64 or not a.is_contiguous():
        raise TypeError("expected a contiguous float64 device tensor")
    if a.numel() != expected:
        raise ValueError(f"expected {expected} elements, got {a.numel()}")
    # Work queued by torch on its stream must be complete before the library reads or writes.
    torch.cuda.current_stream(a.device).synchronize()
    return int(a.data_ptr()), int(a.numel())
