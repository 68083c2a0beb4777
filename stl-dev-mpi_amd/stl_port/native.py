"""Python face of libstl_bridge.so, the MI355X ScatCov engine.

Drop-in for the reference binding's engine API (python/stl_ensemble/stl_port/native.py:
NativeEngine, Snapshot, native_lbfgs): the same constructor keywords, array shapes, identity
rule and error texts, over this library's GPU engine.  The module is organised around one
`Bridge` object per loaded library:

- `Bridge.call(name, *args)` runs a C entry point and turns its failure value (-1, or NULL for
  the handle-returning ones) into `BridgeError` carrying `stlp_last_error()`;
- engines own a handle through `_Owned`, which frees it exactly once;
- every public method of an engine runs under the engine's lock (one caller thread per handle
  is the library's rule) after an open check (`_engine_call`).

Maps and gradients on the device (torch tensors or raw addresses) go through the `*_device`
methods without host copies.  The library has no CPU path and this module has no fallback: a
missing library raises at construction.  Keywords of this library beyond the reference's:
`precision` ("fp64" | "mixed"), `fields` ("keep" | "recompute" | "auto"), `timing`,
`conv_variant`, `families`, `batch_nside` (levels of at most this nside run batched, each step
of a sweep one launch for all of them; 0: none).
"""
from __future__ import annotations

import ctypes as C
import functools
import hashlib
import json
import os
import threading
from pathlib import Path

import numpy as np

DEFAULT_LIBRARY = Path(__file__).resolve().parents[1] / "libstl_bridge.so"

# ---------------------------------------------------------------------------------------------
# The C ABI as data.  One letter per C type: the first letter is the return type, the rest the
# arguments.  p void*, s char* (in), r char* (returned, owned by the library), d double*,
# f float*, z size_t, Z size_t*, u uint64, U uint64*, Q int64*, l int64, i int, I int32*,
# w uint32, W uint32*, b bool, g double, c the L-BFGS callback, - void.
# The stlp_* rows are the names the reference binding binds (native.py:55-108); stlx_* are
# this library's additions.
_CALLBACK = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_double), C.c_void_p)
_CTYPES = {
    "p": C.c_void_p, "s": C.c_char_p, "r": C.c_char_p, "d": C.POINTER(C.c_double), "f": C.POINTER(C.c_float),
    "z": C.c_size_t, "Z": C.POINTER(C.c_size_t), "u": C.c_uint64, "U": C.POINTER(C.c_uint64),
    "Q": C.POINTER(C.c_int64), "l": C.c_int64, "i": C.c_int, "I": C.POINTER(C.c_int32), "w": C.c_uint32,
    "W": C.POINTER(C.c_uint32), "b": C.c_bool, "g": C.c_double, "c": _CALLBACK, "-": None,
}
ABI = {
    "stlp_last_error": "r", "stlp_string_free": "-p", "stlp_new": "ps", "stlp_free": "-p", "stlp_info": "pp",
    "stlp_forward": "ipdzuuubdd", "stlp_vjp": "ipdzuuud",
    "stlp_forward_device": "ipuzuuubdd", "stlp_vjp_device": "ipdzuuuu",
    "stlp_checkpoint": "pp", "stlp_checkpoint_free": "-p", "stlp_checkpoint_bytes": "zp", "stlp_restore": "ippuuu",
    "stlp_set_masks": "ipdz", "stlp_spin_fields": "ipdzzdd", "stlp_forward_levels": "ipdzd",
    "stlp_native_pixels": "ipZ", "stlp_lbfgs": "ldzscp",
    "stlp_patch_new": "ps", "stlp_patch_free": "-p", "stlp_patch_info": "pp", "stlp_pixel_vectors": "izQzd",
    "stlp_patch_select": "ipdzggZ", "stlp_patch_ids": "ipzQ", "stlp_patch_core_pixels": "ipZ",
    "stlp_patch_forward": "ipdzubdd", "stlp_patch_vjp": "ipdzud", "stlp_patch_forward_device": "ipuzubdd",
    "stlp_patch_vjp_device": "ipdzuu",
    "stlp_spec_new": "ps", "stlp_spec_free": "-p", "stlp_spec_info": "pp", "stlp_spec_set_masks": "ipdz",
    "stlp_spec_forward": "ipdzuuud", "stlp_spec_vjp": "ipdzuuud",
    "stlp_unet_new": "ps", "stlp_unet_free": "-p", "stlp_unet_info": "pp", "stlp_unet_get_params": "ipf",
    "stlp_unet_set_params": "ipfz", "stlp_unet_get_grads": "ipf", "stlp_unet_zero_grad": "ip",
    "stlp_unet_adam": "ipg", "stlp_unet_set_cond": "ipfz", "stlp_unet_apply": "ipdzddd",
    "stlp_train_set_masks": "ipdz", "stlp_train_set_common": "ipd", "stlp_train_set_field": "ipwddddddzg",
    "stlp_train_bias": "ipwUdz", "stlp_train_loss": "ipwbd", "stlp_train_step": "ipwgd",
    "stlp_train_live": "ipwUdzUdzbgd", "stlp_train_lbfgs": "ppWzs", "stlp_train_set_joint": "ipddwwsg",
    "stlp_train_front": "ipdzbd",
    "stlx_share": "pps", "stlx_apply_operator": "ipiiiuu", "stlx_apply_operator_noalloc": "ipiiiuu", "stlx_synthesize": "lpuzddsdz",
    "stlx_ensemble_new": "pIIIzs", "stlx_ensemble_free": "-p", "stlx_ensemble_ncoef": "ipZ",
    "stlx_ensemble_forward_device": "ipUzbdd", "stlx_ensemble_vjp_device": "ipdzUz",
    "stlx_map_len": "isZ", "stlx_read_map": "isdz", "stlx_read_map_range": "iszzd", "stlx_read_npy": "isdzZZ",
    "stlx_write_npy": "isdZz", "stlx_write_map": "isdzsb", "stlx_sliced_create": "iszsb",
    "stlx_sliced_write": "isbzdz", "stlx_pixel_normal": "iulzd", "stlx_stream_copy": "iuuzid",
}


class BridgeError(RuntimeError):
    """A failure reported by the library (its stlp_last_error message)."""


class Bridge:
    """One loaded libstl_bridge.so with its entry points typed from `ABI`."""

    _loaded: dict = {}
    _guard = threading.Lock()

    def __init__(self, path: str):
        self.path = path
        self.cdll = C.CDLL(path)
        for name, code in ABI.items():
            fn = getattr(self.cdll, name)
            fn.restype = _CTYPES[code[0]]
            fn.argtypes = [_CTYPES[k] for k in code[1:]]
        self._sha256 = None

    @classmethod
    def open(cls, path=None) -> "Bridge":
        resolved = library_path(path)
        with cls._guard:
            if resolved not in cls._loaded:
                if not os.path.isfile(resolved):
                    raise RuntimeError(f"Native STL library is missing: {resolved}. Build it with "
                                       f"__graft_entry__.build(); no fallback scattering implementation is used.")
                cls._loaded[resolved] = cls(resolved)
            return cls._loaded[resolved]

    @property
    def sha256(self) -> str:
        """SHA-256 of the loaded binary (hex), computed on first use."""
        if self._sha256 is None:
            self._sha256 = hashlib.sha256(Path(self.path).read_bytes()).hexdigest()
        return self._sha256

    def last_error(self) -> str:
        msg = self.cdll.stlp_last_error()
        return msg.decode("utf-8", errors="replace") if msg else "native failure"

    def error(self) -> BridgeError:
        return BridgeError(self.last_error())

    def call(self, name, *args):
        """Run an entry point; its failure value raises BridgeError."""
        out = getattr(self.cdll, name)(*args)
        code = ABI[name][0]
        failed = (code in "il" and out < 0) or (code == "p" and not out)
        if failed:
            raise self.error()
        return out

    def json(self, name, handle):
        """A JSON string returned by the library (freed with stlp_string_free)."""
        ptr = self.call(name, handle)
        try:
            return json.loads(C.string_at(ptr))
        finally:
            self.cdll.stlp_string_free(ptr)


def library_path(path=None) -> str:
    """The library a binding loads: `path`, else $STL_NATIVE_LIBRARY, else the in-tree build."""
    chosen = path if path is not None else os.environ.get("STL_NATIVE_LIBRARY", DEFAULT_LIBRARY)
    return str(Path(chosen).resolve())


def bridge(path=None) -> Bridge:
    return Bridge.open(path)


# ---------------------------------------------------------------------------------------------
# Arrays and identities

def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_CTYPES["d"])


def _host_f64(values, shape, what, allow_single=True):
    """values as a C-contiguous float64 array of `shape` ([1:] accepted for one map)."""
    a = np.ascontiguousarray(values, dtype=np.float64)
    if allow_single and shape[0] == 1 and a.shape == tuple(shape[1:]):
        a = a.reshape(shape)
    if a.shape != tuple(shape):
        raise ValueError(f"expected {what} {tuple(shape)}, got {a.shape}")
    return a


def _device_arg(a, count):
    """(address, count) of a contiguous float64 device tensor, or a raw address given as int.
    Work torch queued on the tensor's stream is finished first: the library uses its own."""
    if isinstance(a, (int, np.integer)) and not isinstance(a, bool):
        return int(a), count
    import torch
    if not (isinstance(a, torch.Tensor) and a.is_cuda):
        raise TypeError("expected a device (cuda) tensor or an integer device address")
    if a.dtype != torch.float64 or not a.is_contiguous():
        raise TypeError("expected a contiguous float64 device tensor")
    if a.numel() != count:
        raise ValueError(f"expected {count} elements, got {a.numel()}")
    torch.cuda.current_stream(a.device).synchronize()
    return int(a.data_ptr()), count


def _u64(value, what):
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, np.integer)) or not 0 <= value < 1 << 64:
        raise ValueError(f"{what} must be an unsigned 64-bit {'evaluation' if what == 'generation' else 'member'} "
                         f"number")
    return int(value)


def objective_id(name: str) -> int:
    """64-bit id of an objective name: the first 8 bytes of its SHA-256, little endian (the
    reference binding's rule, native.py:122-129, so ids agree across bindings)."""
    if not isinstance(name, str) or not name:
        raise ValueError("objective must be a nonempty name")
    return int.from_bytes(hashlib.sha256(name.encode()).digest()[:8], "little")


def _key(generation, member, objective):
    """(generation, member, objective id) identity of a forward state."""
    return _u64(generation, "generation"), _u64(member, "member"), objective_id(objective)


class _Owned:
    """A library object freed exactly once by `free_name`."""

    def __init__(self, lib: Bridge, ptr, free_name: str):
        self.lib, self.ptr, self._free = lib, ptr, free_name

    def release(self):
        ptr, self.ptr = self.ptr, None
        if ptr:
            getattr(self.lib.cdll, self._free)(ptr)

    def __bool__(self):
        return bool(self.ptr)


def _engine_call(method):
    """Run under the engine's lock, after checking it is open."""
    @functools.wraps(method)
    def run(self, *args, **kwargs):
        with self._lock:
            if not self._h:
                raise RuntimeError("native engine is closed")
            return method(self, *args, **kwargs)
    return run


class Snapshot:
    """A host copy of an engine's forward state (stlp_checkpoint); close() frees it.  `key` is
    the identity it was taken under, `owner` the engine's handle address."""

    def __init__(self, lib, ptr, key, owner):
        self._obj = _Owned(lib, ptr, "stlp_checkpoint_free")
        self.key, self.owner = key, owner

    @property
    def valid(self):
        return bool(self._obj)

    def nbytes(self):
        return int(self._obj.lib.cdll.stlp_checkpoint_bytes(self._obj.ptr)) if self._obj else 0

    def close(self):
        self._obj.release()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------------------------
# Engines

_CONSTRUCTOR_SIZES = ("nside", "j", "batch", "kernel", "orientations", "block_pixels", "threads")


class NativeEngine:
    """The ScatCov statistics of a batch of HEALPix NESTED maps and their VJP, on one GPU.

    forward(maps) -> (values complex [batch, ncoeff], spatial variances of the terms [batch,
    ncoeff] with moments=True, else an empty [batch, 0]); vjp(cotangents) -> the gradient of
    Re sum(conj(cotangents) * values) with respect to the maps [batch, npix], for the forward of
    the same (generation, member, objective).  forward_device / vjp_device take float64 device
    tensors (or addresses) instead of host maps.

    cross=True: pairs [batch, 2, npix] and the cross plan (S1, S2, S3, S4 with c1 != c2).
    masks / set_masks: averages over K masks at each coefficient's native level; values,
    variances and cotangents become [batch, K, ncoeff].
    share(): another plan or batch on this engine's fields (no copy of them).
    """

    def __init__(self, nside, j, batch=1, backend="gpu", kernel=5, orientations=4, block_pixels=4096,
                 cache_bytes=1 << 30, threads=1, cache="transposition", device=0, library=None, cross=False,
                 masks=None, spin=0, spin_kernel=5, spin_pyramid="scalar", spin_a=2, precision="fp64",
                 fields="auto", timing=False, conv_variant=None, families=None, batch_nside=None):
        self._lock = threading.RLock()
        self._h = None
        self._lib = bridge(library)
        sizes = dict(nside=nside, j=j, batch=batch, kernel=kernel, orientations=orientations,
                     block_pixels=block_pixels, threads=threads)
        if backend not in ("cpu", "gpu"):
            raise ValueError("backend must be cpu or gpu")
        if not all(_positive_int(v) for v in sizes.values()):
            raise ValueError("engine sizes/thread count must be positive integers")
        if int(cache_bytes) < 0 or int(device) < 0:
            raise ValueError("cache size/device must be nonnegative")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be fp64 or mixed")
        if fields not in ("keep", "recompute", "auto"):
            raise ValueError("fields must be keep, recompute or auto")
        config = {k: int(v) for k, v in sizes.items()}
        config.update(backend=backend, cache_bytes=int(cache_bytes), cache=cache, device=int(device),
                      cross=bool(cross), spin=int(spin), precision=precision, fields=fields, timing=bool(timing))
        if conv_variant is not None:
            config["conv_variant"] = int(conv_variant)
        if families is not None:
            config["families"] = ",".join(families)
        if batch_nside is not None:
            config["batch_nside"] = int(batch_nside)
        self._adopt(self._lib.call("stlp_new", json.dumps(config).encode()), config)
        if masks is not None:
            self.set_masks(masks)

    def _adopt(self, handle, config):
        """Take ownership of a new handle and read its plan."""
        self._h = _Owned(self._lib, handle, "stlp_free")
        self.config = dict(config)
        self.nside, self.j, self.batch = config["nside"], config["j"], config["batch"]
        self.npix = 12 * self.nside ** 2
        self.cross = bool(config["cross"])
        self.backend, self.precision, self.device = config["backend"], config["precision"], config["device"]
        self.spin = config.get("spin", 0)
        self.plan = self.info()["plan"]
        self.plan_digest = hashlib.sha256(json.dumps(self.plan, sort_keys=True).encode()).hexdigest()
        self.ncoeff = len(self.plan)
        self.nmasks = 0
        self._state = None  # identity of the forward state the handle holds

    @property
    def library_sha256(self):
        return self._lib.sha256

    # shapes -------------------------------------------------------------------------------
    def _shape(self):
        """Statistics / cotangent shape."""
        return (self.batch, self.nmasks, self.ncoeff) if self.nmasks else (self.batch, self.ncoeff)

    def _map_shape(self):
        return (self.batch, 2, self.npix) if self.cross else (self.batch, self.npix)

    def _map_count(self):
        return int(np.prod(self._map_shape()))

    def _outputs(self, moments):
        values = np.empty(self._shape(), dtype=np.complex128)
        variance = np.empty(self._shape() if moments else (self.batch, 0), dtype=np.float64)
        return values, variance

    def _cotangents(self, cotangents, key):
        if key != self._state:
            raise RuntimeError("stale/wrong-member/objective forward state")
        c = np.ascontiguousarray(cotangents, dtype=np.complex128)
        shape = self._shape()
        if shape[0] == 1 and c.shape == shape[1:]:
            c = c.reshape(shape)
        if c.shape != shape:
            raise ValueError(f"expected cotangents {shape}, got {c.shape}")
        if not np.isfinite(c).all():
            raise ValueError("non-finite coefficient cotangent")
        return c

    # configuration ------------------------------------------------------------------------
    @_engine_call
    def info(self):
        out = self._lib.json("stlp_info", self._h.ptr)
        out["library_sha256"] = self._lib.sha256
        return out

    @_engine_call
    def share(self, batch=None, cross=None, families=None):
        """Another engine on this engine's level fields and operators (the reference's
        GpuEvaluator::share, crates/stl-gpu/src/engine.rs:1909-1916): its own batch (at most
        this engine's field rows), cross flag, plan families and masks.  The engines take turns
        on the fields: a vjp or checkpoint needs the last forward on them to be its own."""
        request = {}
        if batch is not None:
            request["batch"] = int(batch)
        if cross is not None:
            request["cross"] = bool(cross)
        if families is not None:
            request["families"] = ",".join(families)
        handle = self._lib.call("stlx_share", self._h.ptr, json.dumps(request).encode())
        other = NativeEngine.__new__(NativeEngine)
        other._lock, other._lib = threading.RLock(), self._lib
        config = dict(self.config, batch=request.get("batch", self.batch), cross=request.get("cross", self.cross))
        other._adopt(handle, config)
        return other

    @_engine_call
    def set_masks(self, masks):
        """Average over masks from the next forward on; None restores the whole sky.  masks:
        NESTED id arrays, or indicators [K, npix] / [npix] (nonzero: in the mask).  Level n + 1
        holds the parents of the level-n mask (reference mask.rs:43-70)."""
        weights = self._mask_weights(masks)
        self._state = None
        self._lib.call("stlp_set_masks", self._h.ptr, _dptr(weights), weights.shape[0])
        self.nmasks = weights.shape[0]

    def _mask_weights(self, masks):
        if masks is None:
            return np.zeros((0, self.npix))
        if isinstance(masks, np.ndarray) and masks.dtype != object and masks.ndim in (1, 2) \
                and masks.shape[-1] == self.npix:
            w = np.array(np.atleast_2d(masks), dtype=np.float64, order="C")
        else:
            w = np.zeros((len(masks), self.npix))
            for k, ids in enumerate(masks):
                ids = np.asarray(ids, dtype=np.int64).reshape(-1)
                if ids.size and not (0 <= ids.min() and ids.max() < self.npix):
                    raise ValueError(f"mask {k}: pixel ids out of range")
                w[k, ids] = 1.0
        if not np.isfinite(w).all():
            raise ValueError("mask weights must be finite")
        return w

    @_engine_call
    def native_pixels(self):
        """Pixels each coefficient averages over: [K, ncoeff] with masks, else [ncoeff]."""
        out = np.empty((max(self.nmasks, 1), self.ncoeff), dtype=np.uintp)
        self._lib.call("stlp_native_pixels", self._h.ptr, out.ctypes.data_as(_CTYPES["Z"]))
        out = out.astype(np.int64)
        return out if self.nmasks else out[0]

    # sweeps -------------------------------------------------------------------------------
    @_engine_call
    def forward(self, maps, generation=0, member=0, objective="gaussian", moments=False):
        """Statistics of host maps [batch, npix] ([npix] for one map); the state is kept for vjp."""
        key = _key(generation, member, objective)
        x = _host_f64(maps, self._map_shape(), "maps")
        if not np.isfinite(x).all():
            raise ValueError("input map contains non-finite values")
        values, variance = self._outputs(moments)
        self._state = None
        self._lib.call("stlp_forward", self._h.ptr, _dptr(x), x.size, *key, bool(moments),
                       _dptr(values.view(np.float64)), _dptr(variance))
        self._state = key
        return values, variance

    @_engine_call
    def vjp(self, cotangents, generation=0, member=0, objective="gaussian"):
        """Gradient [batch, npix] of Re sum(conj(cotangents) * values) after forward."""
        key = _key(generation, member, objective)
        c = self._cotangents(cotangents, key)
        grad = np.empty(self._map_shape(), dtype=np.float64)
        self._lib.call("stlp_vjp", self._h.ptr, _dptr(c.view(np.float64)), c.size, *key, _dptr(grad))
        return grad

    @_engine_call
    def forward_device(self, maps, generation=0, member=0, objective="gaussian", moments=False):
        """forward() of a float64 device tensor [batch, npix] (or its address)."""
        key = _key(generation, member, objective)
        addr, count = _device_arg(maps, self._map_count())
        values, variance = self._outputs(moments)
        self._state = None
        self._lib.call("stlp_forward_device", self._h.ptr, addr, count, *key, bool(moments),
                       _dptr(values.view(np.float64)), _dptr(variance))
        self._state = key
        return values, variance

    @_engine_call
    def vjp_device(self, cotangents, out, generation=0, member=0, objective="gaussian"):
        """vjp() written into the float64 device tensor `out` [batch, npix] (or an address)."""
        key = _key(generation, member, objective)
        c = self._cotangents(cotangents, key)
        addr, _ = _device_arg(out, self._map_count())
        self._lib.call("stlp_vjp_device", self._h.ptr, _dptr(c.view(np.float64)), c.size, *key, addr)
        return out

    @_engine_call
    def forward_levels(self, levels):
        """Statistics of maps given at every level (level d: [batch, 12 (nside / 2^d)^2]), with no
        decimation; no vjp may follow (validation entry point)."""
        flat = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64).reshape(-1) for v in levels]))
        values = np.empty(self._shape(), dtype=np.complex128)
        self._state = None
        self._lib.call("stlp_forward_levels", self._h.ptr, _dptr(flat), flat.size, _dptr(values.view(np.float64)))
        return values

    def forward_cross(self, a, b, generation=0, member=0, objective="gaussian", moments=False):
        """Cross statistics of the pairs (a, b), each [batch, npix]; needs cross=True."""
        if not self.cross:
            raise RuntimeError("forward_cross needs NativeEngine(..., cross=True)")
        pair = np.stack([np.asarray(a, dtype=np.float64).reshape(self.batch, self.npix),
                         np.asarray(b, dtype=np.float64).reshape(self.batch, self.npix)], axis=1)
        return self.forward(pair, generation, member, objective, moments)

    def vjp_cross(self, cotangents, generation=0, member=0, objective="gaussian"):
        """(grad_a, grad_b) after forward_cross."""
        g = self.vjp(cotangents, generation, member, objective)
        return g[:, 0], g[:, 1]

    @_engine_call
    def spin_fields(self, maps, level=0):
        """Spin-2 front end: not provided by this library (raises BridgeError)."""
        x = np.ascontiguousarray(maps, dtype=np.float64)
        out = np.empty(1)
        self._lib.call("stlp_spin_fields", self._h.ptr, _dptr(x), x.size, int(level), _dptr(out), _dptr(out))

    # state ---------------------------------------------------------------------------------
    @_engine_call
    def checkpoint(self):
        """Host copy of the current forward state, for restore() before a later vjp."""
        ptr = self._lib.call("stlp_checkpoint", self._h.ptr)
        return Snapshot(self._lib, ptr, self._state, self._h.ptr)

    @_engine_call
    def restore(self, snapshot, generation=0, member=0, objective="gaussian"):
        key = _key(generation, member, objective)
        if not isinstance(snapshot, Snapshot) or not snapshot.valid:
            raise ValueError("checkpoint is closed or invalid")
        if snapshot.key != key:
            raise RuntimeError("stale/wrong-member/objective checkpoint state")
        if snapshot._obj.lib is not self._lib or snapshot.owner != self._h.ptr:
            raise ValueError("checkpoint engine/plan owner mismatch")
        self._lib.call("stlp_restore", self._h.ptr, snapshot._obj.ptr, *key)
        self._state = key

    # device synthesis ----------------------------------------------------------------------
    @_engine_call
    def synthesize(self, x, target, weights=None, lr=1.0, max_iter=20, max_eval=None, tolerance_grad=1e-12,
                   tolerance_change=1e-15, history_size=10, ensemble=False, record="evaluations"):
        """Device L-BFGS of sum w |S(x) - target|^2 from the float64 device tensor x [batch, npix]
        (updated in place); returns the loss of every evaluation.

        target: [batch, K, ncoeff] (K masks; [batch, ncoeff] without), or with ensemble=True one
        row [K, ncoeff] for the mean statistics of the members (the reference's stl-synth
        --ensemble, crates/stl-mpi/src/synthesis.rs:362-470, without its scatter term).
        weights: one per mask and coefficient, [K, ncoeff] (the shape of native_pixels()),
        applied to every map; None weighs all 1.  record: "evaluations" (the loss of every
        evaluation) or "iterations" (the start's loss, then the loss at each iteration's
        accepted step)."""
        rows = max(self.nmasks, 1) * self.ncoeff
        addr, count = _device_arg(x, self._map_count())
        t = np.ascontiguousarray(np.asarray(target, dtype=np.complex128).reshape(-1))
        if t.size != (rows if ensemble else self.batch * rows):
            raise ValueError(f"expected a target of {rows if ensemble else self.batch * rows} coefficients")
        w = None
        if weights is not None:
            w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(-1))
            if w.size != rows:
                raise ValueError(f"expected {rows} weights (masks x coefficients)")
        options = dict(lr=lr, max_iter=max_iter, tolerance_grad=tolerance_grad, tolerance_change=tolerance_change,
                       history_size=history_size, ensemble=bool(ensemble), history=record)
        if max_eval is not None:
            options["max_eval"] = max_eval
        cap = max(max_eval or max_iter * 5 // 4, max_iter) + 64
        history = np.zeros(cap)
        self._state = None
        n = self._lib.call("stlx_synthesize", self._h.ptr, addr, count, _dptr(t.view(np.float64)),
                           None if w is None else _dptr(w), json.dumps(options).encode(), _dptr(history), cap)
        if record == "iterations":
            return history[~np.isnan(history)].copy()
        return history[:min(n, cap)].copy()

    @_engine_call
    def apply_operator(self, kind, level, x, out):
        """Test hook: one level's operator on float64 device tensors (stlx_apply_operator)."""
        self._lib.call("stlx_apply_operator", self._h.ptr, int(kind), int(level), int(x.shape[0]),
                       int(x.data_ptr()), int(out.data_ptr()))
        return out

    # lifetime -----------------------------------------------------------------------------
    def close(self):
        lock = getattr(self, "_lock", None)
        if lock is None:
            return
        with lock:
            if self._h:
                self._h.release()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _positive_int(v):
    return not isinstance(v, bool) and int(v) == v and int(v) > 0


class NativeEnsemble:
    """Maps of uneven nside evaluated together on one GPU.  groups: (nside, batch[, j]) with j
    defaulting to log2(nside) - 1.  forward_device takes one float64 device tensor
    [batch, 12 nside^2] per group and returns one complex array [batch, ncoeff_g] per group
    (and the variances with moments=True); vjp_device takes cotangents of those shapes and
    writes each group's gradient into `out`."""

    def __init__(self, groups, precision="mixed", device=0, library=None, batch_nside=None):
        self._lock = threading.RLock()
        self._h = None
        self._lib = bridge(library)
        spec = []
        for g in groups:
            nside, batch = int(g[0]), int(g[1])
            spec.append((nside, batch, int(g[2]) if len(g) > 2 else nside.bit_length() - 2))
        if not spec:
            raise ValueError("an ensemble needs at least one group")
        self.groups = spec
        cols = [np.ascontiguousarray([s[k] for s in spec], dtype=np.int32) for k in range(3)]
        ints = [c.ctypes.data_as(_CTYPES["I"]) for c in cols]
        options = dict(precision=precision, device=int(device))
        if batch_nside is not None:
            options["batch_nside"] = int(batch_nside)
        options = json.dumps(options).encode()
        self._h = _Owned(self._lib, self._lib.call("stlx_ensemble_new", ints[0], ints[2], ints[1], len(spec),
                                                   options), "stlx_ensemble_free")
        counts = np.zeros(len(spec), dtype=np.uintp)
        self._lib.call("stlx_ensemble_ncoef", self._h.ptr, counts.ctypes.data_as(_CTYPES["Z"]))
        self.ncoeff = [int(c) for c in counts]
        self.npix = [12 * s[0] ** 2 for s in spec]
        self._rows = [s[1] * c for s, c in zip(spec, self.ncoeff)]

    def _device_table(self, arrays):
        if len(arrays) != len(self.groups):
            raise ValueError(f"expected {len(self.groups)} arrays, one per group")
        return np.array([_device_arg(a, s[1] * n)[0] for a, s, n in zip(arrays, self.groups, self.npix)],
                        dtype=np.uint64)

    def _split(self, flat):
        parts = np.split(flat, np.cumsum(self._rows)[:-1])
        return [p.reshape(s[1], -1) for p, s in zip(parts, self.groups)]

    @_engine_call
    def forward_device(self, maps, moments=False):
        table = self._device_table(maps)
        values = np.empty(sum(self._rows), dtype=np.complex128)
        var = np.empty(values.size if moments else 1, dtype=np.float64)
        self._lib.call("stlx_ensemble_forward_device", self._h.ptr, table.ctypes.data_as(_CTYPES["U"]), table.size,
                       bool(moments), _dptr(values.view(np.float64)), _dptr(var))
        return (self._split(values), self._split(var)) if moments else self._split(values)

    @_engine_call
    def vjp_device(self, cotangents, out):
        c = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.complex128).reshape(-1)
                                                 for v in cotangents]))
        if c.size != sum(self._rows):
            raise ValueError("cotangent shapes do not match the ensemble's statistics")
        table = self._device_table(out)
        self._lib.call("stlx_ensemble_vjp_device", self._h.ptr, _dptr(c.view(np.float64)), c.size,
                       table.ctypes.data_as(_CTYPES["U"]), table.size)
        return out

    def close(self):
        lock = getattr(self, "_lock", None)
        if lock is None:
            return
        with lock:
            if self._h:
                self._h.release()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def stream_copy_gbps(nbytes=2 << 30, reps=5, device=0, library=None):
    """HBM rate (bytes read + written per second, GB/s) of the library's hand-written stream
    copy kernel over a buffer of `nbytes` (stlx_stream_copy)."""
    import torch
    words = int(nbytes) // 16
    src = torch.empty(words * 4, dtype=torch.float32, device=f"cuda:{device}").uniform_()
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    ms = C.c_double(0.0)
    bridge(library).call("stlx_stream_copy", int(src.data_ptr()), int(dst.data_ptr()), words * 16, int(reps),
                         C.byref(ms))
    if not torch.equal(src, dst):
        raise RuntimeError("stream copy: destination differs from the source")
    return 2.0 * words * 16 / (ms.value * 1e-3) / 1e9


def ergodic_weights(native_pixels, variance):
    """Ergodic weights of the synthesis loss (reference crates/stl-mpi/src/synthesis.rs:296-314):
    w_i = N_i / (n V_i), V_i the spatial variance of coefficient i's terms on the target
    (forward(..., moments=True)), N_i its native pixel count (native_pixels()), n the number of
    coefficients.  A coefficient with V_i = 0 gets the largest finite weight."""
    counts = np.asarray(native_pixels, dtype=np.float64).reshape(-1)
    var = np.asarray(variance, dtype=np.float64).reshape(-1)
    if counts.shape != var.shape:
        raise ValueError("one variance per coefficient expected")
    with np.errstate(divide="ignore"):
        w = np.where(var > 0, counts / np.where(var > 0, var, 1.0), np.inf)
    top = w[np.isfinite(w)]
    w[~np.isfinite(w)] = top.max() if top.size else 0.0
    return w / w.size


# ---------------------------------------------------------------------------------------------
# Host L-BFGS of a Python objective

_LBFGS_DEFAULTS = dict(lr=1.0, max_iter=20, max_eval=None, tolerance_grad=1e-12, tolerance_change=1e-15,
                       history_size=10)


def _lbfgs_options(options):
    opts = dict(_LBFGS_DEFAULTS)
    extra = sorted(set(options or {}) - set(opts))
    if extra:
        raise ValueError(f"unknown native L-BFGS options: {extra}")
    opts.update(options or {})
    for name in ("max_iter", "history_size"):
        if not _positive_int(opts[name]):
            raise ValueError(f"{name} must be a positive integer")
    if opts["max_eval"] is not None and not _positive_int(opts["max_eval"]):
        raise ValueError("max_eval must be a positive integer or None")
    return opts


class _Objective:
    """The C callback of stlp_lbfgs around a Python objective: it records every loss and keeps
    the first exception (which the library sees as a non-finite loss)."""

    def __init__(self, fn, shape):
        self.fn, self.shape = fn, shape
        self.losses, self.failure = [], None
        self.c_callback = _CALLBACK(self._evaluate)

    def _evaluate(self, x_ptr, n, g_ptr, _ctx):
        if self.failure is not None:
            return float("nan")
        try:
            x = np.ctypeslib.as_array(x_ptr, shape=(n,)).reshape(self.shape)
            x.setflags(write=False)
            loss, grad = self.fn(x)
            grad = np.asarray(grad, dtype=np.float64)
            if grad.shape != self.shape:
                raise ValueError(f"gradient shape {grad.shape} != {self.shape}")
            np.ctypeslib.as_array(g_ptr, shape=(n,))[:] = grad.reshape(-1)
            self.losses.append(float(loss))
            return float(loss)
        except BaseException as exc:  # noqa: BLE001  (handed back to the caller below)
            self.failure = exc
            return float("nan")


def native_lbfgs(callback, x, options=None, library=None):
    """L-BFGS with the strong Wolfe line search (stlp_lbfgs) of callback(x) -> (loss, gradient)
    from x; the reference binding's native_lbfgs (native.py:1007-1052).  callback sees a
    read-only view valid during the call.  x is updated in place when it is a C-contiguous
    float64 array.  Returns (x, report)."""
    lib = bridge(library)
    opts = _lbfgs_options(options)
    x = np.ascontiguousarray(x, dtype=np.float64)
    if x.size == 0 or not np.isfinite(x).all():
        raise ValueError("finite nonempty initial ensemble required")
    objective = _Objective(callback, x.shape)
    request = json.dumps({k: v for k, v in opts.items() if v is not None}).encode()
    evaluations = lib.cdll.stlp_lbfgs(_dptr(x), x.size, request, objective.c_callback, None)
    if objective.failure is not None:
        raise objective.failure
    if evaluations < 0:
        raise lib.error()
    return x, dict(evaluations=int(evaluations), loss_history=objective.losses, options=opts,
                   optimizer="native L-BFGS", line_search="strong_wolfe")
