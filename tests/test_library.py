"""The C-ABI library loads, exports every symbol include/stl_bridge.h declares,
and reports configuration errors through stlp_last_error (no GPU needed)."""
import ctypes as C
import json
import re
from pathlib import Path

import numpy as np
import pytest

from stl_port import native

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "stl_bridge.h"


def header_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(stl[px]_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(native.library_path())
    names = header_symbols()
    assert len(names) >= 55
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_exports_what_the_reference_binding_binds():
    # the reference binding (native.py:55-116) sets argtypes on each of these names at load
    # time, so every one must resolve; the list is data extracted from it
    # (tests/golden/binding/extract_symbols.py)
    ref = json.loads((ROOT / "tests" / "golden" / "binding" / "reference_symbols.json").read_text())["symbols"]
    assert len(ref) == 52
    lib = C.CDLL(native.library_path())
    missing = [n for n in ref if not hasattr(lib, n)]
    assert not missing, missing
    # and this binding types every one of them
    assert set(ref) <= set(native.ABI)


def test_abi_table_matches_the_header():
    # every typed entry point is declared in include/stl_bridge.h with as many parameters
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    for name, code in native.ABI.items():
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", text)
        assert m, name
        params = [p for p in m.group(1).split(",") if p.strip() and p.strip() != "void"]
        assert len(params) == len(code) - 1, (name, params, code)


@pytest.mark.parametrize("config, message", [
    ("{", "valid JSON"),
    ('{"nside": 12}', "power of two"),
    ('{"nside": 16384, "j": 3}', "at most 8192"),
    ('{"nside": 16, "j": 9}', "invalid number of scales"),
    ('{"nside": 16, "j": 3, "kernel": 3}', "kernel"),
    ('{"nside": 16, "j": 3, "orientations": 2}', "orientations"),
    ('{"nside": 16, "j": 3, "backend": "cpu"}', "backend cpu"),
    ('{"nside": 16, "j": 3, "backend": "tpu"}', "backend must be"),
    ('{"nside": 16, "j": 3, "spin": 2}', "spin 2"),
    ('{"nside": 16, "j": 3, "precision": "fp16"}', "precision"),
    ('{"nside": 16, "j": 3, "batch": 0}', "batch"),
    ('{"nside": 16, "j": 3, "conv_variant": 6}', "conv_variant"),
    ('{"nside": 16, "j": 3, "conv_variant": -2}', "conv_variant"),
    ('{"nside": 16, "j": 3, "batch_nside": -1}', "batch_nside"),
    ('{"nside": 16, "j": 3, "fields": "sometimes"}', "fields"),
])
def test_bad_configuration_is_an_error(config, message):
    lib = native.bridge()
    with pytest.raises(native.BridgeError, match=message):
        lib.call("stlp_new", config.encode())


def test_other_products_are_not_implemented():
    lib = native.bridge().cdll
    assert not lib.stlp_patch_new(b"{}")
    assert "not implemented" in lib.stlp_last_error().decode()
    assert lib.stlp_unet_zero_grad(None) == -1
    assert lib.stlp_spin_fields(None, None, 0, 0, None, None) == -1


def test_identity_key_matches_the_reference_hash():
    # native.py:122-129: first 8 bytes of sha256(objective), little endian
    import hashlib
    g, m, o = native._key(3, 7, "gaussian")
    assert native.objective_id("gaussian") == o
    assert (g, m) == (3, 7)
    assert o == int.from_bytes(hashlib.sha256(b"gaussian").digest()[:8], "little")
    with pytest.raises(ValueError):
        native._key(-1, 0, "x")
    with pytest.raises(ValueError):
        native._key(0, 0, "")


def test_lbfgs_failure_value_without_callback_errors():
    lib = native.bridge().cdll
    x = np.zeros(4)
    assert lib.stlp_lbfgs(native._dptr(x), 4, b"{bad", native._CALLBACK(lambda *a: 0.0), None) == -1


@pytest.mark.parametrize("nside, j, batch, options, message", [
    ([16, 12], [3, 3], [1, 1], "{}", "power of two"),
    ([16], [6], [1], "{}", "j must satisfy"),
    ([16], [3], [0], "{}", "batch"),
    ([16], [3], [1], '{"precision": "fp16"}', "precision"),
])
def test_bad_ensemble_is_an_error(nside, j, batch, options, message):
    # validated before any device work
    from stl_port import NativeEnsemble
    with pytest.raises(RuntimeError, match=message):
        NativeEnsemble([(n, b, jj) for n, jj, b in zip(nside, j, batch)],
                       precision=json.loads(options).get("precision", "mixed"))


def test_ensemble_without_groups_is_an_error():
    lib = native.bridge().cdll
    assert not lib.stlx_ensemble_new(None, None, None, 0, b"{}")
    assert "at least one group" in lib.stlp_last_error().decode()
