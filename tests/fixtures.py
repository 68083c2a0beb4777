"""Access to the golden fixtures under tests/golden (copied from the reference's
`fixtures/<suite>/<case>/*.npy`, dumped from STL-Dev by its tools/dump_reference.py)."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


class Case:
    def __init__(self, suite, meta):
        self.meta = meta
        self.dir = GOLDEN / suite / meta["name"]

    @property
    def name(self):
        return self.meta["name"]

    def __getitem__(self, key):
        return self.meta[key]

    def has(self, array):
        return (self.dir / f"{array}.npy").exists()

    def arr(self, array):
        return np.load(self.dir / f"{array}.npy")

    def complex(self, name):
        return self.arr(f"{name}_re") + 1j * self.arr(f"{name}_im")


def cases(suite, kind):
    index = json.loads((GOLDEN / suite / "index.json").read_text())
    out = [Case(suite, m) for m in index if m["kind"] == kind]
    assert out, f"no {kind} case in {suite}"
    return out


def conv_noise(pix2ang_theta, nside):
    """Rounding noise of a convolution output (reference tests/common/mod.rs:78-89):
    acos ulp error amplified by 1/sin(theta), linear over one pixel."""
    s = max(np.sin(pix2ang_theta), 1.0 / nside)
    return max(1e-12, np.finfo(float).eps * 128.0 * nside / s)
