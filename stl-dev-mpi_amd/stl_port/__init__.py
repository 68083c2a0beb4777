"""MI355X ScatCov engine: Python binding of libstl_bridge.so (reference stl_port API)."""
from .native import (NativeEngine, NativeEnsemble, Snapshot, ergodic_weights, library_path,  # noqa: F401
                     native_lbfgs)
