"""MI355X ScatCov engine: Python binding of libstl_bridge.so (reference stl_port API)."""
from .native import (BridgeError, NativeEngine, NativeEnsemble, Snapshot, bridge, ergodic_weights,  # noqa: F401
                     library_path, native_lbfgs)
