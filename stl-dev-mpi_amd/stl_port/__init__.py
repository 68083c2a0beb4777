"""MI355X ScatCov engine: Python binding of libstl_bridge.so (reference stl_port API)."""
from .native import NativeEngine, Snapshot, ergodic_weights, library_path, native_lbfgs  # noqa: F401
