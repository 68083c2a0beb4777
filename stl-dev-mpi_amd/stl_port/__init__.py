"""MI355X ScatCov engine: Python binding of libstl_bridge.so (reference stl_port API)."""
from .native import NativeEngine, Snapshot, library_path  # noqa: F401
