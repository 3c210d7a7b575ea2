"""temf3dt_amd -- MI355X-native FETD forward time stepping for TEMF3DT.

The compute path is libtemfetd.so (hand-written gfx950 HIP kernels behind
the C ABI in include/temfetd.h); this package holds its Python host side.
"""
from .fetd import Fetd, FetdError, load_library, run_root, setup_from_files  # noqa: F401

__all__ = ["Fetd", "FetdError", "load_library", "run_root", "setup_from_files"]
