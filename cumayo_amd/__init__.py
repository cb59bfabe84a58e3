"""MI355X-native batched MAYO engine.

The product is the C ABI in ``include/cumayo.h`` implemented by the HIP
libraries under ``cumayo_amd/lib`` (one per parameter set); ``engine`` is a
thin Python front end over it.
"""
from .engine import Engine, Sizes, VERSIONS, load_library, a128  # noqa: F401

__all__ = ["Engine", "Sizes", "VERSIONS", "load_library", "a128"]
