"""MI355X-native Sentinel-1 Level-0 -> SLC hot path.

The product is the gfx950 library ``libs1gpu.so`` behind the C ABI of
``include/s1gpu.h``; ``capi`` binds it with ctypes, ``pipeline`` holds the
burst-level host logic, ``synth`` makes synthetic packet streams.
"""
__all__ = ["capi", "pipeline", "synth"]
