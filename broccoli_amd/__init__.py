"""MI355X-native second-level permutation-test GLM (BROCCOLI RandomiseGroupLevel path).

The product is the C-ABI library libbpg.so (include/bpg.h) with the
BROCCOLI_LIB drop-in class (include/broccoli_lib.h).  `capi` is its ctypes
binding; `synth` generates synthetic studies of the benchmark shapes.
"""
__all__ = ["capi", "synth"]
