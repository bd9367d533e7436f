"""MI355X (gfx950) implementation of HERO's read-correction hot path.

hero_amd.gpu    ctypes binding of libhero_gpu.so (include/hero_gpu.h): seed index, banded DP,
                alignment records, votes and consensus on the GPU
hero_amd.synth  deterministic synthetic read sets for tests and benchmarks
"""
__all__ = ["gpu", "synth"]
