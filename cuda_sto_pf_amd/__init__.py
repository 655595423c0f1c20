"""MI355X-native time step of the CUDA_STO_PF phase-field simulator.

The compute path is libpfcore.so (hand-written HIP for gfx950) behind the C ABI
of include/pfcore.h; ``pfcore`` is its ctypes binding.
"""
__all__ = ["pfcore", "configs"]
