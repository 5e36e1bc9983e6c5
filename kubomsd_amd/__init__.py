"""MI355X-native backend for KuboMSD's stochastic KPM transport path.

Host code in Python on PyTorch-ROCm device tensors; compute in hand-written
gfx950 HIP kernels behind the C ABI in include/kmsd.h (libkmsd.so).
"""
__version__ = "0.1.0"
