"""MI355X-native Adaptive Banded Event Alignment (ABEA).

The compute path is libabea (HIP kernels for gfx950 behind the C ABI in
include/abea.h); this package is the Python view of it: ctypes binding
(_lib), device-resident batches (engine), the autograd binding (autograd),
synthetic inputs (synth) and f5c dump I/O (dump).
"""
from . import dump, synth  # noqa: F401
from .fixtures import load_golden_subset  # noqa: F401
