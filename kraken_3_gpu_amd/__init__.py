"""MI355X-native batched 2-D XPBD step for cephalopod populations.

libkxpbd.so (csrc/, HIP for gfx950) holds the device paths behind a C ABI
(include/kxpbd.h); this package is the Python host: the gpuworld surface
(gpuworld.py), caller-side worlds and synthetic workloads (world.py).
"""
from . import _lib  # noqa: F401
