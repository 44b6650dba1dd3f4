"""MI355X-native Gay-Berne checkerboard Monte Carlo (libmcgb.so, C ABI in include/mcgb.h).

The compute path is HIP (csrc/); `mcgb` is the ctypes binding used by tests and the benchmark,
`build` compiles the library and the C driver in-tree."""
from . import mcgb  # noqa: F401
