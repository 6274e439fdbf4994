"""us3d: MI355X-native differentiable Gaussian slice renderer (UltraGauss slice splat, UltraG-Ray B-mode).

Host code (Python) drives hand-written HIP kernels for gfx950 through the C ABI of libus3d.so
(include/us3d.h). Importing the package does not load the library; the first render does, and fails loudly
if it has not been built.
"""

__all__ = ["__version__"]
__version__ = "0.1.0"
