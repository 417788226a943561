"""cuszp_amd — MI355X-native (gfx950) cuSZp-compatible fused error-bounded compressor.

The product is libcuszp_amd.so (HIP kernels + C ABI, include/cuSZp.h). This
package holds its sources (csrc/), its build script and the Python binding
(pycuSZp), which mirrors the reference's python/pycuSZp.py.
"""
from ._lib import LIB_PATH, Uint3, load  # noqa: F401

__all__ = ["LIB_PATH", "Uint3", "load"]
