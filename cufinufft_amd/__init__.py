"""MI355X-native NUFFT (types 1 and 2, 1D/2D/3D, FP32/FP64) behind the
cufinufft C ABI.  ``from cufinufft_amd import cufinufft`` is the drop-in for
``from cufinufft import cufinufft``.

Importing the package loads the in-tree HIP library and raises if it has not
been built (``python cufinufft_amd/build.py``); there is no CPU fallback.
"""
from cufinufft_amd.cufinufft import cufinufft  # noqa: F401

__all__ = ["cufinufft"]
