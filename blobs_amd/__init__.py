"""blobs_amd -- MI355X-native (gfx950) drop-in for blobs' `Physics::step`.

The product is libblobs_gpu.so (HIP kernels + the C ABI in
include/blobs_gpu.h).  This package holds its Python binding (`physics`),
the synthetic scene generator (`scenes`) and the build script (`build`).
"""
from . import _lib  # noqa: F401

__all__ = ["physics", "scenes", "build"]
