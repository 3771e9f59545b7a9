"""Device buffers for tests of device-pointer entry points, allocated through
the library's internal hooks (bgx_device_alloc ...), i.e. by the same HIP
runtime the library runs on.  torch's bundled HIP runtime is a different one,
and the two cannot share the GPU in one process on these machines."""
import ctypes as C

import numpy as np

from blobs_amd import _lib as L


class DevBuf:
    """Device memory for the test's inputs and outputs (the library's own HIP runtime)."""

    def __init__(self, lib, nbytes, host=None):
        self.lib, self.n = lib, int(nbytes)
        self.p = C.c_void_p()
        L.check(lib.bgx_device_alloc(max(self.n, 1), C.byref(self.p)), None)
        if host is not None:
            L.check(lib.bgx_copy_to_device(self.p, host.ctypes.data, self.n), None)

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        L.check(self.lib.bgx_copy_to_host(out.ctypes.data, self.p, out.nbytes), None)
        return out

    def __del__(self):
        self.lib.bgx_device_free(self.p)
