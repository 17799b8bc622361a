"""Device buffers a test owns, from the HIP runtime the library links (loaded
already by the library, so one runtime per process; torch's bundled runtime
is kept out of the GPU tests)."""
import ctypes

import numpy as np


class HipBuffers:
    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def upload(self, host):
        host = np.ascontiguousarray(host)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(host.nbytes, 8))) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(host.nbytes), 1) == 0
        return p

    def download(self, p, out):
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []
