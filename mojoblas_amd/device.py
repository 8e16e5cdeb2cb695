"""Device-resident matrices through the library's own buffer entries
(mbh_dev_alloc / mbh_set_matrix / mbh_get_matrix / mbh_stream_*), for
callers that keep operands in HBM across calls without torch.  A
DeviceMatrix can be passed wherever the level-3 / level-2 functions take a
device buffer (it routes to the _dev entries), with its `ld` as the leading
dimension:

    A = DeviceMatrix.from_numpy(a)            # a: 2-D numpy array (rows, cols)
    C = DeviceMatrix(m, n, np.float32)
    with Stream() as s:
        mojoblas_amd.gemm("N", "N", m, n, k, 1.0, A, A.ld, B, B.ld, 0.0, C, C.ld, stream=s.handle)
        c = C.to_numpy(stream=s.handle)
"""
import ctypes

import numpy as np

from . import _lib


class HipBufferError(RuntimeError):
    """A buffer entry of the C-ABI returned an error."""


def _check(rc, what):
    if rc != 0:
        raise HipBufferError("%s failed (%d): %s" % (what, rc, _lib.last_error()))


class Stream:
    """A non-blocking HIP stream owned by the library (mbh_stream_create)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(_lib.lib().mbh_stream_create(ctypes.byref(h)), "mbh_stream_create")
        self.handle = h.value

    def sync(self):
        _check(_lib.lib().mbh_stream_sync(ctypes.c_void_p(self.handle)), "mbh_stream_sync")

    def close(self):
        if self.handle:
            _check(_lib.lib().mbh_stream_destroy(ctypes.c_void_p(self.handle)), "mbh_stream_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if self.handle:
            self.sync()
        self.close()


class DeviceMatrix:
    """A rows x cols column-major matrix in HBM with leading dimension ld
    (default rows), allocated by mbh_dev_alloc.  dtype: float32 or float64
    (the level-3 / level-2 routines), float16 or "bfloat16" (gemm's native
    16-bit path).  bfloat16 has no numpy type: its host side is uint16 bit
    patterns (the upper half of the fp32 bits)."""

    def __init__(self, rows, cols, dtype=np.float32, ld=None):
        if isinstance(dtype, str) and dtype == "bfloat16":
            self.dtype, self._host = "bfloat16", np.dtype(np.uint16)
        else:
            self.dtype = np.dtype(dtype)
            if self.dtype not in (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.float16)):
                raise TypeError("DeviceMatrix holds float32, float64, float16 or bfloat16")
            self._host = self.dtype
        self.rows, self.cols = int(rows), int(cols)
        self.ld = int(ld) if ld is not None else max(1, self.rows)
        if self.rows < 0 or self.cols < 0 or self.ld < max(1, self.rows):
            raise ValueError("bad shape (%d, %d) / ld %d" % (self.rows, self.cols, self.ld))
        p = ctypes.c_void_p()
        nbytes = self.ld * max(self.cols, 0) * self._host.itemsize
        _check(_lib.lib().mbh_dev_alloc(nbytes, ctypes.byref(p)), "mbh_dev_alloc")
        self.ptr = p.value or 0

    @classmethod
    def from_numpy(cls, a, ld=None, stream=None, dtype=None):
        """dtype="bfloat16" takes `a` as uint16 bit patterns."""
        a = np.asarray(a)
        if a.ndim != 2:
            raise ValueError("from_numpy takes a 2-D array (rows, cols)")
        m = cls(a.shape[0], a.shape[1], dtype if dtype is not None else a.dtype, ld)
        m.copy_from(a, stream)
        return m

    def copy_from(self, a, stream=None):
        """Host (rows, cols) array -> this matrix; complete on return."""
        host = np.asfortranarray(np.asarray(a, dtype=self._host))
        if host.shape != (self.rows, self.cols):
            raise ValueError("shape %r, expected %r" % (host.shape, (self.rows, self.cols)))
        s = ctypes.c_void_p(stream)
        _check(_lib.lib().mbh_set_matrix(self.rows, self.cols, self._host.itemsize, host.ctypes.data,
                                        max(1, self.rows), ctypes.c_void_p(self.ptr), self.ld, s),
               "mbh_set_matrix")
        _check(_lib.lib().mbh_stream_sync(s), "mbh_stream_sync")

    def to_numpy(self, stream=None):
        """This matrix -> a new (rows, cols) Fortran-ordered host array,
        after the work queued on `stream` (NULL: the default stream)."""
        out = np.empty((self.rows, self.cols), dtype=self._host, order="F")
        s = ctypes.c_void_p(stream)
        _check(_lib.lib().mbh_get_matrix(self.rows, self.cols, self._host.itemsize, ctypes.c_void_p(self.ptr),
                                        self.ld, out.ctypes.data, max(1, self.rows), s), "mbh_get_matrix")
        _check(_lib.lib().mbh_stream_sync(s), "mbh_stream_sync")
        return out

    def free(self):
        if self.ptr:
            _check(_lib.lib().mbh_dev_free(ctypes.c_void_p(self.ptr)), "mbh_dev_free")
            self.ptr = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
