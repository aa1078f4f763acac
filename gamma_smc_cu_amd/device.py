"""Device memory, streams and events through the C ABI (no GPU framework).

Small RAII wrappers used by the device-resident decode path, the batch and
serving front ends, the benchmark and the GPU tests.
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._core import check, lib


class DeviceArray:
    """A device allocation holding a C-contiguous array of ``dtype``."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().gsmc_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib().gsmc_memcpy_htod(d.ptr, ctypes.c_void_p(a.ctypes.data), d.nbytes))
        return d

    def to_host(self, out=None):
        if out is None:
            out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().gsmc_memcpy_dtoh(ctypes.c_void_p(out.ctypes.data), self.ptr, self.nbytes))
        return out

    def offset(self, elements):
        """Raw device pointer ``elements`` items past the start."""
        return ctypes.c_void_p(self.ptr.value + int(elements) * self.dtype.itemsize)

    def fill_zero(self):
        check(lib().gsmc_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            check(lib().gsmc_free(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(lib().gsmc_stream_create(ctypes.byref(s)))
        self.handle = s

    def synchronize(self):
        check(lib().gsmc_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle is not None:
                lib().gsmc_stream_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(lib().gsmc_event_create(ctypes.byref(e)))
        self.handle = e

    def record(self, stream=None):
        check(lib().gsmc_event_record(self.handle, None if stream is None else stream.handle))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(lib().gsmc_event_elapsed_ms(self.handle, end.handle, ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle is not None:
                lib().gsmc_event_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def synchronize():
    check(lib().gsmc_synchronize())


def run_fb_device(ctx, d_pi, d_pj, n_pairs, core_block_sites, flank_sites, flags, d_mean, d_lower=None,
                  d_upper=None, d_alpha=None, d_beta=None, stream=None):
    """Queue a decode with device-resident pairs and outputs (no sync)."""

    def p(x):
        if x is None:
            return None
        return x.ptr if isinstance(x, DeviceArray) else x

    check(lib().gsmc_run_fb_device(ctx.handle, p(d_pi), p(d_pj), int(n_pairs), int(core_block_sites),
                                   int(flank_sites), int(flags), p(d_mean), p(d_lower), p(d_upper), p(d_alpha),
                                   p(d_beta), None if stream is None else stream.handle))


def run_fwd_device(ctx, d_pi, d_pj, n_pairs, flags, d_mean, d_lower=None, d_upper=None, stream=None):
    """Queue a forward-only decode with device-resident pairs and outputs."""

    def p(x):
        if x is None:
            return None
        return x.ptr if isinstance(x, DeviceArray) else x

    check(lib().gsmc_run_fwd_device(ctx.handle, p(d_pi), p(d_pj), int(n_pairs), int(flags), p(d_mean), p(d_lower),
                                    p(d_upper), None if stream is None else stream.handle))
