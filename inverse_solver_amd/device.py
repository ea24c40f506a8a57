"""Minimal HIP runtime access (device buffers, copies, events) through ctypes.

Keeps benchmarks and the serving loop free of a framework import: the solver
library only needs device pointers and a stream.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

_CANDIDATES = ['libamdhip64.so', '/opt/rocm/lib/libamdhip64.so']


def _load():
    for name in _CANDIDATES + [ctypes.util.find_library('amdhip64') or '']:
        if not name:
            continue
        try:
            return C.CDLL(name)
        except OSError:
            continue
    raise ImportError('libamdhip64.so not found (ROCm runtime required)')


hip = _load()
hip.hipGetErrorString.restype = C.c_char_p
_H2D, _D2H, _D2D = 1, 2, 3


def _chk(err, what):
    if err != 0:
        raise RuntimeError(f'{what}: {hip.hipGetErrorString(err).decode()}')


def set_device(dev):
    _chk(hip.hipSetDevice(int(dev)), 'hipSetDevice')


def device_count():
    n = C.c_int()
    _chk(hip.hipGetDeviceCount(C.byref(n)), 'hipGetDeviceCount')
    return n.value


def synchronize():
    _chk(hip.hipDeviceSynchronize(), 'hipDeviceSynchronize')


def mem_info():
    free, total = C.c_size_t(), C.c_size_t()
    _chk(hip.hipMemGetInfo(C.byref(free), C.byref(total)), 'hipMemGetInfo')
    return free.value, total.value


class DeviceArray:
    """A device buffer with a NumPy dtype and shape (row-major in memory)."""

    def __init__(self, shape, dtype=np.complex64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _chk(hip.hipMalloc(C.byref(p), C.c_size_t(max(self.nbytes, 1))), 'hipMalloc')
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _chk(hip.hipMemcpy(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(d.nbytes), _H2D), 'H2D')
        return d

    def upload(self, a):
        """Copy a host array of the same byte size (in its memory order) into the buffer."""
        a = np.asarray(a)
        if not (a.flags['C_CONTIGUOUS'] or a.flags['F_CONTIGUOUS']) or a.nbytes != self.nbytes:
            raise ValueError('upload needs a contiguous array of %d bytes' % self.nbytes)
        _chk(hip.hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), _H2D), 'H2D')

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        _chk(hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_size_t(self.nbytes), _D2H),
             'D2H')
        return out

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        _chk(hip.hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), C.c_size_t(self.nbytes), _D2D), 'D2D')

    def free(self):
        if self.ptr:
            hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        _chk(hip.hipEventCreate(C.byref(e)), 'hipEventCreate')
        self.e = e

    def record(self, stream):
        _chk(hip.hipEventRecord(self.e, C.c_void_p(stream)), 'hipEventRecord')

    def synchronize(self):
        _chk(hip.hipEventSynchronize(self.e), 'hipEventSynchronize')

    def elapsed_ms(self, end):
        ms = C.c_float()
        _chk(hip.hipEventElapsedTime(C.byref(ms), self.e, end.e), 'hipEventElapsedTime')
        return ms.value

    def __del__(self):
        try:
            hip.hipEventDestroy(self.e)
        except Exception:
            pass


def stream_synchronize(stream):
    _chk(hip.hipStreamSynchronize(C.c_void_p(stream)), 'hipStreamSynchronize')


def device_name(dev=0):
    # hipDeviceProp_t is large; read the name through hipDeviceGetName
    buf = C.create_string_buffer(256)
    _chk(hip.hipDeviceGetName(buf, 256, int(dev)), 'hipDeviceGetName')
    return buf.value.decode()


os.environ.setdefault('HSA_ENABLE_IPC_MODE_LEGACY', '0')
