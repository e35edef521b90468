"""An exportable device allocation for the interop tests: HIP's virtual
memory API (hipMemCreate with a POSIX-descriptor handle type), exported as a
dma-buf file descriptor and mapped at an address of its own.  It stands in for
an exportable Vulkan VkDeviceMemory: the descriptor is what
hl_external_import takes, and the exporter's own mapping is how a test reads
what the library wrote through the import."""
from __future__ import annotations

import ctypes as C
import os


class _Location(C.Structure):
    _fields_ = [("type", C.c_int), ("id", C.c_int)]


class _AllocFlags(C.Structure):
    _fields_ = [("compressionType", C.c_ubyte), ("gpuDirectRDMACapable", C.c_ubyte), ("usage", C.c_ushort)]


class _AllocProp(C.Structure):
    _fields_ = [("type", C.c_int), ("requestedHandleType", C.c_int), ("location", _Location),
                ("win32HandleMetaData", C.c_void_p), ("allocFlags", _AllocFlags)]


class _AccessDesc(C.Structure):
    _fields_ = [("location", _Location), ("flags", C.c_int)]


_PINNED, _POSIX_FD, _LOC_DEVICE, _READ_WRITE, _GRAN_MIN = 1, 1, 1, 3, 0


def _hip():
    # the HIP runtime libhelio_hip.so links (loaded by the library first)
    from helio_amd.device import lib
    lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipGetErrorString.restype = C.c_char_p
    return h


class ExportedAllocation:
    """size bytes (rounded up to the allocation granularity) on `device`,
    with .fd (the exported descriptor, owned here) and .ptr (the exporter's
    mapping)."""

    def __init__(self, size: int, device: int = 0):
        self.hip = h = _hip()
        self.prop = _AllocProp(type=_PINNED, requestedHandleType=_POSIX_FD, location=_Location(_LOC_DEVICE, device))
        gran = C.c_size_t()
        self._ck(h.hipMemGetAllocationGranularity(C.byref(gran), C.byref(self.prop), _GRAN_MIN), "granularity")
        self.size = (int(size) + gran.value - 1) // gran.value * gran.value
        self.handle = C.c_void_p()
        self._ck(h.hipMemCreate(C.byref(self.handle), C.c_size_t(self.size), C.byref(self.prop), C.c_ulonglong(0)),
                 "hipMemCreate")
        fd = C.c_int(-1)
        self._ck(h.hipMemExportToShareableHandle(C.byref(fd), self.handle, _POSIX_FD, C.c_ulonglong(0)),
                 "hipMemExportToShareableHandle")
        self.fd = fd.value
        p = C.c_void_p()
        self._ck(h.hipMemAddressReserve(C.byref(p), C.c_size_t(self.size), C.c_size_t(0), None, C.c_ulonglong(0)),
                 "hipMemAddressReserve")
        self.ptr = p.value
        self._ck(h.hipMemMap(C.c_void_p(self.ptr), C.c_size_t(self.size), C.c_size_t(0), self.handle,
                             C.c_ulonglong(0)), "hipMemMap")
        acc = _AccessDesc(location=_Location(_LOC_DEVICE, device), flags=_READ_WRITE)
        self._ck(h.hipMemSetAccess(C.c_void_p(self.ptr), C.c_size_t(self.size), C.byref(acc), C.c_size_t(1)),
                 "hipMemSetAccess")

    def _ck(self, r: int, what: str) -> None:
        if r != 0:
            raise RuntimeError(f"{what}: {self.hip.hipGetErrorString(r).decode()} ({r})")

    def close(self) -> None:
        h = self.hip
        if self.ptr:
            self._ck(h.hipDeviceSynchronize(), "hipDeviceSynchronize")
            self._ck(h.hipMemUnmap(C.c_void_p(self.ptr), C.c_size_t(self.size)), "hipMemUnmap")
            self._ck(h.hipMemAddressFree(C.c_void_p(self.ptr), C.c_size_t(self.size)), "hipMemAddressFree")
            self.ptr = None
        if self.handle:
            self._ck(h.hipMemRelease(self.handle), "hipMemRelease")
            self.handle = None
        if self.fd >= 0:
            os.close(self.fd)
            self.fd = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
