"""Python host of the device-resident field store (include/cloudsc_field.h).

Mirrors the IFS field API's owner / wrapper objects (ifs-source/contrib/field_api/field_factory_module.fypp
FIELD_NEW / FIELD_DELETE, field_RANKSUFF_module.fypp GET_DEVICE_DATA_RDONLY/_RDWR, GET_HOST_DATA_RDONLY/_RDWR,
SYNC_HOST_*, SYNC_DEVICE_*, GET_VIEW, DELETE_DEVICE and the GPU_STATS counters); the data movement runs in
libcloudsc_hip.so.  `step_fields_run` is the physics step on fields (cloudsc_step_fields_run).
"""
import ctypes

import numpy as np

from . import cloudsc as C

RD, WR, RDWR = 1, 2, 3
DEVFRESH, HSTFRESH, UNALLOCATED = 1, 2, 4
UNINITIALIZED = DEVFRESH | HSTFRESH

_STEP_NAMES = [n for n, _, _ in C.BLOCKED_ARRAYS] + [
    "phrsw", "phrlw", "ptendency_vdf_t", "ptendency_vdf_q", "pvervel",
    "ptendency_satadj_t", "ptendency_satadj_q", "ptendency_satadj_a", "ptendency_satadj_cld"]


class _StepFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in _STEP_NAMES]


class Field:
    """One field: a wrapper around a Fortran-ordered numpy array, or an owner with library host memory."""

    def __init__(self, handle, shape, dtype, wrapped=None):
        self.h = ctypes.c_void_p(handle)
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self._wrapped = wrapped  # keeps the numpy array alive

    # -- factories (FIELD_NEW) --
    @classmethod
    def wrap(cls, array, lbounds=None):
        if not array.flags.f_contiguous:
            raise TypeError("a wrapped array must be Fortran-contiguous")
        rank = array.ndim
        lb = np.array(lbounds if lbounds is not None else [1] * rank, dtype=np.int64)
        ub = lb + np.array(array.shape, dtype=np.int64) - 1
        h = ctypes.c_void_p()
        C._check("cloudsc_field_wrap", C.lib().cloudsc_field_wrap(
            ctypes.c_void_p(array.ctypes.data), ctypes.c_size_t(array.itemsize), ctypes.c_int(rank),
            ctypes.c_void_p(lb.ctypes.data), ctypes.c_void_p(ub.ctypes.data), ctypes.byref(h)))
        return cls(h.value, array.shape, array.dtype, wrapped=array)

    @classmethod
    def owner(cls, shape, dtype, lbounds=None, delayed=False, init_value=None):
        dtype = np.dtype(dtype)
        rank = len(shape)
        lb = np.array(lbounds if lbounds is not None else [1] * rank, dtype=np.int64)
        ub = lb + np.array(shape, dtype=np.int64) - 1
        init = None if init_value is None else np.array([init_value], dtype=dtype)
        h = ctypes.c_void_p()
        C._check("cloudsc_field_owner", C.lib().cloudsc_field_owner(
            ctypes.c_size_t(dtype.itemsize), ctypes.c_int(rank), ctypes.c_void_p(lb.ctypes.data),
            ctypes.c_void_p(ub.ctypes.data), ctypes.c_int(int(bool(delayed))),
            ctypes.c_void_p(None if init is None else init.ctypes.data), ctypes.byref(h)))
        return cls(h.value, shape, dtype)

    # -- data access --
    def get_device_data(self, mode=RD, stream=None):
        p = ctypes.c_void_p()
        C._check("cloudsc_field_get_device_data", C.lib().cloudsc_field_get_device_data(
            self.h, ctypes.c_int(mode), ctypes.byref(p), ctypes.c_void_p(stream)))
        return p.value

    def device_tensor(self, mode=RD, stream=None):
        """The device copy as a torch tensor (no copy), in Fortran element order."""
        import torch

        ptr = self.get_device_data(mode, stream)

        class _Cai:
            __cuda_array_interface__ = {"shape": self.shape[::-1], "typestr": self.dtype.str, "data": (ptr, False),
                                        "version": 3, "strides": None}

        return torch.as_tensor(_Cai(), device="cuda").permute(*reversed(range(len(self.shape))))

    def get_host_data(self, mode=RD, stream=None):
        p = ctypes.c_void_p()
        C._check("cloudsc_field_get_host_data", C.lib().cloudsc_field_get_host_data(
            self.h, ctypes.c_int(mode), ctypes.byref(p), ctypes.c_void_p(stream)))
        return self._host_array(p.value)

    def _host_array(self, addr):
        n = int(np.prod(self.shape))
        buf = (ctypes.c_char * (n * self.dtype.itemsize)).from_address(addr)
        return np.frombuffer(buf, dtype=self.dtype).reshape(self.shape, order="F")

    def sync_device(self, mode=RD, stream=None):
        C._check("cloudsc_field_sync_device", C.lib().cloudsc_field_sync_device(self.h, ctypes.c_int(mode),
                                                                                ctypes.c_void_p(stream)))

    def sync_host(self, mode=RD, stream=None):
        C._check("cloudsc_field_sync_host", C.lib().cloudsc_field_sync_host(self.h, ctypes.c_int(mode),
                                                                            ctypes.c_void_p(stream)))

    def get_view(self, block):
        p = ctypes.c_void_p()
        C._check("cloudsc_field_get_view", C.lib().cloudsc_field_get_view(self.h, ctypes.c_int64(block),
                                                                          ctypes.byref(p)))
        n = int(np.prod(self.shape[:-1]))
        buf = (ctypes.c_char * (n * self.dtype.itemsize)).from_address(p.value)
        return np.frombuffer(buf, dtype=self.dtype).reshape(self.shape[:-1], order="F")

    def delete_device(self):
        C._check("cloudsc_field_delete_device", C.lib().cloudsc_field_delete_device(self.h))

    @property
    def status(self):
        s = ctypes.c_int()
        C._check("cloudsc_field_status", C.lib().cloudsc_field_status(self.h, ctypes.byref(s)))
        return s.value

    @property
    def stats(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        C._check("cloudsc_field_stats", C.lib().cloudsc_field_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def delete(self):
        """FIELD_DELETE (a wrapper first makes the wrapped array current)."""
        if self.h:
            h, self.h = self.h, ctypes.c_void_p()
            C._check("cloudsc_field_delete", C.lib().cloudsc_field_delete(h))


def blocked_fortran_shape(kind, nproma, klev, ngpblks):
    """Fortran shape (NPROMA, KLEV[+1], [n,] NGPBLKS) of a blocked field."""
    return C.blocked_shape(kind, nproma, klev, ngpblks)[::-1]


def step_fields_run(params, precision, ngptot, ptsphy, fields, stream=None):
    """cloudsc_step_fields_run: ``fields`` maps argument names (CLOUDSC's blocked arrays, optionally the
    CLOUD_SATADJ inputs and TENDENCY_SATADJ outputs) to Field objects."""
    s = _StepFields(**{n: (f.h.value if isinstance(f, Field) else None) for n, f in fields.items()})
    prec = {"dp": 0, "sp": 1, "mixed": 2}[precision]
    C._check("cloudsc_step_fields_run", C.lib().cloudsc_step_fields_run(
        ctypes.byref(params), ctypes.c_int(prec), ctypes.c_int(ngptot), ctypes.c_double(ptsphy), ctypes.byref(s),
        ctypes.c_void_p(stream)))
