"""Python host of CLOUD_SUPERSATCHECK and SLTEND's supersaturation adjustment (include/cloudsc_supersat.h).

``CLOUD_SUPERSATCHECK`` mirrors SUBROUTINE CLOUD_SUPERSATCHECK (ifs-source/arpifs/phys_ec/cloud_supersatcheck.F90:9-56)
on numpy arrays of one level; ``sltend_supersat_blocked`` runs section 3 of SLTEND (sltend.F90:370-409) on
device-resident blocked torch tensors.  The arithmetic runs only in libcloudsc_hip.so.
"""
import ctypes

import numpy as np

from . import cloudsc as C

SLTEND_ARRAYS = ["pt", "pq", "pa", "prsf1", "prs1", "pcent", "pcenq", "pcena", "ptent", "ptenq", "ptena", "ptencld",
                 "padja", "padjl", "padji"]


def CLOUD_SUPERSATCHECK(params, kidia, kfdia, klon, klev, kftliqice, pt, pq, pa, pp, psurf, precision="dp"):
    """Returns (PT_ADJ, PQ_ADJ, PA_ADJ, PL_ADJ, PI_ADJ) as numpy arrays of KLON elements (NaN outside KIDIA:KFDIA)."""
    dt = C._real(precision)
    ins = [np.ascontiguousarray(x, dtype=dt) for x in (pt, pq, pa, pp, psurf)]
    for x in ins:
        if x.shape != (klon,):
            raise ValueError(f"arrays must have KLON={klon} elements")
    outs = [np.full(klon, np.nan, dtype=dt) for _ in range(5)]
    fn = f"cloudsc_supersatcheck_{precision}_c"
    C._check(fn, getattr(C.lib(), fn)(ctypes.byref(params), ctypes.c_int(kidia), ctypes.c_int(kfdia),
                                      ctypes.c_int(klon), ctypes.c_int(klev), ctypes.c_int(kftliqice),
                                      *(ctypes.c_void_p(x.ctypes.data) for x in ins + outs)))
    return outs


def sltend_supersat_blocked(arrays, params, tsphy, ngptot, nproma, klev, precision="sp", stream=None):
    """cloudsc_sltend_supersat_blocked_{dp,sp,mixed}_c on a dict of blocked device tensors (``padj*`` optional)."""
    import torch

    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    ngpblks = (ngptot + nproma - 1) // nproma
    fn = f"cloudsc_sltend_supersat_blocked_{precision}_c"
    ptr = lambda n: ctypes.c_void_p(arrays[n].data_ptr() if arrays.get(n) is not None else 0)
    C._check(fn, getattr(C.lib(), fn)(ctypes.byref(params), ctypes.c_int(ngptot), ctypes.c_int(nproma),
                                      ctypes.c_int(klev), ctypes.c_int(ngpblks), C._ctype_real(precision)(tsphy),
                                      *(ptr(n) for n in SLTEND_ARRAYS), ctypes.c_void_p(stream)))
