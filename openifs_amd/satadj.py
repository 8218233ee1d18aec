"""Python host of the gfx950 CLOUD_SATADJ (saturation adjustment ahead of CLOUDSC).

Mirrors SUBROUTINE CLOUD_SATADJ (ifs-source/arpifs/phys_ec/cloud_satadj.F90:13-26): ``CLOUD_SATADJ`` takes one
NPROMA block of Fortran-ordered numpy arrays under the reference's argument names and calls
cloudsc_satadj_{dp,sp,mixed}_c; ``SatadjFields`` + ``satadj_blocked`` run the device-resident blocked entry over
all blocks on torch tensors (torch is the allocator and the stream, nothing more).  The arithmetic runs only in
libcloudsc_hip.so; a missing library raises CloudscLibraryError.
"""
import ctypes

import numpy as np

from . import cloudsc as C

# Per-block argument list after (YDPARAMS, KIDIA, KFDIA, KLON, KLEV, KTYPE, LDDIAG, PTSPHY), reference order
# (cloud_satadj.F90:13-26).  kind: 2d=(KLON,KLEV), h=(KLON,KLEV+1), c4=(KLON,KLEV,4); role i/o/u(nused).
SATADJ_ARRAYS = [
    ("pap", "2d", "i"), ("paph", "h", "i"), ("pt", "2d", "i"), ("pq", "2d", "i"), ("pa", "2d", "i"),
    ("pl", "2d", "i"), ("pi", "2d", "i"),
    ("ptendency_cml_t", "2d", "i"), ("ptendency_cml_q", "2d", "i"), ("ptendency_cml_a", "2d", "i"),
    ("ptendency_cml_l", "2d", "i"), ("ptendency_cml_i", "2d", "i"),
    ("phrsw", "2d", "i"), ("phrlw", "2d", "i"), ("ptendency_vdf_t", "2d", "i"), ("ptendency_vdf_q", "2d", "i"),
    ("pmfu", "2d", "u"), ("pmfd", "2d", "u"), ("pludeli", "c4", "i"), ("pvervel", "2d", "i"),
    ("pgp2dspp", "1d", "u"),
    ("ptendency_loc_t", "2d", "o"), ("ptendency_loc_q", "2d", "o"), ("ptendency_loc_a", "2d", "o"),
    ("ptendency_loc_l", "2d", "o"), ("ptendency_loc_i", "2d", "o"),
]
SATADJ_OUTPUTS = [n for n, k, r in SATADJ_ARRAYS if r == "o"]

# Blocked entry (include/cloudsc_satadj.h, CLOUDSC_SATADJ_BLOCKED_ARGS): liquid/ice live in species 1 and 2 of the
# five-species fields.  kind as cloudsc.blocked_shape.
SATADJ_BLOCKED = [
    ("pap", "2d", "i"), ("paph", "h", "i"), ("pt", "2d", "i"), ("pq", "2d", "i"), ("pa", "2d", "i"),
    ("pclv", "c5", "i"),
    ("ptendency_cml_t", "2d", "i"), ("ptendency_cml_q", "2d", "i"), ("ptendency_cml_a", "2d", "i"),
    ("ptendency_cml_cld", "c5", "i"),
    ("phrsw", "2d", "i"), ("phrlw", "2d", "i"), ("ptendency_vdf_t", "2d", "i"), ("ptendency_vdf_q", "2d", "i"),
    ("pludeli", "c4", "i"), ("pvervel", "2d", "i"),
    ("ptendency_loc_t", "2d", "o"), ("ptendency_loc_q", "2d", "o"), ("ptendency_loc_a", "2d", "o"),
    ("ptendency_loc_cld", "c5", "o"),
]
# Inputs produced by the CLOUDSC synthetic atmosphere; the rest come from cloudsc_satadj_synth_*.
FROM_CLOUDSC_SYNTH = ["pap", "paph", "pt", "pq", "pa", "pclv", "ptendency_cml_t", "ptendency_cml_q",
                      "ptendency_cml_a", "ptendency_cml_cld", "pludeli", "plude"]
SATADJ_SYNTH = ["phrsw", "phrlw", "ptendency_vdf_t", "ptendency_vdf_q", "pvervel"]


def allocate_block(klon, klev, prec="sp", fill=np.nan):
    """Fortran-ordered arrays for one CLOUD_SATADJ call; outputs pre-filled with ``fill``."""
    dt = C._real(prec)
    out = {}
    for name, kind, role in SATADJ_ARRAYS:
        if role == "u":
            continue
        shape = C.fortran_shape(kind, klon, klev)
        out[name] = np.full(shape, fill, dtype=dt, order="F") if role == "o" else np.zeros(shape, dtype=dt, order="F")
    return out


def call_block(func, params, kidia, kfdia, klon, klev, ptsphy, arrays, prec, lddiag=False):
    """Call a CLOUD_SATADJ-signature C function (product or oracle) on numpy arrays."""
    dt = C._real(prec)
    args = [ctypes.byref(params), ctypes.c_int(kidia), ctypes.c_int(kfdia), ctypes.c_int(klon), ctypes.c_int(klev),
            ctypes.c_void_p(0), ctypes.c_int(int(bool(lddiag))), C._ctype_real(prec)(ptsphy)]
    for name, kind, role in SATADJ_ARRAYS:
        a = arrays.get(name)
        if a is None:
            if role != "u":
                raise ValueError(f"missing array {name}")
            args.append(ctypes.c_void_p(0))
            continue
        if a.dtype != dt or not a.flags.f_contiguous or a.shape != C.fortran_shape(kind, klon, klev):
            raise TypeError(f"{name}: need a Fortran-contiguous {np.dtype(dt).name} array of shape "
                            f"{C.fortran_shape(kind, klon, klev)}")
        args.append(ctypes.c_void_p(a.ctypes.data))
    args += [ctypes.c_void_p(0), ctypes.c_int(0)]  # PEXTRA, KFLDX
    return func(*args)


def CLOUD_SATADJ(params, kidia, kfdia, klon, klev, ptsphy, arrays, precision="sp", lddiag=False):
    """One NPROMA block through cloudsc_satadj_{dp,sp,mixed}_c; outputs are written into ``arrays``."""
    if precision not in C.PRECISIONS:
        raise ValueError(precision)
    fn = f"cloudsc_satadj_{precision}_c"
    C._check(fn, call_block(getattr(C.lib(), fn), params, kidia, kfdia, klon, klev, ptsphy, arrays, precision,
                            lddiag))
    return arrays


class SatadjFields(C.BlockedFields):
    """Blocked field set (NPROMA, KLEV[+1], n, NGPBLKS) of CLOUD_SATADJ on one device."""

    def __init__(self, ngptot, nproma, klev, precision="sp", device="cuda", fill=None):
        import torch

        super().__init__(ngptot, nproma, klev, precision, device,
                         names=set(C.SYNTH_NAMES) | {"ptendency_loc_t", "ptendency_loc_q", "ptendency_loc_a",
                                                     "ptendency_loc_cld"}, fill=fill)
        for name in SATADJ_SYNTH:
            self.t[name] = torch.empty(C.blocked_shape("2d", self.nproma, self.klev, self.ngpblks),
                                       dtype=self.dtype, device=device)


def synth_device(fields, params, seed, stream=None):
    """Seeded synthetic inputs: the CLOUDSC atmosphere plus the SATADJ forcings (include/cloudsc_satadj.h)."""
    import torch

    C.synth_device(fields, params, seed, stream)
    if stream is None:
        stream = torch.cuda.current_stream(fields.device).cuda_stream
    fn = "cloudsc_satadj_synth_dp" if fields.precision == "dp" else "cloudsc_satadj_synth_sp"
    ptrs = [ctypes.c_void_p(fields[n].data_ptr()) for n in ["plude"] + SATADJ_SYNTH + ["pludeli"]]
    C._check(fn, getattr(C.lib(), fn)(ctypes.c_uint64(seed), ctypes.c_int(fields.ngptot),
                                      ctypes.c_int(fields.nproma), ctypes.c_int(fields.klev),
                                      ctypes.c_int(fields.ngpblks), *ptrs, ctypes.c_void_p(stream)))
    torch.cuda.synchronize(fields.device)


def blocked_args(arrays, ptr):
    return [ctypes.c_void_p(ptr(arrays[n])) for n, _, _ in SATADJ_BLOCKED]


def satadj_blocked(fields, params, ptsphy, stream=None):
    """One call of cloudsc_satadj_blocked_{dp,sp,mixed}_c over a device-resident SatadjFields."""
    import torch

    if stream is None:
        stream = torch.cuda.current_stream(fields.device).cuda_stream
    fn = f"cloudsc_satadj_blocked_{fields.precision}_c"
    args = [ctypes.byref(params), ctypes.c_int(fields.ngptot), ctypes.c_int(fields.nproma), ctypes.c_int(fields.klev),
            ctypes.c_int(fields.ngpblks), C._ctype_real(fields.precision)(ptsphy)]
    args += blocked_args(fields.t, lambda x: x.data_ptr())
    args.append(ctypes.c_void_p(stream))
    C._check(fn, getattr(C.lib(), fn)(*args))


def algorithmic_bytes(klev, ncldtop, word=4):
    """Algorithmic HBM bytes per column of one blocked call.  On the KLEV-NCLDTOP+1 active levels it reads 18
    level fields (T, Q, A, PAP, L, I, five CML tendencies, PHRSW, PHRLW, two VDF tendencies, PVERVEL,
    PLUDELI(:,:,3:4)) and the half levels NCLDTOP..KLEV+1 of PAPH (the surface value included); it writes the
    five tendencies on all KLEV levels (zeros above NCLDTOP)."""
    active = klev - ncldtop + 1
    return word * (18 * active + (active + 1) + 5 * klev)


# ---------------------------------------------------------------------------
# fused CALLPAR sequence: CLOUD_SATADJ + TENDENCY_CML increment + TENDENCY_SATADJ copy + CLOUDSC, one launch
# ---------------------------------------------------------------------------
SATADJ_EXTRA_INPUTS = ["phrsw", "phrlw", "ptendency_vdf_t", "ptendency_vdf_q", "pvervel"]
SATADJ_TENDENCY_OUT = [("ptendency_satadj_t", "2d"), ("ptendency_satadj_q", "2d"), ("ptendency_satadj_a", "2d"),
                       ("ptendency_satadj_cld", "c5")]


class FusedFields(C.BlockedFields):
    """Every CLOUDSC blocked field plus the CLOUD_SATADJ inputs and the TENDENCY_SATADJ outputs, on one device."""

    def __init__(self, ngptot, nproma, klev, precision="sp", device="cuda", fill=None, satadj_out=True):
        import torch

        super().__init__(ngptot, nproma, klev, precision, device, fill=fill)
        for name in SATADJ_EXTRA_INPUTS:
            self.t[name] = torch.empty(C.blocked_shape("2d", self.nproma, self.klev, self.ngpblks),
                                       dtype=self.dtype, device=device)
        if satadj_out:
            for name, kind in SATADJ_TENDENCY_OUT:
                shape = C.blocked_shape(kind, self.nproma, self.klev, self.ngpblks)
                self.t[name] = (torch.full(shape, fill, dtype=self.dtype, device=device) if fill is not None
                                else torch.empty(shape, dtype=self.dtype, device=device))


def satadj_cloudsc_blocked(fields, params, ptsphy, stream=None):
    """cloudsc_satadj_cloudsc_blocked_{dp,sp,mixed}_c over a FusedFields: the CML tendencies return incremented by
    the saturation adjustment, TENDENCY_LOC holds CLOUDSC's result, TENDENCY_SATADJ the adjustment's."""
    import torch

    if stream is None:
        stream = torch.cuda.current_stream(fields.device).cuda_stream
    fn = f"cloudsc_satadj_cloudsc_blocked_{fields.precision}_c"
    ptr = lambda n: ctypes.c_void_p(fields[n].data_ptr() if n in fields.t else 0)
    args = [ctypes.byref(params), ctypes.c_int(fields.ngptot), ctypes.c_int(fields.nproma), ctypes.c_int(fields.klev),
            ctypes.c_int(fields.ngpblks), C._ctype_real(fields.precision)(ptsphy)]
    args += C.blocked_args(fields.t, lambda x: x.data_ptr())
    args += [ptr(n) for n in SATADJ_EXTRA_INPUTS]
    args += [ptr(n) for n, _ in SATADJ_TENDENCY_OUT]
    args.append(ctypes.c_void_p(stream))
    C._check(fn, getattr(C.lib(), fn)(*args))


def fused_bytes_per_column(klev, ncldtop, word=4):
    """Algorithmic HBM bytes per column of the fused launch: CLOUDSC's 49*KLEV+20 words, plus on the active levels
    the seven SATADJ-only inputs (PHRSW, PHRLW, two VDF tendencies, PVERVEL, PLUDELI(:,:,3:4)) and the five
    incremented CML tendencies written back, plus TENDENCY_SATADJ (five fields, all levels)."""
    active = klev - ncldtop + 1
    return word * ((49 * klev + 20) + 7 * active + 5 * active + 5 * klev)
