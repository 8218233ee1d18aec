"""Host-side interface to the gfx950 CLOUDSC (libcloudsc_hip.so).

Mirrors the reference interface of SUBROUTINE CLOUDSC
(ifs-source/arpifs/phys_ec/cloudsc.F90:14-44): the argument names, their
order and their meaning are the reference's; arrays are Fortran-ordered
(JL fastest).  Three entry levels:

* ``CLOUDSC(...)`` -- one NPROMA block with host (numpy) arrays, the
  positional drop-in (``cloudsc_{dp,sp,mixed}_c``).
* ``cloudsc_blocked(...)`` -- all blocks of an IFS blocked field set resident
  on the GPU (torch tensors), one kernel launch (``cloudsc_blocked_*_c``).
* ``BlockedFields`` -- allocation of such a field set on the device, and
  ``synth_device`` / ``synth_host`` to fill it with a deterministic
  synthetic atmosphere.

Every call goes to the native library; there is no fallback.  A missing or
unloadable library raises ``CloudscLibraryError``, a nonzero status of the
native call raises ``CloudscError`` (the reference's equivalent is ABOR1).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcloudsc_hip.so")

NCLV = 5
PRECISIONS = ("dp", "sp", "mixed")


class CloudscLibraryError(RuntimeError):
    pass


class CloudscError(RuntimeError):
    def __init__(self, func, code, msg):
        super().__init__(f"{func} returned {code}: {msg}")
        self.code = code


_PARAM_FIELDS = (
    "rg rd rv rcpd retv rtt rlvtt rlstt rlmlt ra rpi "
    "r2es r3les r3ies r4les r4ies r5les r5ies r5alvcp r5alscp "
    "ralvdcp ralsdcp ralfdcp rtwat rtice rtwat_rtice_r rkoop1 rkoop2 "
    "ramid ramin rlmin rcldiff rcldiff_convi rclcrit_sea rclcrit_land "
    "rkconv rprc1 rcl_inhomogaut rcl_inhomogacc "
    "rsnowlin1 rsnowlin2 rlcritsnow rvice rvrain rvsnow "
    "rxfallspeedrain rxfallspeedice rnice rccn "
    "rcldtopcf rdepliqrefrate rdepliqrefdepth rcl_overlapliqice rcl_effrime "
    "riceinit rthomo rprecrhmax rpecons rvrfactor rcovpmin "
    "rtaumel rkooptau rssicefactor rdenswat rdensref "
    "rcl_dr rcl_x1r rcl_x2r rcl_x1s rcl_lam1r rcl_lam2r rcl_lam1s rcl_lam2s "
    "rcl_cdenom1 rcl_cdenom2 rcl_cdenom3 rcl_ka273 "
    "rcl_const1r rcl_const2r rcl_const3r rcl_const4r rcl_const5r "
    "rcl_const6r rcl_const7r rcl_const9r rcl_const10r rcl_eff_racw "
    "rcl_kkaac rcl_kkbac rcl_kkaau rcl_kkbauq rcl_kkbaun "
    "rcl_kk_cloud_num_sea rcl_kk_cloud_num_land rcl_fzrab "
    "rcl_const1i rcl_const2i rcl_const3i rcl_const4i rcl_const5i rcl_const6i rcl_const7i "
    "rcl_apb1 rcl_apb2 rcl_apb3 "
    "rcl_const1s rcl_const2s rcl_const3s rcl_const4s rcl_const5s "
    "rcl_const6s rcl_const7s rcl_const8s "
    "rmfadvw rmfadvwdd rcloud_frac_std reisthsc"
).split()
_PARAM_INTS = "ncldtop lmfdsnow lcloud_inhomog lrad_cloud_inhomog laeraware lspp lcldbud pad_".split()


class CloudscParams(ctypes.Structure):
    """cloudsc_params_t of include/cloudsc_hip.h: the flattened TECLDP/TTHF/TCST/... configuration."""

    _fields_ = [(n, ctypes.c_double) for n in _PARAM_FIELDS] + [(n, ctypes.c_int32) for n in _PARAM_INTS]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad_"}


# ---------------------------------------------------------------------------
# Argument lists (reference order).  kind: 2d=(KLON,KLEV), h=(KLON,KLEV+1),
# c5/c4/c2=(KLON,KLEV,n), 1d=(KLON,), role: i=IN, o=OUT, io=INOUT, u=unused.
# ---------------------------------------------------------------------------
CLOUDSC_ARRAYS = [
    ("pt", "2d", "i"), ("pq", "2d", "i"),
    ("ptendency_cml_t", "2d", "i"), ("ptendency_cml_q", "2d", "i"), ("ptendency_cml_a", "2d", "i"),
    ("ptendency_cml_cld", "c5", "i"),
    ("ptendency_loc_t", "2d", "o"), ("ptendency_loc_q", "2d", "o"), ("ptendency_loc_a", "2d", "o"),
    ("ptendency_loc_cld", "c5", "o"),
    ("pvfa", "2d", "u"), ("pvfl", "2d", "i"), ("pvfi", "2d", "i"),
    ("pdyna", "2d", "u"), ("pdynl", "2d", "u"), ("pdyni", "2d", "u"), ("pdynr", "2d", "u"), ("pdyns", "2d", "u"),
    ("phrsw", "2d", "u"), ("phrlw", "2d", "u"),
    ("pvervel", "2d", "u"), ("pap", "2d", "i"), ("paph", "h", "i"), ("pgeom", "2d", "u"),
    ("plsm", "1d", "i"), ("pgaw", "1d", "u"), ("ldcum", "1di", "i"), ("kctop", "1di", "u"),
    ("ktype", "1di", "i"), ("kpbltype", "1di", "u"), ("peis", "1d", "io"),
    ("plu", "2d", "i"), ("plude", "2d", "io"), ("pludeli", "c4", "i"), ("psnde", "c2", "i"),
    ("pmfu", "2d", "i"), ("pmfd", "2d", "i"), ("pgp2dspp", "1d", "u"),
    ("ldslphy", "scalar", "u"),
    ("pa", "2d", "i"), ("pclv", "c5", "i"),
    ("plcrit_aer", "2d", "u"), ("picrit_aer", "2d", "u"), ("pre_ice", "2d", "u"), ("pccn", "2d", "u"),
    ("pnice", "2d", "u"),
    ("pcovptot", "2d", "o"), ("pfsd", "2d", "u"), ("prainfrac_toprfz", "1d", "o"),
    ("pfsqlf", "h", "o"), ("pfsqif", "h", "o"), ("pfcqnng", "h", "o"), ("pfcqlng", "h", "o"),
    ("pfsqrf", "h", "o"), ("pfsqsf", "h", "o"), ("pfcqrng", "h", "o"), ("pfcqsng", "h", "o"),
    ("pfsqltur", "h", "o"), ("pfsqitur", "h", "o"),
    ("pfplsl", "h", "o"), ("pfplsn", "h", "o"), ("pfhpsl", "h", "o"), ("pfhpsn", "h", "o"),
    ("pextra", "extra", "u"),
]
# Arrays of the blocked entry, in the order of CLOUDSC_BLOCKED_ARGS.
BLOCKED_ARRAYS = [a for a in CLOUDSC_ARRAYS if a[2] != "u"]
INPUT_NAMES = [n for n, k, r in BLOCKED_ARRAYS if r in ("i", "io")]
OUTPUT_NAMES = [n for n, k, r in BLOCKED_ARRAYS if r in ("o", "io")]
SYNTH_NAMES = ["pt", "pq", "ptendency_cml_t", "ptendency_cml_q", "ptendency_cml_a", "ptendency_cml_cld", "pvfl",
               "pvfi", "pap", "paph", "plsm", "ldcum", "ktype", "peis", "plu", "plude", "pludeli", "psnde", "pmfu",
               "pmfd", "pa", "pclv"]
KIND_NCOMP = {"c5": 5, "c4": 4, "c2": 2}


def fortran_shape(kind, klon, klev):
    if kind == "2d":
        return (klon, klev)
    if kind == "h":
        return (klon, klev + 1)
    if kind in KIND_NCOMP:
        return (klon, klev, KIND_NCOMP[kind])
    if kind in ("1d", "1di"):
        return (klon,)
    raise ValueError(kind)


def blocked_shape(kind, nproma, klev, ngpblks):
    """C-order shape of a blocked field (Fortran (NPROMA,...,NGPBLKS))."""
    if kind == "2d":
        return (ngpblks, klev, nproma)
    if kind == "h":
        return (ngpblks, klev + 1, nproma)
    if kind in KIND_NCOMP:
        return (ngpblks, KIND_NCOMP[kind], klev, nproma)
    if kind in ("1d", "1di"):
        return (ngpblks, nproma)
    raise ValueError(kind)


_lib = None


def lib():
    """The loaded libcloudsc_hip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CloudscLibraryError(f"{LIB_PATH} is not built (run python -m openifs_amd.build)")
        # One HIP runtime per process: PyTorch ships its own libamdhip64 (SONAME libamdhip64.so.7, needed as
        # libamdhip64.so).  Loaded first, it also satisfies this library's libamdhip64.so.7, so torch tensors and
        # streams and this library share one runtime.  Loaded after /opt/rocm's copy, it would start a second
        # runtime that finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        try:
            _lib = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise CloudscLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        _lib.cloudsc_last_error.restype = ctypes.c_char_p
        _lib.cloudsc_params_default.argtypes = [ctypes.POINTER(CloudscParams), ctypes.c_double, ctypes.c_int,
                                                ctypes.c_void_p]
        # launch configurations recorded by the autotuner (openifs_amd/tuning/cloudsc_tuned.json), if any
        if hasattr(_lib, "cloudsc_tune_set_block"):
            from . import autotune

            autotune.apply(arch="gfx950")
    return _lib


def _check(func, code):
    if code != 0:
        raise CloudscError(func, code, lib().cloudsc_last_error().decode())


# ---------------------------------------------------------------------------
# vertical levels and the default configuration
# ---------------------------------------------------------------------------
def hybrid_levels(klev):
    """(A, B) half-level coefficients for 137 (reference L137) or 91 (project L91) levels."""
    fn = {137: "l137_ab.txt", 91: "l91_ab.txt"}.get(klev)
    if fn is None:
        raise ValueError(f"no hybrid level table for KLEV={klev} (have 137 and 91)")
    d = np.loadtxt(os.path.join(_HERE, "data", fn))
    return d[:, 1].copy(), d[:, 2].copy()


def standard_full_levels(ahalf, bhalf, psurf=101325.0):
    """Full-level pressures of the standard atmosphere (Simmons-Burridge, top level = PAPH(2)/2)."""
    ph = ahalf + bhalf * psurf
    p = np.empty(ph.size - 1)
    p[0] = 0.5 * ph[1]
    pp, pm = ph[2:], ph[1:-1]
    p[1:] = np.exp((pp * np.log(pp) - pm * np.log(pm)) / (pp - pm) - 1.0)
    return p


def default_params(ptsphy=900.0, klev=137, stpre=None):
    """cloudsc_params_t with the reference defaults (cloudsc_params_default)."""
    if stpre is None:
        a, b = hybrid_levels(klev)
        stpre = standard_full_levels(a, b)
    stpre = np.ascontiguousarray(stpre, dtype=np.float64)
    p = CloudscParams()
    _check("cloudsc_params_default", lib().cloudsc_params_default(ctypes.byref(p), float(ptsphy), int(stpre.size),
                                                                  stpre.ctypes.data))
    return p


def _real(prec):
    return np.float64 if prec == "dp" else np.float32


def _ctype_real(prec):
    return ctypes.c_double if prec == "dp" else ctypes.c_float


# ---------------------------------------------------------------------------
# per-block drop-in (host arrays)
# ---------------------------------------------------------------------------
def allocate_block(klon, klev, prec="sp", kfldx=0, fill=np.nan):
    """Fortran-ordered arrays for one CLOUDSC call; outputs pre-filled with ``fill``."""
    dt = _real(prec)
    out = {}
    for name, kind, role in CLOUDSC_ARRAYS:
        if kind == "scalar":
            continue
        if kind == "extra":
            out[name] = np.zeros((klon, klev, max(kfldx, 1)), dtype=dt, order="F")
            continue
        shape = fortran_shape(kind, klon, klev)
        if kind == "1di":
            out[name] = np.zeros(shape, dtype=np.int32, order="F")
        elif role == "o":
            out[name] = np.full(shape, fill, dtype=dt, order="F")
        else:
            out[name] = np.zeros(shape, dtype=dt, order="F")
    return out


def call_block(func, params, kidia, kfdia, klon, klev, ptsphy, arrays, prec, lddiag=False, ldslphy=False, kfldx=0):
    """Call a CLOUDSC-signature C function (product or oracle) on numpy arrays."""
    dt = _real(prec)
    args = [ctypes.byref(params), ctypes.c_int(kidia), ctypes.c_int(kfdia), ctypes.c_int(klon), ctypes.c_int(klev),
            ctypes.c_int(int(bool(lddiag))), _ctype_real(prec)(ptsphy)]
    for name, kind, role in CLOUDSC_ARRAYS:
        if kind == "scalar":
            args.append(ctypes.c_int(int(bool(ldslphy))))
            continue
        a = arrays.get(name)
        if a is None:
            if role != "u":
                raise ValueError(f"missing array {name}")
            args.append(ctypes.c_void_p(0))
            continue
        want = np.int32 if kind == "1di" else dt
        if a.dtype != want or not a.flags.f_contiguous:
            raise TypeError(f"{name}: need a Fortran-contiguous {np.dtype(want).name} array")
        args.append(ctypes.c_void_p(a.ctypes.data))
    args.append(ctypes.c_int(kfldx))
    return func(*args)


def CLOUDSC(params, kidia, kfdia, klon, klev, ptsphy, arrays, precision="sp", lddiag=False, ldslphy=False, kfldx=0):
    """One NPROMA block through cloudsc_{dp,sp,mixed}_c; ``arrays`` maps reference argument names
    (lower case) to Fortran-ordered numpy arrays, updated in place."""
    if precision not in PRECISIONS:
        raise ValueError(precision)
    f = getattr(lib(), f"cloudsc_{precision}_c")
    _check(f"cloudsc_{precision}_c",
           call_block(f, params, kidia, kfdia, klon, klev, ptsphy, arrays, precision, lddiag, ldslphy, kfldx))
    return arrays


# ---------------------------------------------------------------------------
# blocked, device-resident path (torch tensors)
# ---------------------------------------------------------------------------
def last_warnings():
    """(negative-liquid, negative-ice) counts of the calling thread's last per-block CLOUDSC call: the points where
    the reference writes 'CLOUDSC: WARNING Negative cloud liquid!' / 'ice!' (cloudsc.F90:3794, :3802)."""
    nl, ni = ctypes.c_int64(), ctypes.c_int64()
    _check("cloudsc_last_warnings", lib().cloudsc_last_warnings(ctypes.byref(nl), ctypes.byref(ni)))
    return nl.value, ni.value


class BlockedFields:
    """An IFS blocked field set (NPROMA, KLEV[+1], n, NGPBLKS) on one device, as torch tensors."""

    def __init__(self, ngptot, nproma, klev, precision="sp", device="cuda", names=None, fill=None):
        import torch

        self.ngptot, self.nproma, self.klev = int(ngptot), int(nproma), int(klev)
        self.ngpblks = (self.ngptot + self.nproma - 1) // self.nproma
        self.precision = precision
        self.dtype = torch.float64 if precision == "dp" else torch.float32
        self.device = device
        self.t = {}
        for name, kind, role in BLOCKED_ARRAYS:
            if names is not None and name not in names:
                continue
            shape = blocked_shape(kind, self.nproma, self.klev, self.ngpblks)
            dt = torch.int32 if kind == "1di" else self.dtype
            if fill is not None and role == "o":
                self.t[name] = torch.full(shape, fill, dtype=dt, device=device)
            else:
                self.t[name] = torch.empty(shape, dtype=dt, device=device)

    def nbytes(self):
        return sum(x.numel() * x.element_size() for x in self.t.values())

    def __getitem__(self, k):
        return self.t[k]


def synth_desc(seed, ngptot, nproma, klev, ahalf_ptr, bhalf_ptr):
    class Desc(ctypes.Structure):
        _fields_ = [("seed", ctypes.c_uint64), ("ngptot", ctypes.c_int32), ("nproma", ctypes.c_int32),
                    ("klev", ctypes.c_int32), ("ngpblks", ctypes.c_int32), ("ahalf", ctypes.c_void_p),
                    ("bhalf", ctypes.c_void_p)]

    ngpblks = (ngptot + nproma - 1) // nproma
    return Desc(seed, ngptot, nproma, klev, ngpblks, ahalf_ptr, bhalf_ptr)


def synth_device(fields, params, seed, stream=None):
    """Fill the input tensors of ``fields`` with the synthetic atmosphere of ``seed`` (on the device)."""
    import torch

    a, b = hybrid_levels(fields.klev)
    ad = torch.tensor(a, dtype=torch.float64, device=fields.device)
    bd = torch.tensor(b, dtype=torch.float64, device=fields.device)
    d = synth_desc(seed, fields.ngptot, fields.nproma, fields.klev, ad.data_ptr(), bd.data_ptr())
    ptrs = [ctypes.c_void_p(fields[n].data_ptr()) for n in SYNTH_NAMES]
    if stream is None:
        stream = torch.cuda.current_stream(fields.device).cuda_stream
    fn = "cloudsc_synth_dp" if fields.precision == "dp" else "cloudsc_synth_sp"
    _check(fn, getattr(lib(), fn)(ctypes.byref(params), ctypes.byref(d), *ptrs, ctypes.c_void_p(stream)))
    torch.cuda.synchronize(fields.device)
    del ad, bd


def synth_host(params, seed, ngptot, nproma, klev, precision="dp", col0=0, ncol=None):
    """Synthetic atmosphere on the host, blocked layout (dict of C-order numpy arrays)."""
    dt = _real(precision)
    ngpblks = (ngptot + nproma - 1) // nproma
    if ncol is None:
        ncol = ngpblks * nproma - col0
    out = {}
    for name, kind, role in BLOCKED_ARRAYS:
        if name not in SYNTH_NAMES:
            continue
        shape = blocked_shape(kind, nproma, klev, ngpblks)
        out[name] = np.zeros(shape, dtype=np.int32 if kind == "1di" else dt)
    a, b = hybrid_levels(klev)
    d = synth_desc(seed, ngptot, nproma, klev, a.ctypes.data, b.ctypes.data)
    ptrs = [ctypes.c_void_p(out[n].ctypes.data) for n in SYNTH_NAMES]
    fn = "cloudsc_synth_host_dp" if precision == "dp" else "cloudsc_synth_host_sp"
    _check(fn, getattr(lib(), fn)(ctypes.byref(params), ctypes.byref(d), ctypes.c_int64(col0), ctypes.c_int64(ncol),
                                  *ptrs))
    return out


def blocked_args(fields_or_arrays, ptr):
    return [ctypes.c_void_p(ptr(fields_or_arrays[n])) for n, _, _ in BLOCKED_ARRAYS]


def cloudsc_blocked(fields, params, ptsphy, stream=None):
    """One launch of cloudsc_blocked_{dp,sp,mixed}_c over a device-resident BlockedFields."""
    import torch

    if stream is None:
        stream = torch.cuda.current_stream(fields.device).cuda_stream
    fn = f"cloudsc_blocked_{fields.precision}_c"
    f = getattr(lib(), fn)
    args = [ctypes.byref(params), ctypes.c_int(fields.ngptot), ctypes.c_int(fields.nproma), ctypes.c_int(fields.klev),
            ctypes.c_int(fields.ngpblks), _ctype_real(fields.precision)(ptsphy)]
    args += blocked_args(fields.t, lambda x: x.data_ptr())
    args.append(ctypes.c_void_p(stream))
    _check(fn, f(*args))


# ---------------------------------------------------------------------------
# budget diagnostics (LCLDBUD*) into PEXTRA
# ---------------------------------------------------------------------------
# Bits of cloudsc_params_t.lcldbud: the LCLDBUD* switches (cloudsc.F90:706-713), LBUD23 and LCLDBUDGET.
BUD_C, BUD_L, BUD_I, BUD_T, BUD_VERTINT, BUD_TIMEINT, BUD_LBUD23, BUD_CHECK = 1, 2, 4, 8, 16, 32, 64, 128
# Arrays only the budgets read, in the argument order of cloudsc_blocked_diag_*_c (PEXTRA and KFLDX follow, then
# PGAW and PFSD).
BUDGET_ARRAYS = [("pvfa", "2d"), ("pdyna", "2d"), ("pdynl", "2d"), ("pdyni", "2d"), ("pdyns", "2d"),
                 ("phrsw", "2d"), ("phrlw", "2d"), ("kpbltype", "1di"), ("kctop", "1di")]


def budget_fields(flags):
    """PEXTRA fields the switches fill (IS after the last budget block, cloudsc.F90:3832-4101)."""
    n = 26 if flags & BUD_LBUD23 else 0
    for bit, k in ((BUD_VERTINT, 1), (BUD_C, 12), (BUD_L, 22), (BUD_I, 20), (BUD_T, 13)):
        if flags & bit:
            n += k
    return n


def extra_shape(nproma, klev, kfldx, ngpblks):
    """C-order shape of the blocked PEXTRA (Fortran (NPROMA,KLEV,KFLDX,NGPBLKS))."""
    return (ngpblks, kfldx, klev, nproma)


def cloudsc_blocked_diag(fields, diag, params, ptsphy, stream=None):
    """One launch of cloudsc_blocked_diag_{dp,sp,mixed}_c: ``fields`` a BlockedFields, ``diag`` a dict of device
    tensors: the BUDGET_ARRAYS and ``pextra`` (extra_shape) for the budgets, which accumulate into it, and
    ``pgaw`` (NGPBLKS, NPROMA) and ``pfsd`` (NGPBLKS, KLEV, NPROMA) for LRAD_CLOUD_INHOMOG.  Arrays of a diagnostic
    that is off may be left out."""
    import torch

    if stream is None:
        stream = torch.cuda.current_stream(fields.device).cuda_stream
    fn = f"cloudsc_blocked_diag_{fields.precision}_c"
    px = diag.get("pextra")
    kfldx = 0
    if px is not None:
        if px.dim() != 4 or (px.shape[0], px.shape[2], px.shape[3]) != (fields.ngpblks, fields.klev, fields.nproma):
            raise ValueError(f"pextra: need shape (NGPBLKS, KFLDX, KLEV, NPROMA), got {tuple(px.shape)}")
        kfldx = px.shape[1]
    ptr = lambda n: ctypes.c_void_p(diag[n].data_ptr() if diag.get(n) is not None else 0)
    args = [ctypes.byref(params), ctypes.c_int(fields.ngptot), ctypes.c_int(fields.nproma), ctypes.c_int(fields.klev),
            ctypes.c_int(fields.ngpblks), _ctype_real(fields.precision)(ptsphy)]
    args += blocked_args(fields.t, lambda x: x.data_ptr())
    args += [ptr(n) for n, _ in BUDGET_ARRAYS]
    args += [ptr("pextra"), ctypes.c_int(kfldx), ptr("pgaw"), ptr("pfsd"), ctypes.c_void_p(stream)]
    _check(fn, getattr(lib(), fn)(*args))
