"""Host-side interface to the gfx950 CLOUDSC2 (include/cloudsc2.h).

Mirrors SUBROUTINE CLOUDSC2 (ifs-source/arpifs/phys_ec/cloudsc2.F90:10-17): argument names, order and meaning are
the reference's.  ``CLOUDSC2`` runs one NPROMA block on Fortran-ordered numpy arrays (``cloudsc2_{dp,sp}_c``);
``cloudsc2_blocked`` runs all blocks of device-resident blocked fields held as torch tensors
(``cloudsc2_blocked_{dp,sp}_c``).  Every call goes to the native library; there is no fallback.
"""
import ctypes

import numpy as np

from openifs_amd import cloudsc as C

PRECISIONS = ("dp", "sp")


class Cloudsc2Params(ctypes.Structure):
    """cloudsc2_params_t: the TECLDP/TEPHLI/TPHNC/TTHF components CLOUDSC2 reads besides cloudsc_params_t."""

    _fields_ = [("rclcrit", ctypes.c_double), ("rlptrc", ctypes.c_double), ("rlpal1", ctypes.c_double),
                ("rlpal2", ctypes.c_double), ("rvtmp2", ctypes.c_double), ("lphylin", ctypes.c_int32),
                ("levapls2", ctypes.c_int32), ("lepcld2", ctypes.c_int32), ("pad_", ctypes.c_int32)]


def default_params2(params):
    p2 = Cloudsc2Params()
    C._check("cloudsc2_params_default", C.lib().cloudsc2_params_default(ctypes.byref(p2), ctypes.byref(params)))
    return p2


# (name, kind, role) in argument order after KLEV/LDRAIN1D/PTSPHY; kind: 2d = (KLON,KLEV), h = (KLON,KLEV+1)
CLOUDSC2_ARRAYS = [
    ("paphp1", "h", "i"), ("papp1", "2d", "i"), ("pqm1", "2d", "i"), ("pqs", "2d", "i"), ("ptm1", "2d", "i"),
    ("pl", "2d", "i"), ("pi", "2d", "i"), ("plude", "2d", "i"), ("plu", "2d", "i"), ("pmfu", "2d", "i"),
    ("pmfd", "2d", "i"), ("ptent", "2d", "o"), ("pgtent", "2d", "i"), ("ptenq", "2d", "o"), ("pgtenq", "2d", "i"),
    ("ptenl", "2d", "o"), ("pgtenl", "2d", "i"), ("pteni", "2d", "o"), ("pgteni", "2d", "i"), ("pclc", "2d", "o"),
    ("pfplsl", "h", "o"), ("pfplsn", "h", "o"), ("pfhpsl", "h", "o"), ("pfhpsn", "h", "o"), ("pcovptot", "2d", "o"),
]
INPUT_NAMES = [n for n, k, r in CLOUDSC2_ARRAYS if r == "i"]
OUTPUT_NAMES = [n for n, k, r in CLOUDSC2_ARRAYS if r == "o"]


def reference_eta(klev):
    """CETA: full-level pressure of the standard atmosphere (surface VP00 = 101325 Pa) over VP00
    (suphec.F90:261-271, sucld.F90:125-127)."""
    a, b = C.hybrid_levels(klev)
    ps = 101325.0
    return C.standard_full_levels(a, b, ps) / ps


def _real(prec):
    return np.float64 if prec == "dp" else np.float32


def _ctype_real(prec):
    return ctypes.c_double if prec == "dp" else ctypes.c_float


def block_args(params, params2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d, ptsphy, arrays, precision):
    """ctypes argument list of CLOUDSC2_ARGS for Fortran-ordered numpy ``arrays``; ``ceta`` must stay alive."""
    if precision not in PRECISIONS:
        raise ValueError(precision)
    dt = _real(precision)
    if ceta.dtype != dt or ceta.shape != (klev,):
        raise TypeError(f"ceta: need a {np.dtype(dt).name} array of KLEV values")
    args = [ctypes.byref(params), ctypes.byref(params2), ctypes.c_void_p(ceta.ctypes.data), ctypes.c_int(kidia),
            ctypes.c_int(kfdia), ctypes.c_int(klon), ctypes.c_int(ktdia), ctypes.c_int(klev),
            ctypes.c_int(int(bool(ldrain1d))), _ctype_real(precision)(ptsphy)]
    for name, kind, role in CLOUDSC2_ARRAYS:
        a = arrays[name]
        want = (klon, klev + 1) if kind == "h" else (klon, klev)
        if a.dtype != dt or not a.flags.f_contiguous or a.shape != want:
            raise TypeError(f"{name}: need a Fortran-contiguous {np.dtype(dt).name} array of shape {want}")
        args.append(ctypes.c_void_p(a.ctypes.data))
    return args


def CLOUDSC2(params, params2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d, ptsphy, arrays, precision="sp"):
    """One NPROMA block through cloudsc2_{dp,sp}_c; ``arrays`` maps the reference's argument names (lower case) to
    Fortran-ordered numpy arrays, outputs updated in place."""
    ceta = np.ascontiguousarray(ceta, dtype=_real(precision))
    args = block_args(params, params2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d, ptsphy, arrays, precision)
    fn = f"cloudsc2_{precision}_c"
    C._check(fn, getattr(C.lib(), fn)(*args))
    return arrays


def check_device_fields(fields, spec, nproma, klev, ngpblks, precision, ceta=None, shape=None):
    """Validates device tensors before a blocked launch: every array of ``spec`` ((name, kind, ...) tuples) that is
    present must be a contiguous CUDA tensor of the precision's dtype and of its blocked shape, and ``ceta`` a
    contiguous CUDA tensor of KLEV values of that dtype.  A wrong tensor would make the kernel index past its
    allocation, so this raises TypeError before any launch."""
    import torch

    if precision not in PRECISIONS:
        raise ValueError(precision)
    dt = torch.float64 if precision == "dp" else torch.float32
    shape = shape or blocked_shape

    def need(name, t, want):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dt or not t.is_contiguous() \
                or tuple(t.shape) != tuple(want):
            got = (f"{tuple(t.shape)} {t.dtype} {'cuda' if t.is_cuda else 'cpu'} "
                   f"{'contiguous' if t.is_contiguous() else 'strided'}") if isinstance(t, torch.Tensor) else type(t)
            raise TypeError(f"{name}: need a contiguous CUDA {dt} tensor of shape {tuple(want)}, got {got}")

    if ceta is not None:
        need("ceta", ceta, (klev,))
    for entry in spec:
        name, kind = entry[0], entry[1]
        if name in fields:
            need(name, fields[name], shape(kind, nproma, klev, ngpblks))


def cloudsc2_blocked(fields, params, params2, ceta, ngptot, nproma, klev, ptsphy, precision="sp", ktdia=1,
                     ldrain1d=False, stream=None):
    """One launch of cloudsc2_blocked_{dp,sp}_c: ``fields`` maps the argument names to device tensors in blocked
    layout ((NGPBLKS, KLEV[+1], NPROMA) in C order), ``ceta`` a device tensor of KLEV values."""
    import torch

    ngpblks = (ngptot + nproma - 1) // nproma
    missing = [n for n, _, _ in CLOUDSC2_ARRAYS if n not in fields]
    if missing:
        raise KeyError(f"cloudsc2_blocked: missing {missing}")
    check_device_fields(fields, CLOUDSC2_ARRAYS, nproma, klev, ngpblks, precision, ceta)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    args = [ctypes.byref(params), ctypes.byref(params2), ctypes.c_void_p(ceta.data_ptr()), ctypes.c_int(ngptot),
            ctypes.c_int(nproma), ctypes.c_int(ktdia), ctypes.c_int(klev), ctypes.c_int(ngpblks),
            ctypes.c_int(int(bool(ldrain1d))), _ctype_real(precision)(ptsphy)]
    args += [ctypes.c_void_p(fields[n].data_ptr()) for n, _, _ in CLOUDSC2_ARRAYS]
    args.append(ctypes.c_void_p(stream))
    fn = f"cloudsc2_blocked_{precision}_c"
    C._check(fn, getattr(C.lib(), fn)(*args))


def blocked_shape(kind, nproma, klev, ngpblks):
    return (ngpblks, klev + 1, nproma) if kind == "h" else (ngpblks, klev, nproma)


def _qsat(params, t, p, xp=np):
    """Saturation specific humidity over the mixed phase (the FOEEWM form of SATUR), for synthetic inputs."""
    q = params
    a = xp.clip(((xp.clip(t, q.rtice, q.rtwat) - q.rtice) * q.rtwat_rtice_r) ** 2, None, 1.0)
    e = q.r2es * (a * xp.exp(q.r3les * (t - q.rtt) / (t - q.r4les)) +
                  (1 - a) * xp.exp(q.r3ies * (t - q.rtt) / (t - q.r4ies)))
    qs = xp.clip(e / p, None, 0.5)
    return qs / (1.0 - q.retv * qs)


def _derive(params, s, u, ptsphy, xp):
    """CLOUDSC2 inputs from a CLOUDSC synthetic atmosphere ``s`` (blocked arrays) and uniform draws ``u`` in [0, 1)
    of the shape of PMFD."""
    out = {"paphp1": s["paph"], "papp1": s["pap"], "pqm1": xp.clip(s["pq"], None, 0.02), "ptm1": s["pt"],
           "pl": s["pclv"][:, 0], "pi": s["pclv"][:, 1], "plu": s["plu"], "pmfu": s["pmfu"], "pmfd": s["pmfd"],
           "plude": s["plude"], "pgtent": s["ptendency_cml_t"], "pgtenq": s["ptendency_cml_q"],
           "pgtenl": s["ptendency_cml_cld"][:, 0], "pgteni": s["ptendency_cml_cld"][:, 1]}
    # humidity bounded by 20 g/kg: the CLOUDSC atmosphere sets q from a relative humidity, which is unbounded
    # where the saturation mixing ratio is clipped in the upper stratosphere.  The downdraught mass flux is
    # negative; convection is confined below 150 hPa, as mass fluxes at every level would feed the compensating
    # subsidence term without bound where the density is small.
    out["pmfd"] = -abs(out["pmfd"]) * (0.5 * u)
    conv = out["papp1"] > 15000.0
    for n in ("pmfu", "pmfd", "plu", "plude"):
        out[n] = xp.where(conv, out[n], 0.0 * out[n])
    # PQS as CLOUD_S_LAYER recomputes it with SATUR at T + PTSPHY * PGTENT (cloud_s_layer.F90:113-118)
    out["pqs"] = _qsat(params, out["ptm1"] + ptsphy * out["pgtent"], out["papp1"], xp)
    return out


def synth_host(params, seed, ngptot, nproma, klev, precision="dp", ptsphy=900.0):
    """Synthetic blocked inputs of CLOUDSC2 from the CLOUDSC synthetic atmosphere of ``seed`` (dict of C-order
    numpy arrays; outputs NaN): state, convective fields and accumulated tendencies as CLOUD_S_LAYER passes them."""
    s = C.synth_host(params, seed, ngptot, nproma, klev, precision="dp")
    u = np.random.default_rng(seed + 104729).uniform(0.0, 1.0, s["pmfd"].shape)
    dt = _real(precision)
    res = {k: np.ascontiguousarray(v, dtype=dt) for k, v in _derive(params, s, u, ptsphy, np).items()}
    ngpblks = (ngptot + nproma - 1) // nproma
    for name, kind, role in CLOUDSC2_ARRAYS:
        if role == "o":
            res[name] = np.full(blocked_shape(kind, nproma, klev, ngpblks), np.nan, dtype=dt)
    return res


def synth_device(params, seed, ngptot, nproma, klev, precision="sp", device="cuda", ptsphy=900.0, fill=float("nan")):
    """synth_host's construction on the device (dict of torch tensors), from the device CLOUDSC atmosphere; the
    PMFD draws come from torch's generator, so the values differ from synth_host's."""
    import torch

    f = C.BlockedFields(ngptot, nproma, klev, precision, device, names=C.SYNTH_NAMES)
    C.synth_device(f, params, seed)
    g = torch.Generator(device=device).manual_seed(seed + 104729)
    u = torch.rand(f["pmfd"].shape, generator=g, device=device, dtype=f.dtype)
    out = {k: v.contiguous().clone() for k, v in _derive(params, f.t, u, ptsphy, torch).items()}
    del f, u
    ngpblks = (ngptot + nproma - 1) // nproma
    dt = torch.float64 if precision == "dp" else torch.float32
    for name, kind, role in CLOUDSC2_ARRAYS:
        if role == "o":
            out[name] = torch.full(blocked_shape(kind, nproma, klev, ngpblks), fill, dtype=dt, device=device)
    return out


# CLOUD_S_LAYER (LEPCLD2) argument arrays after TSPHY: (name, kind, role, optional); kind as CLOUDSC2_ARRAYS plus
# "1d" = (NPROMA,) per block
LAYER_ARRAYS = [
    ("prs1", "h", "i", False), ("prsf1", "2d", "i", False), ("pq", "2d", "i", False), ("pt", "2d", "i", False),
    ("pa", "2d", "i", True), ("pl", "2d", "i", False), ("pi", "2d", "i", False), ("plude", "2d", "i", False),
    ("plu", "2d", "i", False), ("pmfu", "2d", "i", True), ("pmfd", "2d", "i", True),
    ("ptendency_t", "2d", "i", False), ("ptendency_q", "2d", "i", False), ("ptendency_l", "2d", "i", True),
    ("ptendency_i", "2d", "i", True), ("ptendency_loc_t", "2d", "o", False), ("ptendency_loc_q", "2d", "o", False),
    ("ptendency_loc_a", "2d", "o", True), ("ptendency_loc_l", "2d", "o", False), ("ptendency_loc_i", "2d", "o", False),
    ("ptendency_loc_r", "2d", "o", True), ("ptendency_loc_s", "2d", "o", True), ("pqsat", "2d", "o", True),
    ("pclc", "2d", "o", True), ("pfplsl", "h", "o", False), ("pfplsn", "h", "o", False), ("pfhpsl", "h", "o", False),
    ("pfhpsn", "h", "o", False), ("pcovptot", "2d", "o", False), ("pprectype", "1d", "o", True),
    ("pfzra", "1d", "o", True), ("pqiwp", "2d", "o", True), ("pqlwp", "2d", "o", True),
]
# PMFU, PMFD and PTENDENCY_L/I are read only when LEPCLD2 selects CLOUDSC2

# CLOUDST (cloudst.F90:9-16) arrays after PTSPHY, in the order of CLOUDST_BLOCKED_ARGS: (name, kind, role)
CLOUDST_ARRAYS = [
    ("paphp1", "h", "i"), ("papp1", "2d", "i"), ("pqm1", "2d", "i"), ("pqs", "2d", "i"), ("ptm1", "2d", "i"),
    ("plude", "2d", "i"), ("plu", "2d", "i"), ("ptent", "2d", "o"), ("pgtent", "2d", "i"), ("ptenq", "2d", "o"),
    ("pgtenq", "2d", "i"), ("pclc", "2d", "o"), ("pqiwp", "2d", "o"), ("pqlwp", "2d", "o"), ("pfplsl", "h", "o"),
    ("pfplsn", "h", "o"), ("pfhpsl", "h", "o"), ("pfhpsn", "h", "o"), ("pcovptot", "2d", "o"),
]


def cloudst_blocked(fields, params, params2, ceta, ngptot, nproma, klev, ptsphy, precision="sp", ktdia=1,
                    ldrain1d=False, stream=None):
    """One launch of cloudst_blocked_{dp,sp}_c on device tensors in blocked layout named as CLOUDST_ARRAYS."""
    import torch

    ngpblks = (ngptot + nproma - 1) // nproma
    missing = [n for n, _, _ in CLOUDST_ARRAYS if n not in fields]
    if missing:
        raise KeyError(f"cloudst_blocked: missing {missing}")
    check_device_fields(fields, CLOUDST_ARRAYS, nproma, klev, ngpblks, precision, ceta)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    args = [ctypes.byref(params), ctypes.byref(params2), ctypes.c_void_p(ceta.data_ptr()), ctypes.c_int(ngptot),
            ctypes.c_int(nproma), ctypes.c_int(ktdia), ctypes.c_int(klev), ctypes.c_int(ngpblks),
            ctypes.c_int(int(bool(ldrain1d))), _ctype_real(precision)(ptsphy)]
    args += [ctypes.c_void_p(fields[n].data_ptr()) for n, _, _ in CLOUDST_ARRAYS]
    args.append(ctypes.c_void_p(stream))
    fn = f"cloudst_blocked_{precision}_c"
    C._check(fn, getattr(C.lib(), fn)(*args))
# how CLOUD_S_LAYER passes its fields to CLOUDSC2 (cloud_s_layer.F90:126-139)
LAYER_FROM_CLOUDSC2 = {"prs1": "paphp1", "prsf1": "papp1", "pq": "pqm1", "pt": "ptm1", "pl": "pl", "pi": "pi",
                       "plude": "plude", "plu": "plu", "pmfu": "pmfu", "pmfd": "pmfd", "ptendency_t": "pgtent",
                       "ptendency_q": "pgtenq", "ptendency_l": "pgtenl", "ptendency_i": "pgteni"}


def layer_shape(kind, nproma, klev, ngpblks):
    return (ngpblks, nproma) if kind == "1d" else blocked_shape(kind, nproma, klev, ngpblks)


def layer_args(ptr, fields, required_only=False):
    """Pointers of LAYER_ARRAYS in order; optional arrays missing from ``fields`` are NULL."""
    out = []
    for name, kind, role, opt in LAYER_ARRAYS:
        if name in fields and not (required_only and opt):
            out.append(ctypes.c_void_p(ptr(fields[name])))
        elif opt:
            out.append(ctypes.c_void_p(0))
        else:
            raise KeyError(name)
    return out


def cloud_s_layer_blocked(fields, params, params2, ceta, ngptot, nproma, klev, tsphy, precision="sp", ktdia=1,
                          ldrain1d=False, stream=None):
    """CLOUD_S_LAYER with LEPCLD2 in one launch of cloud_s_layer_blocked_{dp,sp}_c on device tensors in blocked
    layout; optional outputs absent from ``fields`` are skipped."""
    import torch

    ngpblks = (ngptot + nproma - 1) // nproma
    check_device_fields(fields, LAYER_ARRAYS, nproma, klev, ngpblks, precision, ceta, shape=layer_shape)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    args = [ctypes.byref(params), ctypes.byref(params2), ctypes.c_void_p(ceta.data_ptr()), ctypes.c_int(ngptot),
            ctypes.c_int(nproma), ctypes.c_int(ktdia), ctypes.c_int(klev), ctypes.c_int(ngpblks),
            ctypes.c_int(int(bool(ldrain1d))), _ctype_real(precision)(tsphy)]
    args += layer_args(lambda t: t.data_ptr(), fields)
    args.append(ctypes.c_void_p(stream))
    fn = f"cloud_s_layer_blocked_{precision}_c"
    C._check(fn, getattr(C.lib(), fn)(*args))


def layer_host(c2_inputs, pa, precision="dp"):
    """CLOUD_S_LAYER input arrays from CLOUDSC2 inputs (synth_host) and a cloud fraction PA; outputs NaN."""
    dt = _real(precision)
    res = {lay: np.ascontiguousarray(c2_inputs[c2], dtype=dt) for lay, c2 in LAYER_FROM_CLOUDSC2.items()}
    res["pa"] = np.ascontiguousarray(pa, dtype=dt)
    ngpblks, klev, nproma = res["pq"].shape
    for name, kind, role, opt in LAYER_ARRAYS:
        if role == "o":
            res[name] = np.full(layer_shape(kind, nproma, klev, ngpblks), np.nan, dtype=dt)
    return res


def CLOUDST(params, params2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d, ptsphy, arrays, precision="sp",
            func=None):
    """One NPROMA block of CLOUDST through cloudst_{dp,sp}_c (or ``func`` with the same signature); ``arrays`` maps
    the CLOUDST_ARRAYS names to Fortran-ordered numpy arrays, outputs updated in place."""
    dt = _real(precision)
    ceta = np.ascontiguousarray(ceta, dtype=dt)
    args = [ctypes.byref(params), ctypes.byref(params2), ctypes.c_void_p(ceta.ctypes.data), ctypes.c_int(kidia),
            ctypes.c_int(kfdia), ctypes.c_int(klon), ctypes.c_int(ktdia), ctypes.c_int(klev),
            ctypes.c_int(int(bool(ldrain1d))), _ctype_real(precision)(ptsphy)]
    for name, kind, role in CLOUDST_ARRAYS:
        a = arrays[name]
        want = (klon, klev + 1) if kind == "h" else (klon, klev)
        if a.dtype != dt or not a.flags.f_contiguous or a.shape != want:
            raise TypeError(f"{name}: need a Fortran-contiguous {np.dtype(dt).name} array of shape {want}")
        args.append(ctypes.c_void_p(a.ctypes.data))
    if func is None:
        fn = f"cloudst_{precision}_c"
        C._check(fn, getattr(C.lib(), fn)(*args))
    else:
        rc = func(*args)
        if rc != 0:
            raise RuntimeError(f"CLOUDST returned {rc}")
    return arrays


def CLOUD_S_LAYER(params, params2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d, tsphy, arrays, precision="sp"):
    """One NPROMA block of CLOUD_S_LAYER through cloud_s_layer_{dp,sp}_c: ``arrays`` maps LAYER_ARRAYS names to
    Fortran-ordered numpy arrays ((KLON,KLEV[+1]), (KLON) for PPRECTYPE/PFZRA); optional ones may be absent."""
    dt = _real(precision)
    ceta = np.ascontiguousarray(ceta, dtype=dt)
    args = [ctypes.byref(params), ctypes.byref(params2), ctypes.c_void_p(ceta.ctypes.data), ctypes.c_int(kidia),
            ctypes.c_int(kfdia), ctypes.c_int(klon), ctypes.c_int(ktdia), ctypes.c_int(klev),
            ctypes.c_int(int(bool(ldrain1d))), _ctype_real(precision)(tsphy)]
    for name, kind, role, opt in LAYER_ARRAYS:
        a = arrays.get(name)
        if a is None:
            if not opt:
                raise KeyError(name)
            args.append(ctypes.c_void_p(0))
            continue
        want = (klon,) if kind == "1d" else ((klon, klev + 1) if kind == "h" else (klon, klev))
        if a.dtype != dt or not a.flags.f_contiguous or a.shape != want:
            raise TypeError(f"{name}: need a Fortran-contiguous {np.dtype(dt).name} array of shape {want}")
        args.append(ctypes.c_void_p(a.ctypes.data))
    fn = f"cloud_s_layer_{precision}_c"
    C._check(fn, getattr(C.lib(), fn)(*args))
    return arrays
