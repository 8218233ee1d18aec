"""Python binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (cloudsc_oracle.hpp) restates the reference CLOUDSC default path
on the CPU.  It is the checker of the HIP kernels: only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py import this
module.  The product (openifs_amd, libcloudsc_hip.so) never does.

Parity status of each oracle function (what pins it; see tests/test_oracle_*.py):
  Thermo statement functions     pinned: closed forms and Clausius-Clapeyron derivative checks
  Lanczos Gamma / PSD constants  pinned: math.gamma, hand-derived closed forms of SUCLDP
  CLOUD_SUPERSATCHECK            pinned: invariants (q+l+i conserved, enthalpy, exact
                                 saturation after adjustment to first order), no-op cases
  lu5_solve                      pinned: numpy.linalg.solve on diagonally dominant systems
  truncate_sinks                 pinned: brute-force positivity and ordering properties
  cloudsc (whole column)         pinned by column water/enthalpy conservation, the reference's
                                 write set and special cases (dry, clear-sky, warm, saturated)
  section-4 processes            pinned by single-process columns against the published rate laws with the
                                 SUCLDP constants, size-distribution integrals by quadrature and the reference's
                                 time discretisation (tests/test_oracle_processes.py, test_oracle_processes2.py),
                                 each checked to fail against a planted mistake: snow sedimentation, KK2000
                                 autoconversion and accretion, homogeneous freezing, turbulent erosion (rate and
                                 saturation limit), ice deposition (Meyers nuclei, Rotstayn growth), snow deposition,
                                 ice->snow autoconversion, melting, rain freezing (Wisner, and the heat-limited
                                 refreezing), riming, PSACR, PIACR, rain evaporation, snow sublimation
  budget diagnostics (LCLDBUD*)  pinned (tests/test_budget.py): liquid, ice and cloud-cover closure, the
                                 temperature budget against the liquid/ice fields, vertical integrals against the
                                 integrated 3D fields, counter closed forms, process regimes, TIMEINT/LBUD23/
                                 accumulation; the rain-only rates are pinned only through those identities
  condensate FSD (PFSD)          ice FSD pinned (tests/test_fsd.py): structure recovered from two PGAW values
                                 (detrainment ratio, layer-depth/total-water factor, overcast/broken branch),
                                 defaults and bounds; liquid FSD pinned by the structure of the Boutle et al. (2013)
                                 form (power laws in the grid length, linear in cover, overcast ratio;
                                 tests/test_oracle_processes2.py)
  LCLOUD_INHOMOG factors         pinned (tests/test_inhomog.py): autoconversion and accretion enhancement equal the
                                 lognormal and bivariate-lognormal moment ratios by Gauss-Hermite quadrature in an
                                 overcast warm layer; no change without liquid
  CLOUDSC2 (cloudsc2_oracle.hpp) pinned (tests/test_cloudsc2.py): per-level water closure of the tendencies,
                                 the precipitation fluxes and the detrainment, the enthalpy fluxes, the dry column,
                                 no supersaturation left after CUADJTQ/CUADJTQS, cloud cover 0/1 above the
                                 tropopause and monotone in total water, melting of snow, the write set and KTDIA;
                                 the autoconversion and evaporation rates only through the water closure
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_cloudsc.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def cloudsc_block(precision, params, kidia, kfdia, klon, klev, ptsphy, arrays, kfldx=0):
    """One CLOUDSC block on the CPU, same argument list as the product's cloudsc_*_c."""
    from openifs_amd.cloudsc import call_block

    f = getattr(lib(), f"oracle_cloudsc_{precision}")
    rc = call_block(f, params, kidia, kfdia, klon, klev, ptsphy, arrays, precision, kfldx=kfldx)
    if rc != 0:
        raise RuntimeError(f"oracle_cloudsc_{precision} returned {rc}")
    return arrays


def last_warnings():
    """(negative-liquid, negative-ice) warning counts of the calling thread's last cloudsc_block call."""
    nl, ni = ctypes.c_int64(), ctypes.c_int64()
    lib().oracle_last_warnings(ctypes.byref(nl), ctypes.byref(ni))
    return nl.value, ni.value


def cloudsc_blocked(precision, params, ngptot, nproma, klev, ptsphy, arrays, nthreads=0):
    """All blocks of a blocked field set (dict of C-order numpy arrays, see openifs_amd.cloudsc.blocked_shape)."""
    from openifs_amd.cloudsc import BLOCKED_ARRAYS

    ngpblks = (ngptot + nproma - 1) // nproma
    real = ctypes.c_double if precision == "dp" else ctypes.c_float
    args = [ctypes.byref(params), ctypes.c_int(ngptot), ctypes.c_int(nproma), ctypes.c_int(klev),
            ctypes.c_int(ngpblks), real(ptsphy)]
    for name, kind, role in BLOCKED_ARRAYS:
        a = arrays[name]
        want = np.int32 if kind == "1di" else (np.float64 if precision == "dp" else np.float32)
        if a.dtype != want or not a.flags.c_contiguous:
            raise TypeError(f"{name}: need a C-contiguous {np.dtype(want).name} array")
        args.append(ctypes.c_void_p(a.ctypes.data))
    args.append(ctypes.c_int(nthreads))
    rc = getattr(lib(), f"oracle_cloudsc_blocked_{precision}")(*args)
    if rc != 0:
        raise RuntimeError(f"oracle_cloudsc_blocked_{precision} returned {rc}")
    return arrays


def cloudsc_blocked_diag(precision, params, ngptot, nproma, klev, ptsphy, arrays, nthreads=0):
    """cloudsc_blocked plus the diagnostics: ``arrays`` may also hold the BUDGET_ARRAYS of openifs_amd.cloudsc and
    ``pextra`` as (NGPBLKS, KFLDX, KLEV, NPROMA), and ``pgaw`` (NGPBLKS, NPROMA) and ``pfsd`` (NGPBLKS, KLEV,
    NPROMA); missing ones are passed as NULL.  Returns the return code."""
    from openifs_amd.cloudsc import BLOCKED_ARRAYS, BUDGET_ARRAYS

    ngpblks = (ngptot + nproma - 1) // nproma
    real = ctypes.c_double if precision == "dp" else ctypes.c_float
    rdt = np.float64 if precision == "dp" else np.float32
    args = [ctypes.byref(params), ctypes.c_int(ngptot), ctypes.c_int(nproma), ctypes.c_int(klev),
            ctypes.c_int(ngpblks), real(ptsphy)]

    def ptr(name, kind):
        a = arrays.get(name)
        if a is None:
            return ctypes.c_void_p(0)
        want = np.int32 if kind == "1di" else rdt
        if a.dtype != want or not a.flags.c_contiguous:
            raise TypeError(f"{name}: need a C-contiguous {np.dtype(want).name} array")
        return ctypes.c_void_p(a.ctypes.data)

    for name, kind, _ in BLOCKED_ARRAYS:
        if name not in arrays:
            raise KeyError(name)
        args.append(ptr(name, kind))
    args += [ptr(n, k) for n, k in BUDGET_ARRAYS]
    px = arrays.get("pextra")
    kfldx = 0
    if px is not None:
        if px.shape != (ngpblks, px.shape[1], klev, nproma):
            raise ValueError("pextra: need (NGPBLKS, KFLDX, KLEV, NPROMA)")
        kfldx = px.shape[1]
    args += [ptr("pextra", "extra"), ctypes.c_int(kfldx), ptr("pgaw", "1d"), ptr("pfsd", "2d"),
             ctypes.c_int(nthreads)]
    return getattr(lib(), f"oracle_cloudsc_blocked_diag_{precision}")(*args)


def supersatcheck(params, t, q, a, pp, psurf, precision="dp"):
    dt = np.float64 if precision == "dp" else np.float32
    t, q, a, pp, psurf = (np.ascontiguousarray(x, dtype=dt) for x in (t, q, a, pp, psurf))
    n = t.size
    outs = [np.zeros(n, dtype=dt) for _ in range(5)]
    f = getattr(lib(), f"oracle_supersatcheck_{precision}")
    f(ctypes.byref(params), ctypes.c_int(n), *(ctypes.c_void_p(x.ctypes.data) for x in (t, q, a, pp, psurf)),
      *(ctypes.c_void_p(x.ctypes.data) for x in outs))
    return outs  # t_adj, q_adj, a_adj, l_adj, i_adj


def lu5(lhs, rhs):
    lhs = np.ascontiguousarray(lhs, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.zeros(5)
    lib().oracle_lu5_dp(ctypes.c_void_p(lhs.ctypes.data), ctypes.c_void_p(rhs.ctypes.data),
                        ctypes.c_void_p(x.ctypes.data))
    return x


def truncate_sinks(qa, qx):
    qa = np.ascontiguousarray(qa, dtype=np.float64).copy()
    qx = np.ascontiguousarray(qx, dtype=np.float64)
    lib().oracle_truncate_sinks_dp(ctypes.c_void_p(qa.ctypes.data), ctypes.c_void_p(qx.ctypes.data))
    return qa


THERMO = {"foedelta": 0, "foeew": 1, "foealfa": 2, "foeewm": 3, "foedem": 4, "foeldcpm": 5, "foeeliq": 6,
          "foeeice": 7, "foeeliq2ice": 8, "fokoop": 9}


def thermo(params, name, t):
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = np.zeros_like(t)
    rc = lib().oracle_thermo_dp(ctypes.byref(params), ctypes.c_int(THERMO[name]), ctypes.c_int(t.size),
                                ctypes.c_void_p(t.ctypes.data), ctypes.c_void_p(out.ctypes.data))
    if rc != 0:
        raise RuntimeError(name)
    return out


def satadj_block(precision, params, kidia, kfdia, klon, klev, ptsphy, arrays):
    """One CLOUD_SATADJ block on the CPU, same argument list as the product's cloudsc_satadj_*_c."""
    from openifs_amd.satadj import call_block

    rc = call_block(getattr(lib(), f"oracle_satadj_{precision}"), params, kidia, kfdia, klon, klev, ptsphy, arrays,
                    precision)
    if rc != 0:
        raise RuntimeError(f"oracle_satadj_{precision} returned {rc}")
    return arrays


def satadj_blocked(precision, params, ngptot, nproma, klev, ptsphy, arrays, nthreads=0):
    """All blocks of a CLOUD_SATADJ blocked field set (dict of C-order numpy arrays, openifs_amd.satadj)."""
    from openifs_amd.satadj import SATADJ_BLOCKED

    ngpblks = (ngptot + nproma - 1) // nproma
    dt = np.float64 if precision == "dp" else np.float32
    real = ctypes.c_double if precision == "dp" else ctypes.c_float
    args = [ctypes.byref(params), ctypes.c_int(ngptot), ctypes.c_int(nproma), ctypes.c_int(klev),
            ctypes.c_int(ngpblks), real(ptsphy)]
    for name, kind, role in SATADJ_BLOCKED:
        a = arrays[name]
        if a.dtype != dt or not a.flags.c_contiguous:
            raise TypeError(f"{name}: need a C-contiguous {np.dtype(dt).name} array")
        args.append(ctypes.c_void_p(a.ctypes.data))
    args.append(ctypes.c_int(nthreads))
    rc = getattr(lib(), f"oracle_satadj_blocked_{precision}")(*args)
    if rc != 0:
        raise RuntimeError(f"oracle_satadj_blocked_{precision} returned {rc}")
    return arrays


def qsatmixadj(params, t, q, pp):
    """QSATMIXADJ (two Newton iterations) on arrays of points, FP64."""
    t, q, pp = (np.ascontiguousarray(x, dtype=np.float64) for x in (t, q, pp))
    out = np.zeros_like(t)
    lib().oracle_qsatmixadj_dp(ctypes.byref(params), ctypes.c_int(t.size), *(ctypes.c_void_p(x.ctypes.data)
                                                                             for x in (t, q, pp, out)))
    return out


def supersatcheck_k(params, kftliqice, t, q, a, pp, psurf):
    """CLOUD_SUPERSATCHECK in FP64 with an explicit KFTLIQICE; returns (t, q, a, l, i) adjustments."""
    t, q, a, pp, psurf = (np.ascontiguousarray(x, dtype=np.float64) for x in (t, q, a, pp, psurf))
    outs = [np.zeros(t.size) for _ in range(5)]
    lib().oracle_supersatcheck_k_dp(ctypes.byref(params), ctypes.c_int(t.size), ctypes.c_int(kftliqice),
                                    *(ctypes.c_void_p(x.ctypes.data) for x in (t, q, a, pp, psurf)),
                                    *(ctypes.c_void_p(x.ctypes.data) for x in outs))
    return outs


def sltend_supersat_blocked(precision, params, ngptot, nproma, klev, tsphy, arrays, nthreads=0):
    """SLTEND section 3 on blocked C-order numpy arrays (names of openifs_amd.supersat.SLTEND_ARRAYS; padj*
    optional), updated in place."""
    from openifs_amd.supersat import SLTEND_ARRAYS

    ngpblks = (ngptot + nproma - 1) // nproma
    dt = np.float64 if precision == "dp" else np.float32
    real = ctypes.c_double if precision == "dp" else ctypes.c_float
    args = [ctypes.byref(params), ctypes.c_int(ngptot), ctypes.c_int(nproma), ctypes.c_int(klev),
            ctypes.c_int(ngpblks), real(tsphy)]
    for n in SLTEND_ARRAYS:
        a = arrays.get(n)
        if a is None:
            args.append(ctypes.c_void_p(0))
            continue
        if a.dtype != dt or not a.flags.c_contiguous:
            raise TypeError(f"{n}: need a C-contiguous {np.dtype(dt).name} array")
        args.append(ctypes.c_void_p(a.ctypes.data))
    args.append(ctypes.c_int(nthreads))
    rc = getattr(lib(), f"oracle_sltend_supersat_blocked_{precision}")(*args)
    if rc != 0:
        raise RuntimeError(f"oracle_sltend_supersat_blocked_{precision} returned {rc}")
    return arrays


def cloudsc2_block(precision, params, params2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d, ptsphy, arrays):
    """One CLOUDSC2 block on the CPU, same argument list as the product's cloudsc2_*_c (Fortran-ordered arrays)."""
    from openifs_amd.cloudsc2 import block_args

    ceta = np.ascontiguousarray(ceta, dtype=np.float64 if precision == "dp" else np.float32)
    args = block_args(params, params2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d, ptsphy, arrays, precision)
    rc = getattr(lib(), f"oracle_cloudsc2_{precision}")(*args)
    if rc != 0:
        raise RuntimeError(f"oracle_cloudsc2_{precision} returned {rc}")
    return arrays


def cloudsc2_blocked(precision, params, params2, ceta, ngptot, nproma, ktdia, klev, ldrain1d, ptsphy, arrays,
                     nthreads=0):
    """CLOUDSC2 over all blocks of blocked host arrays (C-order (NGPBLKS, KLEV[+1], NPROMA), the argument names of
    openifs_amd.cloudsc2.CLOUDSC2_ARRAYS); outputs written in place."""
    from openifs_amd.cloudsc2 import CLOUDSC2_ARRAYS

    ngpblks = (ngptot + nproma - 1) // nproma
    real = ctypes.c_double if precision == "dp" else ctypes.c_float
    rdt = np.float64 if precision == "dp" else np.float32
    ceta = np.ascontiguousarray(ceta, dtype=rdt)
    args = [ctypes.byref(params), ctypes.byref(params2), ctypes.c_void_p(ceta.ctypes.data), ctypes.c_int(ngptot),
            ctypes.c_int(nproma), ctypes.c_int(ktdia), ctypes.c_int(klev), ctypes.c_int(ngpblks),
            ctypes.c_int(int(bool(ldrain1d))), real(ptsphy)]
    for name, kind, _ in CLOUDSC2_ARRAYS:
        a = arrays[name]
        want = (ngpblks, klev + 1 if kind == "h" else klev, nproma)
        if a.dtype != rdt or not a.flags.c_contiguous or a.shape != want:
            raise TypeError(f"{name}: need a C-contiguous {np.dtype(rdt).name} array of shape {want}")
        args.append(ctypes.c_void_p(a.ctypes.data))
    args.append(ctypes.c_int(nthreads))
    rc = getattr(lib(), f"oracle_cloudsc2_blocked_{precision}")(*args)
    if rc != 0:
        raise RuntimeError(f"oracle_cloudsc2_blocked_{precision} returned {rc}")
    return arrays


def cloud_s_layer_blocked(precision, params, params2, ceta, ngptot, nproma, ktdia, klev, ldrain1d, tsphy, arrays,
                          nthreads=0):
    """CLOUD_S_LAYER (LEPCLD2) over all blocks of blocked host arrays named as openifs_amd.cloudsc2.LAYER_ARRAYS;
    optional arrays missing from ``arrays`` are passed as NULL."""
    from openifs_amd.cloudsc2 import LAYER_ARRAYS, layer_shape

    ngpblks = (ngptot + nproma - 1) // nproma
    real = ctypes.c_double if precision == "dp" else ctypes.c_float
    rdt = np.float64 if precision == "dp" else np.float32
    ceta = np.ascontiguousarray(ceta, dtype=rdt)
    args = [ctypes.byref(params), ctypes.byref(params2), ctypes.c_void_p(ceta.ctypes.data), ctypes.c_int(ngptot),
            ctypes.c_int(nproma), ctypes.c_int(ktdia), ctypes.c_int(klev), ctypes.c_int(ngpblks),
            ctypes.c_int(int(bool(ldrain1d))), real(tsphy)]
    for name, kind, _, opt in LAYER_ARRAYS:
        a = arrays.get(name)
        if a is None:
            if not opt:
                raise KeyError(name)
            args.append(ctypes.c_void_p(0))
            continue
        want = layer_shape(kind, nproma, klev, ngpblks)
        if a.dtype != rdt or not a.flags.c_contiguous or a.shape != want:
            raise TypeError(f"{name}: need a C-contiguous {np.dtype(rdt).name} array of shape {want}")
        args.append(ctypes.c_void_p(a.ctypes.data))
    args.append(ctypes.c_int(nthreads))
    rc = getattr(lib(), f"oracle_cloud_s_layer_blocked_{precision}")(*args)
    if rc != 0:
        raise RuntimeError(f"oracle_cloud_s_layer_blocked_{precision} returned {rc}")
    return arrays


def cloudst_blocked(precision, params, params2, ceta, ngptot, nproma, ktdia, klev, ldrain1d, ptsphy, arrays,
                    nthreads=0):
    """CLOUDST over all blocks of blocked host arrays named as openifs_amd.cloudsc2.CLOUDST_ARRAYS."""
    from openifs_amd.cloudsc2 import CLOUDST_ARRAYS

    ngpblks = (ngptot + nproma - 1) // nproma
    real = ctypes.c_double if precision == "dp" else ctypes.c_float
    rdt = np.float64 if precision == "dp" else np.float32
    ceta = np.ascontiguousarray(ceta, dtype=rdt)
    args = [ctypes.byref(params), ctypes.byref(params2), ctypes.c_void_p(ceta.ctypes.data), ctypes.c_int(ngptot),
            ctypes.c_int(nproma), ctypes.c_int(ktdia), ctypes.c_int(klev), ctypes.c_int(ngpblks),
            ctypes.c_int(int(bool(ldrain1d))), real(ptsphy)]
    for name, kind, _ in CLOUDST_ARRAYS:
        a = arrays[name]
        want = (ngpblks, klev + 1 if kind == "h" else klev, nproma)
        if a.dtype != rdt or not a.flags.c_contiguous or a.shape != want:
            raise TypeError(f"{name}: need a C-contiguous {np.dtype(rdt).name} array of shape {want}")
        args.append(ctypes.c_void_p(a.ctypes.data))
    args.append(ctypes.c_int(nthreads))
    rc = getattr(lib(), f"oracle_cloudst_blocked_{precision}")(*args)
    if rc != 0:
        raise RuntimeError(f"oracle_cloudst_blocked_{precision} returned {rc}")
    return arrays
