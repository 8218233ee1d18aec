"""CLOUD_SATADJ inputs for the tests.

blocked_case: the CLOUDSC synthetic atmosphere (host generator of libcloudsc_hip.so) plus numpy-drawn forcings
with the distributions of cloudsc_satadj_synth_*; block_arrays cuts one block out of it in Fortran order.
controlled_block: a (KLON, KLEV) block with everything zero except a uniform pressure grid, for the
closed-form cases.
"""
import numpy as np

from openifs_amd import cloudsc as C
from openifs_amd import satadj as S


def real_dtype(prec):
    return np.float64 if prec == "dp" else np.float32


def blocked_case(params, seed, ngptot, nproma, klev, prec="dp"):
    """Blocked C-order arrays of every SATADJ input; outputs NaN."""
    base = C.synth_host(params, seed, ngptot, nproma, klev, precision="dp")
    ngpblks = (ngptot + nproma - 1) // nproma
    rng = np.random.default_rng(seed)
    s2 = C.blocked_shape("2d", nproma, klev, ngpblks)
    arr = {n: base[n] for n in S.FROM_CLOUDSC_SYNTH}
    arr["phrsw"] = rng.uniform(0.0, 2e-5, s2)
    arr["phrlw"] = rng.normal(-2e-5, 1e-5, s2)
    arr["ptendency_vdf_t"] = rng.normal(0.0, 1e-5, s2)
    arr["ptendency_vdf_q"] = rng.normal(0.0, 1e-8, s2)
    arr["pvervel"] = rng.normal(0.0, 0.3, s2)
    lde = arr["pludeli"].copy()
    lde[:, 2] = 1e-2 * arr["plude"]
    lde[:, 3] = 1e3 * arr["plude"]
    arr["pludeli"] = lde
    dt = real_dtype(prec)
    out = {k: v.astype(dt) for k, v in arr.items()}
    for name, kind, role in S.SATADJ_BLOCKED:
        if role == "o":
            out[name] = np.full(C.blocked_shape(kind, nproma, klev, ngpblks), np.nan, dtype=dt)
    return out


def block_arrays(arr, b, prec="dp"):
    """Block ``b`` of a blocked case as the Fortran-ordered per-block argument set of CLOUD_SATADJ."""
    dt = real_dtype(prec)
    F = lambda x: np.array(x.T, dtype=dt, order="F")
    out = {n: F(arr[n][b]) for n in ("pap", "paph", "pt", "pq", "pa", "ptendency_cml_t", "ptendency_cml_q",
                                      "ptendency_cml_a", "phrsw", "phrlw", "ptendency_vdf_t", "ptendency_vdf_q",
                                      "pludeli", "pvervel")}
    out["pl"], out["pi"] = F(arr["pclv"][b, 0]), F(arr["pclv"][b, 1])
    out["ptendency_cml_l"], out["ptendency_cml_i"] = F(arr["ptendency_cml_cld"][b, 0]), F(arr["ptendency_cml_cld"][b, 1])
    for n in S.SATADJ_OUTPUTS:
        out[n] = np.full(out["pt"].shape, np.nan, dtype=dt, order="F")
    return out


def controlled_block(klon, klev, p_top=50000.0, p_surf=100000.0):
    """Zero inputs on a uniform pressure grid (half levels from p_top to p_surf)."""
    a = S.allocate_block(klon, klev, "dp", fill=np.nan)
    ph = p_top + (p_surf - p_top) * np.arange(klev + 1) / klev
    a["paph"][:] = ph[None, :]
    a["pap"][:] = 0.5 * (ph[:-1] + ph[1:])[None, :]
    return a


def qsat_mix(params, t, pp):
    """Saturation specific humidity over the mixed phase, from the pinned FOEEWM."""
    import oracle

    t = np.asarray(t, dtype=np.float64)
    es = np.minimum(oracle.thermo(params, "foeewm", t.ravel()).reshape(t.shape) / pp, 0.5)
    return es / (1.0 - params.retv * es)


def saturation_root(params, t0, q0, pp, iters=200):
    """Warm (liquid-only) isenthalpic saturation point: the q' with q' = qsat(t0 + Lv/cp (q0 - q'), pp), by
    bisection (exact to rounding, independent of any Newton step)."""
    t0, q0, pp = (np.asarray(x, dtype=np.float64) for x in (t0, q0, pp))
    lo = np.zeros_like(q0)
    hi = np.maximum(q0, qsat_mix(params, t0, pp)) * 2.0 + 1e-3
    f = lambda q: q - qsat_mix(params, t0 + params.ralvdcp * (q0 - q), pp)
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        neg = f(mid) < 0
        lo = np.where(neg, mid, lo)
        hi = np.where(neg, hi, mid)
    return 0.5 * (lo + hi)
