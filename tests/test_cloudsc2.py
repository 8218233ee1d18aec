"""CLOUDSC2 (cloudsc2.F90): the oracle pinned by conservation laws and special cases, the C-ABI checks, and the
gfx950 kernel against the oracle.

Oracle pins (CPU):
- water: on every level, (PTENQ + PTENL + PTENI) dp/g plus the divergence of the rain + snow flux equals the
  detrained condensate PLUDE, whatever the processes did (condensation, autoconversion, melting, evaporation, the
  final adjustment);
- PFHPSL = -RLVTT PFPLSL and PFHPSN = -RLSTT PFPLSN (cloudsc2.F90:694-699);
- a dry column (subsaturated, no condensate, no convection) gets exactly zero tendencies, cover and fluxes;
- no supersaturation is left after CUADJTQ / CUADJTQS at the final state T + dt (PGTENT + PTENT),
  q + dt (PGTENQ + PTENQ), although the first guess is up to 40 % supersaturated;
- cloud cover is 0 or 1 above the tropopause (critical humidity 1) and nondecreasing in total water, 0 below the
  lowest critical humidity and 1 at saturation;
- snow falling into a layer whose heat content above RTT + 2 exceeds the latent heat of the snow flux melts
  entirely;
- the write set: tendencies on KTDIA..KLEV, PCLC and PCOVPTOT on every level, the fluxes at the top and below
  KTDIA, the padding columns of the last block untouched.
The individual rates (autoconversion, evaporation) are pinned only through the water closure.

GPU parity: FP64 relative RMS <= 1e-12 per field with the same write set; FP32 by the FP32 oracle's own distance
to FP64 (GPU relative RMS against the FP64 oracle <= 2 x the FP32 oracle's + 1e-6).
"""
import math

import numpy as np
import pytest

import oracle
from openifs_amd import cloudsc as C
from openifs_amd import cloudsc2 as C2
from tests.helpers import relrms

PTSPHY = 900.0
# (LPHYLIN, LEVAPLS2, LDRAIN1D): the default path, precipitation evaporation, the linearised thermodynamics
# (tanh phase function and CUADJTQS) and the 1D rain switch, which turns on both
VARIANTS = [(0, 0, False), (0, 1, False), (1, 0, False), (0, 0, True)]
VARIANT_IDS = ["default", "levapls2", "lphylin", "ldrain1d"]


def params2(p, lphylin=0, levapls2=0):
    q = C2.default_params2(p)
    q.lphylin, q.levapls2 = lphylin, levapls2
    return q


def run_oracle(prec, p, q2, arr, ngptot, nproma, klev, ktdia=1, ldrain1d=False):
    dt = np.float64 if prec == "dp" else np.float32
    a = {k: v.astype(dt) for k, v in arr.items()}
    oracle.cloudsc2_blocked(prec, p, q2, C2.reference_eta(klev), ngptot, nproma, ktdia, klev, ldrain1d, PTSPHY, a,
                            nthreads=8)
    return a


def water_residual(p, a, ktdia=1):
    """Per-level water closure residual over its scale (levels KTDIA..KLEV)."""
    k0 = ktdia - 1
    dp = np.diff(a["paphp1"], axis=1)[:, k0:]
    tq, tl, ti = (a[n][:, k0:].astype(np.float64) for n in ("ptenq", "ptenl", "pteni"))
    f = (a["pfplsl"] + a["pfplsn"]).astype(np.float64)
    fdn, fup = f[:, k0 + 1:], f[:, k0:-1].copy()
    fup[:, 0] = 0.0  # the sweep starts with no precipitation at KTDIA
    lude = a["plude"][:, k0:]
    res = (tq + tl + ti) * dp / p.rg + fdn - fup - lude
    scale = (abs(tq) + abs(tl) + abs(ti)) * dp / p.rg + abs(fdn) + abs(fup) + abs(lude)
    return abs(res) / (scale + 1e-300)


def final_rh(p, a, linear):
    """q / qsat at the final state and the points it is checked on.  The default path adjusts to the mixed-phase
    saturation (FOEEWM): all points.  CUADJTQS adjusts to water (above RTT) or ice saturation chosen from the
    temperature before the adjustment, which the outputs do not hold, and the tendency then books the condensate
    with the latent heat of the phase after it: checked against the saturation of the final temperature's phase
    where the first-guess and the final temperature are on the same side of RTT, 5 K or more from it."""
    t1 = a["ptm1"] + PTSPHY * a["pgtent"]
    t = t1 + PTSPHY * a["ptent"]
    q = a["pqm1"] + PTSPHY * (a["pgtenq"] + a["ptenq"])
    if linear:
        liq = t > p.rtt
        r3, r4 = np.where(liq, p.r3les, p.r3ies), np.where(liq, p.r4les, p.r4ies)
        qs = np.minimum(0.5, p.r2es * np.exp(r3 * (t - p.rtt) / (t - r4)) / a["papp1"])
        qs = qs / (1.0 - p.retv * qs)
        m = (abs(t - p.rtt) > 5.0) & (abs(t1 - p.rtt) > 5.0) & ((t > p.rtt) == (t1 > p.rtt))
    else:
        qs = C2._qsat(p, t, a["papp1"])
        m = np.ones(t.shape, bool)
    return (q / qs)[m]


@pytest.fixture(scope="module")
def case137(params137):
    return C2.synth_host(params137, 7, 512, 32, 137)


# ----------------------------------------------------------------------------------------------------------------
# oracle pins
# ----------------------------------------------------------------------------------------------------------------
def test_params2_defaults(params137):
    """sucldp.F90:264, suphli.F90:107,125-127, suphec.F90:229, su0phy.F90:1361."""
    q = C2.default_params2(params137)
    assert q.rclcrit == 4.0e-4
    assert q.rlptrc == pytest.approx(params137.rtice + (params137.rtwat - params137.rtice) / math.sqrt(2.0), abs=0)
    assert (q.rlpal1, q.rlpal2, q.rvtmp2, q.lphylin, q.levapls2) == (0.15, 20.0, 0.0, 0, 0)
    assert q.rlptrc == pytest.approx(266.4234559672906, rel=1e-15)


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("ktdia", [1, 12])
def test_oracle_water_closure(variant, ktdia, params137, case137):
    lphylin, levap, ldr = variant
    a = run_oracle("dp", params137, params2(params137, lphylin, levap), case137, 512, 32, 137, ktdia, ldr)
    r = water_residual(params137, a, ktdia)
    assert r.max() < 1e-12
    # the closure is not vacuous: precipitation reaches the surface and the detrainment is active
    assert (a["pfplsl"][:, -1] + a["pfplsn"][:, -1] > 1e-6).mean() > 0.3
    assert (a["plude"] > 0).any()
    # PFHPSL / PFHPSN on every half level from the precipitation fluxes
    np.testing.assert_array_equal(a["pfhpsl"], -a["pfplsl"] * params137.rlvtt)
    np.testing.assert_array_equal(a["pfhpsn"], -a["pfplsn"] * params137.rlstt)


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_oracle_no_supersaturation_left(variant, params137, case137):
    lphylin, levap, ldr = variant
    p = params137
    q1 = case137["pqm1"] + PTSPHY * case137["pgtenq"]
    rh1 = q1 / C2._qsat(p, case137["ptm1"] + PTSPHY * case137["pgtent"], case137["papp1"])
    assert rh1.max() > 1.3  # the first guess is supersaturated
    a = run_oracle("dp", p, params2(p, lphylin, levap), case137, 512, 32, 137, 1, ldr)
    rh = final_rh(p, a, lphylin or ldr)
    # two Newton steps leave a residual of second order in the initial supersaturation
    assert rh.max() < 1.005
    assert (rh > 0.999).sum() > 1000  # the adjustment acted


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_oracle_dry_column(variant, params137, case137):
    lphylin, levap, ldr = variant
    a = {k: v.copy() for k, v in case137.items()}
    for n in ("pl", "pi", "pgtenl", "pgteni", "pgtenq", "plude", "plu", "pmfu", "pmfd"):
        a[n][...] = 0.0
    a["pqm1"] = 0.05 * a["pqs"]
    a = run_oracle("dp", params137, params2(params137, lphylin, levap), a, 512, 32, 137, 1, ldr)
    for n in C2.OUTPUT_NAMES:
        assert np.all(a[n] == 0.0), n


def test_oracle_cloud_cover_sweep(params137, case137):
    """One level, 64 identical columns but for the total water, swept from 0 to 1.3 x saturation."""
    p, k, n = params137, 100, 64
    a = {key: np.ascontiguousarray(np.broadcast_to(v[:1, :, :1], (1, v.shape[1], n))) for key, v in case137.items()}
    for key in ("pl", "pi", "pgtenl", "pgteni", "pgtenq", "plude"):
        a[key][0, k, :] = 0.0
    qs = a["pqs"][0, k, 0]
    t1 = a["ptm1"][0, k, 0] + PTSPHY * a["pgtent"][0, k, 0]
    assert t1 > p.rtice  # no ice supersaturation factor at this level
    qt = np.sort(np.r_[np.linspace(0.0, 1.3 * qs, n - 1), qs * (1.0 - 1e-9)])
    a["pqm1"][0, k, :] = qt
    a = run_oracle("dp", p, params2(p), a, n, n, 137)
    clc = a["pclc"][0, k]
    assert np.all(np.diff(clc) >= 0.0)
    assert np.all(clc[qt <= 0.3 * qs] == 0.0)  # the critical humidity is at least 0.31
    assert np.all(clc[qt >= qs] == 1.0)
    partial = (clc > 0.0) & (clc < 1.0)
    i0 = int(np.argmax(partial))
    assert partial.sum() >= 5 and np.all(partial[i0:] | (qt[i0:] >= qs))
    # continuous at the critical humidity and at saturation
    assert clc[i0] < 0.05 and clc[np.argmin(abs(qt - qs * (1.0 - 1e-9)))] > 0.999


def test_oracle_cloud_cover_above_tropopause(params137, case137):
    """Above eta 0.1 (always above the tropopause) the critical humidity is 1: cover is 0 or 1."""
    a = run_oracle("dp", params137, params2(params137), case137, 512, 32, 137)
    top = C2.reference_eta(137) < 0.1
    clc = a["pclc"][:, top]
    assert np.all((clc == 0.0) | (clc == 1.0))
    assert (clc == 1.0).any() and (clc == 0.0).any()


def test_oracle_snow_melts_in_warm_layers(params137, case137):
    p = params137
    a = run_oracle("dp", p, params2(p), case137, 512, 32, 137)
    t1 = a["ptm1"] + PTSPHY * a["pgtent"]
    dp = np.diff(a["paphp1"], axis=1)
    heat = dp / (p.rg * PTSPHY) * (p.rcpd / p.rlmlt) * (t1 - (p.rtt + 2.0))  # kg m-2 s-1 of snow it can melt
    fin = a["pfplsn"][:, :-1]
    sel = (t1 > p.rtt + 2.0) & (fin > 0.0) & (heat > fin)
    assert sel.sum() > 100
    assert np.all(a["pfplsn"][:, 1:][sel] == 0.0)
    assert (a["pfplsl"][:, 1:][sel] >= fin[sel]).all()


def test_oracle_write_set_and_ktdia(params137, case137):
    ktdia, ngptot, nproma = 12, 500, 32  # 500 columns: the last block has 20 of 32
    a = run_oracle("dp", params137, params2(params137), {k: v[:16] for k, v in case137.items()}, ngptot, nproma,
                   137, ktdia)
    k0 = ktdia - 1
    real = np.zeros((16, nproma), bool)
    real.reshape(-1)[:ngptot] = True
    for n in ("ptent", "ptenq", "ptenl", "pteni"):
        assert np.all(np.isnan(a[n][:, :k0])), n
        assert np.all(np.isfinite(a[n][:, k0:].transpose(0, 2, 1)[real])), n
    for n in ("pclc", "pcovptot"):
        assert np.all(a[n].transpose(0, 2, 1)[real][:, :k0] == 0.0), n
    for n in ("pfplsl", "pfplsn"):
        x = a[n].transpose(0, 2, 1)[real]
        assert np.all(x[:, 0] == 0.0) and np.all(np.isnan(x[:, 1:ktdia])) and np.all(np.isfinite(x[:, ktdia:])), n
    for n in C2.OUTPUT_NAMES:
        assert np.all(np.isnan(a[n].transpose(0, 2, 1)[~real])), n


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_oracle_precipitation_fraction(variant, params137, case137):
    """PCOVPTOT is written only with precipitation evaporation (LEVAPLS2 or LDRAIN1D), and then lies between the
    level's cloud cover and 1."""
    lphylin, levap, ldr = variant
    a = run_oracle("dp", params137, params2(params137, lphylin, levap), case137, 512, 32, 137, 1, ldr)
    cov = a["pcovptot"]
    if not (levap or ldr):
        assert np.all(cov == 0.0)
    else:
        on = cov > 0.0
        assert on.mean() > 0.2
        assert np.all(cov[on] >= a["pclc"][on]) and np.all(cov <= 1.0)


def test_oracle_fp32_tracks_fp64(params137, case137):
    p = params137
    a64 = run_oracle("dp", p, params2(p), case137, 512, 32, 137)
    a32 = run_oracle("sp", p, params2(p), case137, 512, 32, 137)
    for n in ("ptent", "ptenq", "pclc", "pfplsl", "pfplsn"):
        assert relrms(a32[n], a64[n]) < 1e-2, n
    assert water_residual(p, a32).mean() < 1e-5


def test_entry_checks(params137):
    """Bad dimensions and NULL arrays are rejected before any device work (no GPU needed)."""
    klon, klev = 8, 137
    arr = {}
    for name, kind, _ in C2.CLOUDSC2_ARRAYS:
        arr[name] = np.zeros((klon, klev + 1) if kind == "h" else (klon, klev), order="F")
    ceta = C2.reference_eta(klev)
    q2 = params2(params137)
    for kidia, kfdia, ktdia in [(0, 8, 1), (3, 2, 1), (1, 9, 1), (1, 8, 0), (1, 8, 138)]:
        with pytest.raises(C.CloudscError) as e:
            C2.CLOUDSC2(params137, q2, ceta, kidia, kfdia, klon, ktdia, klev, False, PTSPHY, arr, "dp")
        assert e.value.code == -1  # CLOUDSC_EDIMS
    with pytest.raises(TypeError):
        C2.CLOUDSC2(params137, q2, ceta, 1, 8, klon, 1, klev, False, PTSPHY, dict(arr, ptent=np.zeros((klon, klev))),
                    "dp")


# ----------------------------------------------------------------------------------------------------------------
# GPU parity
# ----------------------------------------------------------------------------------------------------------------
def gpu_run(p, q2, host, ngptot, nproma, klev, prec, ktdia=1, ldrain1d=False):
    import torch

    dt = torch.float64 if prec == "dp" else torch.float32
    f = {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda", dt) for k, v in host.items()}
    ceta = torch.from_numpy(C2.reference_eta(klev)).to("cuda", dt)
    C2.cloudsc2_blocked(f, p, q2, ceta, ngptot, nproma, klev, PTSPHY, prec, ktdia, ldrain1d)
    torch.cuda.synchronize()
    return {k: f[k].double().cpu().numpy() for k in C2.OUTPUT_NAMES}


@pytest.mark.gpu
@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("ngptot,nproma,klev,ktdia,seed", [(2048, 64, 137, 1, 1), (777, 32, 91, 5, 2)])
def test_gpu_fp64_matches_oracle(variant, ngptot, nproma, klev, ktdia, seed, params137, params91):
    p = params137 if klev == 137 else params91
    lphylin, levap, ldr = variant
    q2 = params2(p, lphylin, levap)
    host = C2.synth_host(p, seed, ngptot, nproma, klev, "dp")
    print(f"cloudsc2 dp {VARIANT_IDS[VARIANTS.index(variant)]} {ngptot}x{klev} ktdia {ktdia}", flush=True)
    out = gpu_run(p, q2, host, ngptot, nproma, klev, "dp", ktdia, ldr)
    ref = run_oracle("dp", p, q2, host, ngptot, nproma, klev, ktdia, ldr)
    for n in C2.OUTPUT_NAMES:
        np.testing.assert_array_equal(np.isnan(out[n]), np.isnan(ref[n]), err_msg=n)
        assert relrms(out[n], ref[n]) <= 1e-12, n


@pytest.mark.gpu
@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_gpu_fp32_as_accurate_as_fp32_oracle(variant, params137):
    p = params137
    lphylin, levap, ldr = variant
    q2 = params2(p, lphylin, levap)
    ngptot, nproma, klev = 4096, 128, 137
    host = C2.synth_host(p, 11, ngptot, nproma, klev, "dp")
    print(f"cloudsc2 sp {VARIANT_IDS[VARIANTS.index(variant)]} {ngptot}x{klev}", flush=True)
    host32 = {k: v.astype(np.float32) for k, v in host.items()}
    out = gpu_run(p, q2, host32, ngptot, nproma, klev, "sp", 1, ldr)
    r64 = run_oracle("dp", p, q2, host32, ngptot, nproma, klev, 1, ldr)
    r32 = run_oracle("sp", p, q2, host32, ngptot, nproma, klev, 1, ldr)
    report = []
    for n in C2.OUTPUT_NAMES:
        np.testing.assert_array_equal(np.isnan(out[n]), np.isnan(r64[n]), err_msg=n)
        e_gpu, e_cpu = relrms(out[n], r64[n]), relrms(r32[n], r64[n])
        report.append((n, e_gpu, e_cpu))
        assert e_gpu <= 2.0 * e_cpu + 1e-6, f"{n}: GPU {e_gpu:.3e} vs FP32 oracle {e_cpu:.3e}"
    print("\n".join(f"{n:10s} gpu {a:.3e}  fp32-oracle {b:.3e}" for n, a, b in report))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["dp", "sp"])
def test_gpu_per_block_drop_in(prec, params137):
    """cloudsc2_*_c (host arrays, one block, KIDIA:KFDIA window) equals the oracle of the same precision and leaves
    the columns outside the window untouched."""
    p, klon, klev = params137, 48, 137
    q2 = params2(p, 0, 1)
    host = C2.synth_host(p, 9, klon, klon, klev, prec)
    dt = np.float64 if prec == "dp" else np.float32

    def fortran(h):
        return {k: np.asfortranarray(v[0].T.astype(dt)) for k, v in h.items()}

    a_gpu, a_ref = fortran(host), fortran(host)
    ceta = C2.reference_eta(klev)
    print(f"cloudsc2 per-block {prec}", flush=True)
    C2.CLOUDSC2(p, q2, ceta, 3, 45, klon, 2, klev, False, PTSPHY, a_gpu, prec)
    oracle.cloudsc2_block(prec, p, q2, ceta, 3, 45, klon, 2, klev, False, PTSPHY, a_ref)
    for n in C2.OUTPUT_NAMES:
        g, r = a_gpu[n], a_ref[n]
        np.testing.assert_array_equal(np.isnan(g), np.isnan(r), err_msg=n)
        assert np.all(np.isnan(g[:2])) and np.all(np.isnan(g[45:])), n
        assert relrms(g, r) <= (1e-12 if prec == "dp" else 1e-4), n


@pytest.mark.gpu
def test_gpu_full_size_properties(params137):
    """655,360 columns at L137 in FP32: finite outputs, the per-level water closure to FP32 rounding, PFHPSL/PFHPSN
    from the fluxes, and a sample of blocks equal to the oracle by the FP32 criterion."""
    import torch

    p, ngptot, nproma, klev = params137, 655360, 128, 137
    q2 = params2(p)
    print("cloudsc2 sp 655360x137", flush=True)
    f = C2.synth_device(p, 21, ngptot, nproma, klev, "sp")
    ceta = torch.from_numpy(C2.reference_eta(klev)).to("cuda", torch.float32)
    C2.cloudsc2_blocked(f, p, q2, ceta, ngptot, nproma, klev, PTSPHY, "sp")
    torch.cuda.synchronize()
    for n in C2.OUTPUT_NAMES:
        assert torch.isfinite(f[n]).all(), n
    assert torch.equal(f["pfhpsl"], -f["pfplsl"] * np.float32(p.rlvtt))
    dp = (f["paphp1"][:, 1:] - f["paphp1"][:, :-1]).double()
    tw = (f["ptenq"].double() + f["ptenl"].double() + f["pteni"].double()) * dp / p.rg
    fl = (f["pfplsl"] + f["pfplsn"]).double()
    res = (tw + fl[:, 1:] - fl[:, :-1] - f["plude"].double()).abs()
    scale = (f["ptenq"].double().abs() + f["ptenl"].double().abs() + f["pteni"].double().abs()) * dp / p.rg \
        + fl[:, 1:].abs() + fl[:, :-1].abs() + f["plude"].double().abs()
    r = res / (scale + 1e-300)
    # each tendency is an FP32 sum of a few process terms; a nearly cancelling sum carries its largest term's
    # rounding: the bound is on the worst level and on the mean
    assert float(r.max()) < 1e-3 and float(r.mean()) < 1e-6
    del dp, tw, fl, res, scale, r
    blocks = [0, 1717, 3333, 5119]
    host = {k: f[k][blocks].cpu().numpy().copy() for k in C2.INPUT_NAMES}
    out = {k: f[k][blocks].double().cpu().numpy() for k in C2.OUTPUT_NAMES}
    del f
    torch.cuda.empty_cache()
    n = len(blocks) * nproma
    for k in C2.OUTPUT_NAMES:
        host[k] = np.full_like(out[k], np.nan, dtype=np.float32)
    r64, r32 = run_oracle("dp", p, q2, host, n, nproma, klev), run_oracle("sp", p, q2, host, n, nproma, klev)
    for k in C2.OUTPUT_NAMES:
        assert relrms(out[k], r64[k]) <= 2.0 * relrms(r32[k], r64[k]) + 1e-6, k
