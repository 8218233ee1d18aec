"""CLOUD_S_LAYER with LEPCLD2 (cloud_s_layer.F90:112-176): SATUR at T + TSPHY*TENDENCY%T, CLOUDSC2 on that
saturation humidity, and the wrap-up, in one launch.

CPU pins of the oracle:
  - SATUR (satur.F90:110-141):
    - at T = RTT every Tetens branch gives e = R2ES, so q_sat = (R2ES/p)/(1 - RETV R2ES/p);
    - where e/p exceeds 0.5, q_sat = 0.5/(1 - 0.5 RETV);
    - the LPHYLIN form (water and ice pressures weighted by FOEALFA) equals FOEEWM to rounding.
  - the layer's tendencies, fluxes and cover equal CLOUDSC2's on the PQSAT the layer reports;
  - the wrap-up:
    - PA + TSPHY PTENDENCY_LOC_A = PCLC on every level (PCLC = 0 above KTDIA);
    - rain and snow tendencies are zero on every column of every block, padding included;
    - PPRECTYPE and PFZRA are zero;
    - the per-level water closure of CLOUDSC2 holds.
  - optional outputs may be NULL without changing the others.
GPU: FP64 equals the oracle to 1e-12 (default and LPHYLIN, KTDIA 1 and 5, ragged last block); FP32 is as accurate
as the FP32 oracle.
"""
import numpy as np
import pytest

import oracle
from openifs_amd import cloudsc2 as C2
from tests.helpers import relrms
from tests.test_cloudsc2 import PTSPHY, params2, water_residual


def layer_case(p, seed, ngptot, nproma, klev, prec="dp"):
    c2 = C2.synth_host(p, seed, ngptot, nproma, klev, "dp")
    rng = np.random.default_rng(seed + 7)
    pa = rng.uniform(0.0, 1.0, c2["pqm1"].shape)
    return C2.layer_host(c2, pa, prec)


def run_layer(prec, p, q2, arr, ngptot, nproma, klev, ktdia=1, ldrain1d=False, drop=()):
    dt = np.float64 if prec == "dp" else np.float32
    a = {k: v.astype(dt) for k, v in arr.items() if k not in drop}
    oracle.cloud_s_layer_blocked(prec, p, q2, C2.reference_eta(klev), ngptot, nproma, ktdia, klev, ldrain1d, PTSPHY,
                                 a, nthreads=8)
    return a


@pytest.fixture(scope="module")
def case(params137):
    return layer_case(params137, 5, 500, 32, 137)  # 16 blocks, the last with 20 of 32 columns


def test_satur_at_triple_point_and_clip(params137, case):
    p = params137
    a = {k: v.copy() for k, v in case.items()}
    a["pt"][:, 60] = p.rtt
    a["ptendency_t"][:, 60] = 0.0
    a["pt"][:, 2] = 250.0  # upper stratosphere: e/p > 0.5
    a["ptendency_t"][:, 2] = 0.0
    for lphylin in (0, 1):
        out = run_layer("dp", p, params2(p, lphylin), a, 500, 32, 137)
        full = slice(0, 15)  # the 15 full blocks; the padding of the last one is not written
        e = p.r2es / a["prsf1"][full, 60]
        np.testing.assert_allclose(out["pqsat"][full, 60], e / (1.0 - p.retv * e), rtol=1e-14)
        assert np.all(a["prsf1"][:, 2] < 2.0 * p.r2es)  # the clip is reached
        np.testing.assert_allclose(out["pqsat"][full, 2], 0.5 / (1.0 - 0.5 * p.retv), rtol=1e-15)


def test_satur_lphylin_form_equals_foeewm(params137, case):
    a0 = run_layer("dp", params137, params2(params137, 0), case, 500, 32, 137)
    q2 = params2(params137, 0)
    q2.lphylin = 1
    # the phase function and thresholds of CLOUDSC2 change with LPHYLIN, but SATUR's value does not
    a1 = run_layer("dp", params137, q2, case, 500, 32, 137)
    real = ~np.isnan(a0["pqsat"])
    np.testing.assert_allclose(a1["pqsat"][real], a0["pqsat"][real], rtol=1e-13)


@pytest.mark.parametrize("lphylin,ktdia", [(0, 1), (1, 1), (0, 5)])
def test_layer_is_cloudsc2_on_its_pqsat(params137, case, lphylin, ktdia):
    p, q2 = params137, params2(params137, lphylin)
    lay = run_layer("dp", p, q2, case, 500, 32, 137, ktdia)
    c2 = {c2n: case[ln].copy() for ln, c2n in C2.LAYER_FROM_CLOUDSC2.items()}
    c2["pqs"] = np.where(np.isnan(lay["pqsat"]), 0.3, lay["pqsat"])  # levels above KTDIA are not read
    for name, kind, role in C2.CLOUDSC2_ARRAYS:
        if role == "o":
            c2[name] = np.full(C2.blocked_shape(kind, 32, 137, 16), np.nan)
    oracle.cloudsc2_blocked("dp", p, q2, C2.reference_eta(137), 500, 32, ktdia, 137, False, PTSPHY, c2, nthreads=8)
    for ln, c2n in [("ptendency_loc_t", "ptent"), ("ptendency_loc_q", "ptenq"), ("ptendency_loc_l", "ptenl"),
                    ("ptendency_loc_i", "pteni"), ("pclc", "pclc"), ("pfplsl", "pfplsl"), ("pfplsn", "pfplsn"),
                    ("pfhpsl", "pfhpsl"), ("pfhpsn", "pfhpsn"), ("pcovptot", "pcovptot")]:
        np.testing.assert_array_equal(lay[ln], c2[c2n], err_msg=ln)


def test_wrap_up(params137, case):
    p = params137
    a = run_layer("dp", p, params2(p, 0, 1), case, 500, 32, 137, ktdia=5)
    real = np.zeros((16, 32), bool)
    real.reshape(-1)[:500] = True
    lhs = (a["pa"] + PTSPHY * a["ptendency_loc_a"]).transpose(0, 2, 1)[real]
    np.testing.assert_allclose(lhs, a["pclc"].transpose(0, 2, 1)[real], rtol=0, atol=1e-14)
    assert np.all(a["pclc"].transpose(0, 2, 1)[real][:, :4] == 0.0)
    for n in ("ptendency_loc_r", "ptendency_loc_s"):
        assert np.all(a[n] == 0.0), n  # every column, padding included
    for n in ("pprectype", "pfzra"):
        assert np.all(a[n][real] == 0.0) and np.all(np.isnan(a[n][~real])), n
    layer_as_c2 = {"ptenq": a["ptendency_loc_q"], "ptenl": a["ptendency_loc_l"], "pteni": a["ptendency_loc_i"],
                   "pfplsl": a["pfplsl"], "pfplsn": a["pfplsn"], "paphp1": a["prs1"], "plude": a["plude"]}
    r = water_residual(p, {k: v[:15] for k, v in layer_as_c2.items()}, ktdia=5)
    assert r.max() < 1e-12


def test_optional_outputs_may_be_null(params137, case):
    full = run_layer("dp", params137, params2(params137), case, 500, 32, 137)
    opt = [n for n, k, r, o in C2.LAYER_ARRAYS if o and r == "o"]
    part = run_layer("dp", params137, params2(params137), case, 500, 32, 137, drop=opt)
    for n, k, r, o in C2.LAYER_ARRAYS:
        if r == "o" and not o:
            np.testing.assert_array_equal(part[n], full[n], err_msg=n)


# ----------------------------------------------------------------------------------------------------------------
# GPU parity
# ----------------------------------------------------------------------------------------------------------------
def gpu_layer(p, q2, host, ngptot, nproma, klev, prec, ktdia=1, drop=()):
    import torch

    dt = torch.float64 if prec == "dp" else torch.float32
    f = {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda", dt) for k, v in host.items() if k not in drop}
    ceta = torch.from_numpy(C2.reference_eta(klev)).to("cuda", dt)
    C2.cloud_s_layer_blocked(f, p, q2, ceta, ngptot, nproma, klev, PTSPHY, prec, ktdia)
    torch.cuda.synchronize()
    return {n: f[n].double().cpu().numpy() for n, k, r, o in C2.LAYER_ARRAYS if r == "o" and n in f}


OUTS = [n for n, k, r, o in C2.LAYER_ARRAYS if r == "o"]


@pytest.mark.gpu
@pytest.mark.parametrize("lphylin,ktdia,ngptot,nproma,klev", [(0, 1, 2048, 64, 137), (1, 1, 2048, 64, 137),
                                                                (0, 5, 777, 32, 91)])
def test_gpu_layer_fp64_matches_oracle(params137, params91, lphylin, ktdia, ngptot, nproma, klev):
    p = params137 if klev == 137 else params91
    q2 = params2(p, lphylin, 1)
    host = layer_case(p, 3, ngptot, nproma, klev, "dp")
    print(f"cloud_s_layer dp lphylin={lphylin} ktdia={ktdia} {ngptot}x{klev}", flush=True)
    out = gpu_layer(p, q2, host, ngptot, nproma, klev, "dp", ktdia)
    ref = run_layer("dp", p, q2, host, ngptot, nproma, klev, ktdia)
    for n in OUTS:
        np.testing.assert_array_equal(np.isnan(out[n]), np.isnan(ref[n]), err_msg=n)
        assert relrms(out[n], ref[n]) <= 1e-12, n


@pytest.mark.gpu
def test_gpu_layer_fp32_as_accurate_as_fp32_oracle(params137):
    p, ngptot, nproma, klev = params137, 4096, 128, 137
    q2 = params2(p)
    host = layer_case(p, 9, ngptot, nproma, klev, "sp")
    print("cloud_s_layer sp", flush=True)
    out = gpu_layer(p, q2, host, ngptot, nproma, klev, "sp")
    r64 = run_layer("dp", p, q2, host, ngptot, nproma, klev)
    r32 = run_layer("sp", p, q2, host, ngptot, nproma, klev)
    for n in OUTS:
        np.testing.assert_array_equal(np.isnan(out[n]), np.isnan(r64[n]), err_msg=n)
        assert relrms(out[n], r64[n]) <= 2.0 * relrms(r32[n], r64[n]) + 1e-6, n


@pytest.mark.gpu
def test_gpu_layer_optional_outputs_null(params137):
    p, ngptot, nproma, klev = params137, 1000, 64, 137
    q2 = params2(p)
    host = layer_case(p, 13, ngptot, nproma, klev, "sp")
    opt = [n for n, k, r, o in C2.LAYER_ARRAYS if o and r == "o"]
    print("cloud_s_layer sp, optional outputs NULL", flush=True)
    full = gpu_layer(p, q2, host, ngptot, nproma, klev, "sp")
    part = gpu_layer(p, q2, host, ngptot, nproma, klev, "sp", drop=opt)
    for n in part:
        np.testing.assert_array_equal(part[n], full[n], err_msg=n)
