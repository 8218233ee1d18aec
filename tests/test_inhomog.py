"""LCLOUD_INHOMOG: the KK2000 autoconversion and accretion rates enhanced for sub-grid lognormal cloud and rain
water (cloudsc.F90:1937-1971, 2075-2098; Boutle et al. 2013), instead of the constants RCL_INHOMOGAUT/ACC.

CPU pins of the oracle:
  - the enhancement factors are moments of lognormal distributions:
      E[x^b] / E[x]^b for the cloud water, with fractional standard deviation f_c (autoconversion), and
      E[x^b y^b] / (E[x]^b E[y]^b) for correlated cloud and rain water (accretion), with correlation
      1 - 0.8 a between the logarithms.
    Both are evaluated here by Gauss-Hermite quadrature of the definition, not from the closed forms. In an
    overcast warm layer whose only process is warm rain, the liquid left by the implicit step,
    ql / (1 + dt (k_aut + k_acc)), must match these factors. f_c is the liquid FSD the same run reports in PFSD
    (LRAD_CLOUD_INHOMOG). The rain FSD is Boutle et al.'s fit in grid length and cloud fraction.
  - the switch needs PGAW: the entries without it reject the configuration.
GPU: the diagnostics entry against the oracle (FP64 1e-12; FP32 as accurate as the FP32 oracle) and the per-block
entry against the blocked one bit for bit.
"""
import math

import numpy as np
import pytest

import oracle
from openifs_amd import cloudsc as C
from tests.helpers import columns, relrms
from tests.test_fsd import fsd_inputs, run
from tests.test_budget import TSPHY
from tests.test_oracle_processes import DT, _column

NC_SEA = 50.0  # cm^-3


def inhomog_params(p, lrad=1):
    q = type(p).from_buffer_copy(p)
    q.lcloud_inhomog = 1
    q.lrad_cloud_inhomog = lrad
    return q


def gridlen_to_pgaw(p, km):
    """PGAW for a grid length of `km` km, ZGRIDLEN = 2 RA sqrt(RPI PGAW) / 1000 (:1944)."""
    return (km * 1000.0 / (2.0 * p.ra)) ** 2 / p.rpi


def lognormal_moment_ratio(f, b, n=120):
    """E[x^b] / E[x]^b for lognormal x with fractional standard deviation f, by Gauss-Hermite quadrature."""
    z, w = np.polynomial.hermite_e.hermegauss(n)
    w = w / math.sqrt(2.0 * math.pi)
    s = math.sqrt(math.log(1.0 + f * f))
    x = np.exp(-0.5 * s * s + s * z)  # mean 1
    return float(np.sum(w * x ** b)) / float(np.sum(w * x)) ** b


def bivariate_moment_ratio(fc, fr, rho, b, n=100):
    """E[x^b y^b] / (E[x]^b E[y]^b) for lognormal x, y with correlation rho between log x and log y."""
    z, w = np.polynomial.hermite_e.hermegauss(n)
    w = w / math.sqrt(2.0 * math.pi)
    z1, z2 = np.meshgrid(z, z, indexing="ij")
    ww = np.outer(w, w)
    sc, sr = math.sqrt(math.log(1.0 + fc * fc)), math.sqrt(math.log(1.0 + fr * fr))
    x = np.exp(-0.5 * sc * sc + sc * z1)
    y = np.exp(-0.5 * sr * sr + sr * (rho * z1 + math.sqrt(1.0 - rho * rho) * z2))
    return float(np.sum(ww * x ** b * y ** b)) / (float(np.sum(ww * x)) ** b * float(np.sum(ww * y)) ** b)


def test_quadrature_reproduces_known_moments():
    # f = 0: no enhancement; b = 1: the mean itself; and E[x^2]/E[x]^2 = 1 + f^2 for any distribution
    assert lognormal_moment_ratio(0.0, 2.47) == pytest.approx(1.0, abs=1e-12)
    assert lognormal_moment_ratio(0.8, 1.0) == pytest.approx(1.0, rel=1e-12)
    assert lognormal_moment_ratio(0.8, 2.0) == pytest.approx(1.64, rel=1e-10)
    # independent x, y: the joint moment factorises
    assert bivariate_moment_ratio(0.7, 0.5, 0.0, 1.15) == pytest.approx(
        lognormal_moment_ratio(0.7, 1.15) * lognormal_moment_ratio(0.5, 1.15), rel=1e-10)


def _warm_layer(p, qr):
    klev = 137
    # overcast (no erosion or clear-sky processes), slightly subsaturated, warm; rain in every layer
    arrs = _column(p, klev, lambda s: 288.0 + 0.0 * s, 0.99, 1.0, [5e-4, 0.0, qr, 0.0])
    arrs["plsm"][0] = 0.0  # sea
    arrs["pgaw"][0] = gridlen_to_pgaw(p, 25.0)
    return arrs


def test_oracle_autoconversion_factor_is_lognormal_moment(params137):
    p = params137
    q = inhomog_params(p)
    arrs = _warm_layer(p, 0.0)
    ql = arrs["pclv"][0, :, 0].copy()
    oracle.cloudsc_block("dp", q, 1, 1, 1, 137, DT, arrs)
    k0 = p.ncldtop - 1
    fsd = arrs["pfsd"][0]
    assert np.all((fsd[k0:] > 0.1) & (fsd[k0:] < 3.575))  # the liquid FSD, not clipped
    for k in range(k0, 137):
        e_aut = lognormal_moment_ratio(fsd[k], p.rcl_kkbauq)
        k_aut = e_aut * p.rcl_kkaau * ql[k] ** (p.rcl_kkbauq - 1.0) * NC_SEA ** p.rcl_kkbaun
        want = ql[k] / (1.0 + DT * k_aut)
        got = ql[k] + DT * arrs["ptendency_loc_cld"][0, k, 0]
        assert got == pytest.approx(want, rel=1e-9), k
    # the enhancement differs from the constant RCL_INHOMOGAUT it replaces
    assert abs(lognormal_moment_ratio(fsd[k0], p.rcl_kkbauq) - p.rcl_inhomogaut) > 0.05


def test_oracle_accretion_factor_is_bivariate_moment(params137):
    p = params137
    q = inhomog_params(p)
    qr = 2e-4
    arrs = _warm_layer(p, qr)
    ql = arrs["pclv"][0, :, 0].copy()
    oracle.cloudsc_block("dp", q, 1, 1, 1, 137, DT, arrs)
    k0 = p.ncldtop - 1
    fsd = arrs["pfsd"][0]
    za = 1.0
    glen = 25.0
    # rain FSD, Boutle et al. (2013) fit, overcast branch (:2081-2088)
    phir = (glen * za) ** (1.0 / 3.0) * ((0.11 * glen * za) ** 1.14 + 1.0) ** -0.22
    fr = 0.3 * phir
    rho = 1.0 - 0.8 * za
    for k in range(k0 + 1, 137):
        e_aut = lognormal_moment_ratio(fsd[k], p.rcl_kkbauq)
        e_acc = bivariate_moment_ratio(fsd[k], fr, rho, p.rcl_kkbac)
        k_aut = e_aut * p.rcl_kkaau * ql[k] ** (p.rcl_kkbauq - 1.0) * NC_SEA ** p.rcl_kkbaun
        k_acc = e_acc * p.rcl_kkaac * ql[k] ** (p.rcl_kkbac - 1.0) * qr ** p.rcl_kkbac
        want = ql[k] / (1.0 + DT * (k_aut + k_acc))
        got = ql[k] + DT * arrs["ptendency_loc_cld"][0, k, 0]
        assert got == pytest.approx(want, rel=1e-8), k
    assert k_acc > k_aut  # the accretion term matters in this layer


def test_entries_without_pgaw_reject(params137):
    """LCLOUD_INHOMOG needs PGAW: the per-block entry with PGAW = NULL and the plain blocked entry refuse it (checked
    before any device work)."""
    import torch

    q = inhomog_params(params137, lrad=0)
    arrs = C.allocate_block(8, 137, "dp")
    del arrs["pgaw"]
    with pytest.raises(C.CloudscError) as e:
        C.CLOUDSC(q, 1, 8, 8, 137, TSPHY, arrs, precision="dp")
    assert e.value.code == -3 and "PGAW" in str(e.value)  # CLOUDSC_ENULL
    f = C.BlockedFields(64, 32, 137, "sp", "cpu")
    with pytest.raises(C.CloudscError) as e:
        C.cloudsc_blocked(f, q, TSPHY, stream=0)
    assert "LCLOUD_INHOMOG" in str(e.value)
    del torch


def test_oracle_switch_changes_only_warm_rain_dependents(params137):
    """With LCLOUD_INHOMOG the state differs from the constant-factor run only where warm rain is active; with
    warm rain off everywhere (no liquid) the two runs agree exactly."""
    p, ngptot, nproma, klev = params137, 256, 32, 137
    inp = fsd_inputs(p, ngptot, nproma, klev, 17)
    for n in ("pclv", "ptendency_cml_cld"):
        inp[n] = inp[n].copy()
        inp[n][:, 0] = 0.0  # no liquid
    inp["plude"] = np.zeros_like(inp["plude"])
    inp["pludeli"] = np.zeros_like(inp["pludeli"])
    q0 = type(p).from_buffer_copy(p)
    q0.lrad_cloud_inhomog = 1
    a = run("dp", q0, inp, ngptot, nproma, klev)
    b = run("dp", inhomog_params(p), inp, ngptot, nproma, klev)
    for n in C.OUTPUT_NAMES:
        np.testing.assert_array_equal(a[n], b[n], err_msg=n)


# ----------------------------------------------------------------------------------------------------------------
# GPU parity
# ----------------------------------------------------------------------------------------------------------------
def _gpu_run(p, q, prec, ngptot, nproma, klev, seed):
    import torch

    f = C.BlockedFields(ngptot, nproma, klev, prec, "cuda", fill=float("nan"))
    C.synth_device(f, p, seed)
    inp = {n: f[n].cpu().numpy().copy() for n in C.INPUT_NAMES}
    extra = fsd_inputs(p, ngptot, nproma, klev, seed)
    dt = torch.float64 if prec == "dp" else torch.float32
    diag = {"pgaw": torch.from_numpy(extra["pgaw"]).to("cuda", dt)}
    if q.lrad_cloud_inhomog:
        diag["pfsd"] = torch.full((f.ngpblks, klev, nproma), float("nan"), dtype=dt, device="cuda")
    C.cloudsc_blocked_diag(f, diag, q, TSPHY)
    torch.cuda.synchronize()
    out = {n: f[n].cpu().numpy() for n in C.OUTPUT_NAMES}
    if "pfsd" in diag:
        out["pfsd"] = diag["pfsd"].cpu().numpy()
    inp["pgaw"] = diag["pgaw"].cpu().numpy()
    return inp, out


FIELDS = ["ptendency_loc_t", "ptendency_loc_q", "ptendency_loc_a", "pfplsl", "pfplsn", "pcovptot"]


@pytest.mark.gpu
@pytest.mark.parametrize("lrad", [0, 1])
def test_gpu_inhomog_fp64_matches_oracle(params137, lrad):
    p, ngptot, nproma, klev = params137, 1500, 64, 137
    q = inhomog_params(p, lrad)
    print(f"inhomog fp64 lrad={lrad}", flush=True)
    inp, out = _gpu_run(p, q, "dp", ngptot, nproma, klev, 61)
    ref = run("dp", q, inp, ngptot, nproma, klev)
    names = FIELDS + (["pfsd"] if lrad else [])
    for n in names:
        g, r = columns(out[n], nproma)[:ngptot], columns(ref[n], nproma)[:ngptot]
        np.testing.assert_array_equal(np.isnan(g), np.isnan(r), err_msg=n)
        assert relrms(g, r) <= 1e-12, n
    g, r = columns(out["ptendency_loc_cld"], nproma)[:ngptot], columns(ref["ptendency_loc_cld"], nproma)[:ngptot]
    assert relrms(g, r) <= 1e-12
    # the switch is active: the warm-rain tendencies differ from the constant-factor run
    q0 = type(p).from_buffer_copy(p)
    q0.lrad_cloud_inhomog = 1
    base = run("dp", q0, inp, ngptot, nproma, klev)
    assert relrms(columns(base["pfplsl"], nproma), columns(ref["pfplsl"], nproma)) > 1e-3


@pytest.mark.gpu
def test_gpu_inhomog_fp32_as_accurate_as_fp32_oracle(params137):
    p, ngptot, nproma, klev = params137, 4096, 128, 137
    q = inhomog_params(p, 1)
    print("inhomog sp", flush=True)
    inp, out = _gpu_run(p, q, "sp", ngptot, nproma, klev, 63)
    r64 = run("dp", q, inp, ngptot, nproma, klev)
    r32 = run("sp", q, inp, ngptot, nproma, klev)
    for n in FIELDS + ["pfsd"]:
        g, a64, a32 = (columns(x[n], nproma) for x in (out, r64, r32))
        e_gpu, e_cpu = relrms(g, a64), relrms(a32, a64)
        print(f"{n}: gpu {e_gpu:.3e} fp32 oracle {e_cpu:.3e}")
        assert e_gpu <= 2.0 * e_cpu + 1e-5, n


@pytest.mark.gpu
def test_gpu_inhomog_per_block_equals_blocked(params137):
    p, nproma, klev = params137, 32, 137
    q = inhomog_params(p, 1)
    inp, out = _gpu_run(p, q, "sp", 3 * nproma, nproma, klev, 67)
    for blk in (0, 2):
        arrs = C.allocate_block(nproma, klev, "sp")
        for n, kind, role in C.BLOCKED_ARRAYS:
            if role != "o":
                arrs[n][...] = inp[n][blk].T
        arrs["pgaw"][...] = inp["pgaw"][blk]
        arrs["pfsd"][...] = np.nan
        C.CLOUDSC(q, 1, nproma, nproma, klev, TSPHY, arrs, precision="sp")
        np.testing.assert_array_equal(arrs["pfsd"], out["pfsd"][blk].T)
        for n in ("ptendency_loc_t", "ptendency_loc_cld", "pfplsl"):
            np.testing.assert_array_equal(arrs[n], out[n][blk].T, err_msg=n)
