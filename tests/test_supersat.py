"""CLOUD_SUPERSATCHECK with both KFTLIQICE modes and SLTEND's supersaturation adjustment (sltend.F90:370-409).

CPU: pins of the oracle (KFTLIQICE 2 against KFTLIQICE 1 and the RTHOMO split; SLTEND: conservation, no-op below
saturation, saturation of the adjusted state, levels above NCLDTOP untouched).  GPU: the gfx950 entries against
the oracle, FP64 to 1e-12 and FP32/mixed by the FP32 oracle's own distance to FP64."""
import numpy as np
import pytest

import oracle
from openifs_amd import cloudsc as C
from tests.helpers import relrms

TSPHY = 900.0


def _points(p, n, rng, tmin=200.0, tmax=310.0, rhmin=0.5, rhmax=1.5):
    t = rng.uniform(tmin, tmax, n)
    pp = rng.uniform(20000.0, 100000.0, n)
    es = np.minimum(oracle.thermo(p, "foeewm", t) / pp, 0.5)
    q = rng.uniform(rhmin, rhmax, n) * es / (1 - p.retv * es)
    a = rng.uniform(0.0, 1.0, n)
    return t, q, a, pp, np.full(n, 100000.0)


def test_kftliqice2_split(params137):
    """KFTLIQICE 2 changes only the partition: the same T, Q, A adjustments, all condensate liquid where
    T + T_ADJ >= RTHOMO and all ice below, summing exactly to -Q_ADJ."""
    p = params137
    t, q, a, pp, ps = _points(p, 5000, np.random.default_rng(1))
    k1 = oracle.supersatcheck_k(p, 1, t, q, a, pp, ps)
    k2 = oracle.supersatcheck_k(p, 2, t, q, a, pp, ps)
    for x, y in zip(k1[:3], k2[:3]):
        np.testing.assert_array_equal(x, y)
    tq, l2, i2 = k2[1], k2[3], k2[4]
    np.testing.assert_array_equal(l2 + i2, -tq)
    liquid = t + k2[0] >= p.rthomo
    assert np.all(i2[liquid] == 0) and np.all(l2[~liquid] == 0)
    assert np.any(liquid & (tq < 0)) and np.any(~liquid & (tq < 0))
    np.testing.assert_allclose(k1[3] + k1[4], -k1[1], rtol=1e-14, atol=0)


def sltend_case(p, ngptot, nproma, klev, seed, prec="dp"):
    """Blocked SLTEND inputs: the synthetic atmosphere (RH up to 1.05, so some points are supersaturated) with
    random dynamics (PCEN*) and physics (PTEN*) tendencies."""
    s = C.synth_host(p, seed, ngptot, nproma, klev, precision="dp")
    rng = np.random.default_rng(seed)
    sh = s["pt"].shape
    arr = {"pt": s["pt"], "pq": s["pq"], "pa": s["pa"], "prsf1": s["pap"], "prs1": s["paph"],
           "pcent": rng.normal(0, 1e-4, sh), "pcenq": s["pq"] * rng.normal(0, 1e-4, sh),
           "pcena": rng.normal(0, 1e-4, sh), "ptent": rng.normal(0, 1e-4, sh),
           "ptenq": s["pq"] * rng.normal(0, 1e-4, sh), "ptena": rng.normal(0, 1e-5, sh),
           "ptencld": rng.normal(0, 1e-9, s["pclv"].shape)}
    dt = np.float64 if prec == "dp" else np.float32
    out = {k: np.ascontiguousarray(v, dtype=dt) for k, v in arr.items()}
    for n in ("padja", "padjl", "padji"):
        out[n] = np.full(sh, np.nan, dtype=dt)
    return out


def test_sltend_oracle_properties(params137):
    p, nproma, klev = params137, 32, 137
    a0 = sltend_case(p, 64, nproma, klev, 3)
    a = {k: v.copy() for k, v in a0.items()}
    oracle.sltend_supersat_blocked("dp", p, 64, nproma, klev, TSPHY, a)
    k0 = p.ncldtop - 1
    d = {n: a[n] - a0[n] for n in ("ptent", "ptenq", "ptena")}
    dl, di = a["ptencld"][:, 0] - a0["ptencld"][:, 0], a["ptencld"][:, 1] - a0["ptencld"][:, 1]
    # above NCLDTOP nothing is touched
    for n in ("ptent", "ptenq", "ptena"):
        np.testing.assert_array_equal(a[n][:, :k0], a0[n][:, :k0])
    assert np.all(np.isnan(a["padja"][:, :k0]))
    np.testing.assert_array_equal(a["ptencld"][:, 2:], a0["ptencld"][:, 2:])
    # the diagnostics are the increments
    np.testing.assert_allclose(a["padjl"][:, k0:], dl[:, k0:], rtol=1e-12, atol=1e-24)
    np.testing.assert_allclose(a["padja"][:, k0:], d["ptena"][:, k0:], rtol=1e-12, atol=1e-24)
    # water and enthalpy
    scale = np.abs(d["ptenq"]) + np.abs(dl) + np.abs(di) + 1e-30
    assert np.max(np.abs(d["ptenq"] + dl + di) / scale) < 1e-10
    h = d["ptent"] - p.ralvdcp * dl - p.ralsdcp * di
    assert np.max(np.abs(h) / (np.abs(d["ptent"]) + 1e-30)) < 1e-10
    # some points adjust, many do not
    act = d["ptenq"][:, k0:] != 0
    assert 0.02 < act.mean() < 0.9
    # a second pass on the adjusted tendencies leaves a residual of second order only
    b = {k: v.copy() for k, v in a.items()}
    oracle.sltend_supersat_blocked("dp", p, 64, nproma, klev, TSPHY, b)
    r1 = np.abs(d["ptenq"][:, k0:][act])
    r2 = np.abs((b["ptenq"] - a["ptenq"])[:, k0:][act])
    assert np.median(r2 / r1) < 1e-2


def test_sltend_noop_when_subsaturated(params137):
    p, nproma, klev = params137, 16, 137
    a0 = sltend_case(p, 16, nproma, klev, 5)
    a0["pq"] *= 0.5
    a0["pcenq"][:] = 0
    a0["ptenq"][:] = 0
    a = {k: v.copy() for k, v in a0.items()}
    oracle.sltend_supersat_blocked("dp", p, 16, nproma, klev, TSPHY, a)
    for n in ("ptent", "ptenq", "ptena", "ptencld"):
        np.testing.assert_array_equal(a[n], a0[n])


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("kft", [1, 2])
@pytest.mark.parametrize("prec", ["dp", "sp", "mixed"])
def test_supersatcheck_drop_in(params137, kft, prec):
    """cloudsc_supersatcheck_*_c on one level (KIDIA = 3, KFDIA = KLON-2): FP64 and mixed equal the FP64 oracle
    on the same (FP32-rounded, for mixed) inputs; FP32 is as accurate as the FP32 oracle."""
    from openifs_amd import supersat as S

    p = params137
    n = 4096
    t, q, a, pp, ps = _points(p, n, np.random.default_rng(7))
    got = S.CLOUD_SUPERSATCHECK(p, 3, n - 2, n, 137, kft, t, q, a, pp, ps, precision=prec)
    for g in got:
        assert np.all(np.isnan(g[:2])) and np.all(np.isnan(g[-2:]))
    got = [g[2:-2].astype(np.float64) for g in got]
    sl = slice(2, n - 2)
    if prec == "dp":
        ref = oracle.supersatcheck_k(p, kft, t[sl], q[sl], a[sl], pp[sl], ps[sl])
        for g, r in zip(got, ref):
            assert relrms(g, r) <= 1e-12
    elif prec == "mixed":
        r32 = [x[sl].astype(np.float32).astype(np.float64) for x in (t, q, a, pp, ps)]
        ref = oracle.supersatcheck_k(p, kft, *r32)
        for g, r in zip(got, ref):
            assert relrms(g, r) <= 1e-6
    else:
        ref = oracle.supersatcheck_k(p, kft, t[sl], q[sl], a[sl], pp[sl], ps[sl])
        r32 = oracle.supersatcheck(p, t[sl], q[sl], a[sl], pp[sl], ps[sl], precision="sp")  # KFTLIQICE 1
        r64 = oracle.supersatcheck_k(p, 1, t[sl], q[sl], a[sl], pp[sl], ps[sl])
        for j in range(3):
            assert relrms(got[j], ref[j]) <= 2 * relrms(r32[j], r64[j]) + 1e-6
        bound = 2 * relrms(r32[1], r64[1]) + 1e-6
        assert relrms(got[3] + got[4], ref[3] + ref[4]) <= bound


@pytest.mark.gpu
@pytest.mark.parametrize("prec,ngptot,nproma,klev", [("dp", 2048, 64, 137), ("sp", 8192, 128, 137),
                                                     ("mixed", 8192, 128, 137), ("dp", 700, 128, 91)])
def test_sltend_matches_oracle(params137, params91, prec, ngptot, nproma, klev):
    import torch

    from openifs_amd import supersat as S

    p = params137 if klev == 137 else params91
    host = sltend_case(p, ngptot, nproma, klev, 11, prec)
    dev = {k: torch.from_numpy(v.copy()).cuda() for k, v in host.items()}
    S.sltend_supersat_blocked(dev, p, TSPHY, ngptot, nproma, klev, prec)
    torch.cuda.synchronize()
    ref = oracle.sltend_supersat_blocked(prec, p, ngptot, nproma, klev, TSPHY, {k: v.copy() for k, v in host.items()})
    ref64 = oracle.sltend_supersat_blocked("dp", p, ngptot, nproma, klev, TSPHY,
                                           {k: v.astype(np.float64) for k, v in host.items()})
    k0 = p.ncldtop - 1
    for n in ("ptent", "ptenq", "ptena", "ptencld", "padja", "padjl", "padji"):
        g = dev[n].cpu().numpy()
        # real columns only (the last block may be partial)
        col = lambda x: (x.transpose(0, 3, 2, 1) if x.ndim == 4 else x.transpose(0, 2, 1)).reshape(-1, *x.shape[1:-1][::-1])[:ngptot]
        g, r, r64 = col(g), col(ref[n]), col(ref64[n])
        np.testing.assert_array_equal(np.isnan(g), np.isnan(r), err_msg=n)
        if n.startswith("pten"):  # compare the increments, not the unchanged background
            b = col(host[n].astype(np.float64))
            g, r, r64 = g - b, r - b, r64 - b
        if prec == "dp":
            assert relrms(g, r) <= 1e-12, n
        else:
            assert relrms(g, r64) <= 2 * relrms(r, r64) + 1e-6, n
