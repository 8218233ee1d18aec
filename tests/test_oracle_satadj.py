"""Pins of the CLOUD_SATADJ oracle (no GPU): QSATMIXADJ against an exact bisection root, the per-point water
and enthalpy budgets, the reference's write set, and closed-form cases for each process of the default path
(inert clear air, supersaturated clear air, evaporation under heating, condensation in an overcast layer under
cooling, clear-sky condensation, negative-condensate clean-up)."""
import numpy as np
import pytest

import oracle
from openifs_amd import satadj as S

from tests.satadj_cases import block_arrays, blocked_case, controlled_block, qsat_mix, saturation_root

DT = 900.0


def run(params, arrays, kidia=1, kfdia=None, prec="dp"):
    klon, klev = arrays["pt"].shape
    oracle.satadj_block(prec, params, kidia, kfdia or klon, klon, klev, DT, arrays)
    return arrays


def test_qsatmixadj_matches_exact_root(params137):
    p = params137
    rng = np.random.default_rng(11)
    n = 4000
    t = rng.uniform(285.0, 305.0, n)  # stays above RTWAT along the adjustment: liquid only
    pp = rng.uniform(50000.0, 100000.0, n)
    s = rng.uniform(-0.2, 0.2, n)  # initial relative supersaturation
    q = qsat_mix(p, t, pp) * (1.0 + s)
    qn = oracle.qsatmixadj(p, t, q, pp)
    root = saturation_root(p, t, q, pp)
    err = np.abs(qn - root) / root
    # two Newton steps: the error is fourth order in the initial supersaturation (measured constant 0.0075);
    # a dropped or mis-signed dqs/dT term would leave a first- or second-order error
    m = np.abs(s) > 1e-3
    assert np.max(err[m] / s[m] ** 4) < 0.03
    assert np.max(err) < 3e-5


def test_qsatmixadj_fixed_point_at_saturation(params137):
    p = params137
    t = np.linspace(200.0, 310.0, 500)
    pp = np.full_like(t, 70000.0)
    qs = qsat_mix(p, t, pp)
    np.testing.assert_allclose(oracle.qsatmixadj(p, t, qs, pp), qs, rtol=1e-14)


@pytest.mark.parametrize("prec", ["dp", "sp", "mixed"])
def test_satadj_budgets_and_write_set(params137, prec):
    """Every process moves water between vapour and condensate and heats by the latent heat of what it moves:
    per point tq + tl + ti = 0 and tt = Lv/cp tl + Ls/cp ti.  Levels above NCLDTOP are zero and columns outside
    KIDIA:KFDIA are not written."""
    p = params137
    arr = blocked_case(p, 7, 96, 48, 137, "dp")
    a = block_arrays(arr, 1, prec)
    run(p, a, 3, 45, prec)
    o = {n: a[n].astype(np.float64) for n in S.SATADJ_OUTPUTS}
    for n in S.SATADJ_OUTPUTS:
        assert np.all(np.isnan(o[n][:2])) and np.all(np.isnan(o[n][45:])), n
        assert np.all(o[n][2:45, :p.ncldtop - 1] == 0.0), n
        assert np.all(np.isfinite(o[n][2:45])), n
    tt, tq, tl, ti = (o[f"ptendency_loc_{x}"][2:45] for x in "tqli")
    tol = 1e-12 if prec == "dp" else 2e-6
    wscale = np.abs(tq) + np.abs(tl) + np.abs(ti) + 1e-30
    assert np.max(np.abs(tq + tl + ti) / wscale) < tol
    h = tt - p.ralvdcp * tl - p.ralsdcp * ti
    assert np.max(np.abs(h) / (np.abs(tt) + p.ralsdcp * (np.abs(tl) + np.abs(ti)) + 1e-30)) < tol
    # the synthetic case exercises condensation and evaporation
    assert np.mean(tl + ti > 0) > 0.05 and np.mean(tl + ti < 0) > 0.05


def test_satadj_blocked_equals_per_block(params137):
    p = params137
    arr = blocked_case(p, 3, 100, 32, 137, "dp")
    oracle.satadj_blocked("dp", p, 100, 32, 137, DT, arr, nthreads=2)
    for b in range(4):
        a = block_arrays(arr, b)
        kf = min(32, 100 - 32 * b)
        run(p, a, 1, kf)
        np.testing.assert_array_equal(a["ptendency_loc_t"][:kf], arr["ptendency_loc_t"][b].T[:kf])
        np.testing.assert_array_equal(a["ptendency_loc_l"][:kf], arr["ptendency_loc_cld"][b, 0].T[:kf])
        np.testing.assert_array_equal(a["ptendency_loc_i"][:kf], arr["ptendency_loc_cld"][b, 1].T[:kf])
    # species 3..5 of the local tendency are not written
    assert np.all(np.isnan(arr["ptendency_loc_cld"][:, 2:]))


def test_inert_subsaturated_clear_air(params137):
    p = params137
    a = controlled_block(16, 137)
    rng = np.random.default_rng(2)
    a["pt"][:] = rng.uniform(240.0, 300.0, a["pt"].shape)
    a["pq"][:] = qsat_mix(p, a["pt"], a["pap"]) * rng.uniform(0.1, 0.7, a["pt"].shape)
    run(p, a)
    for n in S.SATADJ_OUTPUTS:
        np.testing.assert_array_equal(a[n], 0.0)


def test_supersaturated_clear_warm_air_ends_saturated(params137):
    p = params137
    a = controlled_block(16, 137)
    rng = np.random.default_rng(4)
    a["pt"][:] = rng.uniform(280.0, 300.0, a["pt"].shape)
    qs = qsat_mix(p, a["pt"], a["pap"])
    a["pq"][:] = qs * rng.uniform(1.01, 1.05, a["pt"].shape)
    run(p, a)
    k0 = p.ncldtop - 1
    tt, tq, ta, tl, ti = (a[f"ptendency_loc_{x}"][:, k0:] for x in "tqali")
    t1 = a["pt"][:, k0:] + DT * tt
    q1 = a["pq"][:, k0:] + DT * tq
    excess = a["pq"][:, k0:] - qs[:, k0:]
    # the linearised adjustment leaves a second-order residual only
    resid = np.abs(q1 - qsat_mix(p, t1, a["pap"][:, k0:]))
    assert np.max(resid / excess) < 0.02
    np.testing.assert_allclose(ta * DT, 1.0, rtol=1e-14)  # warm: the layer becomes overcast
    np.testing.assert_array_equal(ti, 0.0)
    assert np.all(tl > 0)


def test_evaporation_under_heating(params137):
    """Warm cloud heated at the same rate in the CML tendency and the radiation forcing: the cloudy part evaporates
    min(A * dqs, L + I) with dqs the rise of the saturation humidity for the forced temperature, split in the
    proportion of liquid and ice.  A <= 0.85 keeps the end state subsaturated even where ice evaporation cools by
    Ls/cp, so the supersaturation check stays idle."""
    p = params137
    a = controlled_block(32, 137)
    rng = np.random.default_rng(5)
    shp = a["pt"].shape
    a["pt"][:] = rng.uniform(280.0, 300.0, shp)
    qs = qsat_mix(p, a["pt"], a["pap"])
    a["pq"][:] = qs
    a["pa"][:] = rng.uniform(0.2, 0.85, shp)
    a["pl"][:] = 1e-5 * np.exp(rng.normal(0, 1.5, shp))
    a["pi"][:] = 1e-5 * np.exp(rng.normal(0, 1.5, shp))
    heat = rng.uniform(1e-6, 2e-5, shp)
    a["phrsw"][:] = heat
    a["ptendency_cml_t"][:] = heat
    run(p, a)
    k0 = p.ncldtop - 1
    sl = np.s_[:, k0:]
    dqs = saturation_root(p, a["pt"] + DT * heat, qs, a["pap"]) - qs
    assert np.all(dqs[sl] > 0)
    evap = np.minimum(a["pa"] * dqs, a["pl"] + a["pi"])
    liq = a["pl"] / (a["pl"] + a["pi"])
    np.testing.assert_allclose(-a["ptendency_loc_l"][sl] * DT, (liq * evap)[sl], rtol=1e-5)
    np.testing.assert_allclose(-a["ptendency_loc_i"][sl] * DT, ((1 - liq) * evap)[sl], rtol=1e-5)
    np.testing.assert_array_equal(a["ptendency_loc_a"][sl], 0.0)


def test_overcast_cooling_condenses_to_saturation(params137):
    """Saturated overcast layer cooled by the same rate in the CML tendency and the radiation forcing: the
    cloudy condensation plus the supersaturation check leave the final state saturated and conserve water
    and enthalpy."""
    p = params137
    a = controlled_block(32, 137)
    rng = np.random.default_rng(6)
    shp = a["pt"].shape
    a["pt"][:] = rng.uniform(280.0, 300.0, shp)
    qs = qsat_mix(p, a["pt"], a["pap"])
    a["pq"][:] = qs
    a["pa"][:] = 1.0
    a["pl"][:] = 1e-5
    cool = -rng.uniform(1e-6, 1e-5, shp)
    a["phrlw"][:] = cool
    a["ptendency_cml_t"][:] = cool
    run(p, a)
    k0 = p.ncldtop - 1
    sl = np.s_[:, k0:]
    t1 = (a["pt"] + DT * (cool + a["ptendency_loc_t"]))[sl]
    q1 = (a["pq"] + DT * a["ptendency_loc_q"])[sl]
    cond = (a["ptendency_loc_l"] * DT)[sl]
    assert np.all(cond > 0)
    np.testing.assert_allclose(q1, qsat_mix(p, t1, a["pap"][sl]), rtol=1e-6)
    np.testing.assert_allclose(a["ptendency_loc_t"][sl], p.ralvdcp * a["ptendency_loc_l"][sl], rtol=1e-12)


@pytest.mark.parametrize("forcing", ["radiation", "ascent"])
def test_clear_sky_condensation_relation(params137, forcing):
    """Clear air between the critical and full saturation humidity, cooled at the same rate in the CML tendency
    and in the forcing (radiation, or adiabatic ascent dT/dt = (R/cp) T omega / p): new cloud forms with an
    increase dA and condensate -dqs dA / 2 (a uniform humidity distribution in the clear part, Tiedtke 1993)."""
    p = params137
    a = controlled_block(32, 137)
    rng = np.random.default_rng(8)
    shp = a["pt"].shape
    a["pt"][:] = rng.uniform(280.0, 300.0, shp)
    qs = qsat_mix(p, a["pt"], a["pap"])
    a["pq"][:] = qs * rng.uniform(0.93, 0.97, shp)
    cool = -rng.uniform(3e-5, 1.5e-4, shp)  # below the limit dp/dt * R T / (cp p) on this grid
    if forcing == "radiation":
        a["phrlw"][:] = cool
    else:
        a["pvervel"][:] = cool * a["pap"] / (p.rd / p.rcpd * a["pt"])
    a["ptendency_cml_t"][:] = cool
    run(p, a)
    k0 = p.ncldtop - 1
    sl = np.s_[:, k0:]
    dqs = (saturation_root(p, a["pt"] + DT * cool, qs, a["pap"]) - qs)[sl]
    ta, tl = a["ptendency_loc_a"][sl] * DT, a["ptendency_loc_l"][sl] * DT
    act = ta > 0
    assert np.mean(act) > 0.5
    np.testing.assert_allclose(tl[act], -0.5 * dqs[act] * ta[act], rtol=1e-5)
    assert np.all(ta <= 1.0)


def test_negative_liquid_is_cleaned_up(params137):
    """Inert subsaturated air whose liquid would go negative: the tidy-up returns exactly the deficit
    L + dt*CML_L to vapour (cloud_satadj.F90:827-833), cooling by Lv/cp times it."""
    p = params137
    a = controlled_block(8, 137)
    a["pt"][:] = 285.0
    a["pq"][:] = 0.5 * qsat_mix(p, a["pt"], a["pap"])
    a["pl"][:] = 1e-7
    a["ptendency_cml_l"][:] = -2e-7 / DT
    run(p, a)
    k0 = p.ncldtop - 1
    d = (a["pl"] + DT * a["ptendency_cml_l"])[:, k0:]
    np.testing.assert_allclose(a["ptendency_loc_l"][:, k0:], -d, rtol=1e-14)
    np.testing.assert_allclose(a["ptendency_loc_q"][:, k0:], d, rtol=1e-14)
    np.testing.assert_allclose(a["ptendency_loc_t"][:, k0:], -p.ralvdcp * d, rtol=1e-14)


def test_unsupported_configuration_is_rejected(params137):
    import copy

    p = copy.copy(params137)
    p.lspp = 1
    a = controlled_block(4, 137)
    rc = S.call_block(oracle.lib().oracle_satadj_dp, p, 1, 4, 4, 137, DT, a, "dp")
    assert rc != 0
