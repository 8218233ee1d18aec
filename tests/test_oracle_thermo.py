"""Pins of the oracle's thermodynamic statement functions (fcttre.func.h,
fccld.func.h) by closed forms and by independent saturation formulas."""
import numpy as np
import pytest

import oracle

T = np.linspace(160.0, 330.0, 1701)


def murphy_koop_liquid(t):
    """Saturation vapour pressure over liquid water, Murphy & Koop (2005), Pa."""
    return np.exp(54.842763 - 6763.22 / t - 4.21 * np.log(t) + 0.000367 * t
                  + np.tanh(0.0415 * (t - 218.8)) * (53.878 - 1331.22 / t - 9.44523 * np.log(t) + 0.014025 * t))


def murphy_koop_ice(t):
    """Saturation vapour pressure over ice, Murphy & Koop (2005), Pa."""
    return np.exp(9.550426 - 5723.265 / t + 3.53068 * np.log(t) - 0.00728332 * t)


def test_foeeliq_is_saturation_pressure_over_water(params137):
    # FOEELIQ carries the factor RD/RV (e_s * RD/RV); Tetens is within 1 % of Murphy-Koop
    # over 233-320 K (0.3 % above 250 K)
    p = params137
    t = T[(T > 233.0) & (T < 320.0)]
    es = oracle.thermo(p, "foeeliq", t) * p.rv / p.rd
    dev = np.abs(es / murphy_koop_liquid(t) - 1.0)
    assert np.max(dev) < 1e-2 and np.max(dev[t > 250.0]) < 3e-3


def test_foeeice_is_saturation_pressure_over_ice(params137):
    p = params137
    t = T[(T > 190.0) & (T < 273.16)]
    es = oracle.thermo(p, "foeeice", t) * p.rv / p.rd
    assert np.max(np.abs(es / murphy_koop_ice(t) - 1.0)) < 5e-3


def test_triple_point_values(params137):
    p = params137
    for f in ("foeeliq", "foeeice", "foeewm", "foeew"):
        assert oracle.thermo(p, f, np.array([p.rtt]))[0] == pytest.approx(p.r2es, rel=1e-15)
    assert oracle.thermo(p, "foeeliq2ice", np.array([p.rtt]))[0] == 1.0


def test_foedelta_and_foeew(params137):
    p = params137
    d = oracle.thermo(p, "foedelta", T)
    assert np.all(d == np.where(T >= p.rtt, 1.0, 0.0))
    ew = oracle.thermo(p, "foeew", T)
    ref = np.where(T >= p.rtt, oracle.thermo(p, "foeeliq", T), oracle.thermo(p, "foeeice", T))
    np.testing.assert_allclose(ew, ref, rtol=1e-15)


def test_foealfa_shape(params137):
    p = params137
    a = oracle.thermo(p, "foealfa", T)
    assert np.all(a[T <= p.rtice] == 0.0) and np.all(a[T >= p.rtwat] == 1.0)
    mid = (T > p.rtice) & (T < p.rtwat)
    np.testing.assert_allclose(a[mid], ((T[mid] - p.rtice) / 23.0) ** 2, rtol=1e-13)


def test_foeewm_is_alpha_blend(params137):
    p = params137
    a = oracle.thermo(p, "foealfa", T)
    ref = a * oracle.thermo(p, "foeeliq", T) + (1 - a) * oracle.thermo(p, "foeeice", T)
    np.testing.assert_allclose(oracle.thermo(p, "foeewm", T), ref, rtol=1e-14)


def test_foedem_is_latent_weighted_log_derivative(params137):
    """FOEDEM = alpha L_v/c_p dln(e_l)/dT + (1-alpha) L_s/c_p dln(e_i)/dT (Clausius-Clapeyron form)."""
    p = params137
    h = 1e-4
    dl = (np.log(oracle.thermo(p, "foeeliq", T + h)) - np.log(oracle.thermo(p, "foeeliq", T - h))) / (2 * h)
    di = (np.log(oracle.thermo(p, "foeeice", T + h)) - np.log(oracle.thermo(p, "foeeice", T - h))) / (2 * h)
    a = oracle.thermo(p, "foealfa", T)
    ref = a * p.ralvdcp * dl + (1 - a) * p.ralsdcp * di
    np.testing.assert_allclose(oracle.thermo(p, "foedem", T), ref, rtol=1e-6)


def test_foeldcpm(params137):
    p = params137
    a = oracle.thermo(p, "foealfa", T)
    np.testing.assert_allclose(oracle.thermo(p, "foeldcpm", T), a * p.rlvtt / p.rcpd + (1 - a) * p.rlstt / p.rcpd,
                               rtol=1e-14)


def test_koop_limit(params137):
    p = params137
    ratio = oracle.thermo(p, "foeeliq", T) / oracle.thermo(p, "foeeice", T)
    np.testing.assert_allclose(oracle.thermo(p, "foeeliq2ice", T), ratio, rtol=1e-12)
    np.testing.assert_allclose(oracle.thermo(p, "fokoop", T), np.minimum(p.rkoop1 - p.rkoop2 * T, ratio), rtol=1e-12)
    # between 200 and 230 K the homogeneous-nucleation line is the binding limit
    cold = (T > 200.0) & (T < 230.0)
    np.testing.assert_allclose(oracle.thermo(p, "fokoop", T)[cold], (p.rkoop1 - p.rkoop2 * T)[cold], rtol=1e-15)
