"""Process-level pins of the oracle (FP64): single-process columns whose outcome follows from a published
formula or from a conservation law, not from the oracle's own expressions.

* Snow sedimentation: with snow the only active species, each level obeys the flux-form continuity equation
  discretised implicitly upwind, (q' - q)/dt = g (F_in - rho v q')/dp with rho = p/(R_d T) and the constant
  snow fall speed, and the flux leaving the level is rho v q'.
* Warm-rain autoconversion: Khairoutdinov and Kogan (2000, Mon. Wea. Rev. 128, eq. 29),
  dq_r/dt = 1350 q_c^2.47 N_c^-1.79 (N_c in cm^-3), with the IFS inhomogeneity enhancement factor, applied
  implicitly over the time step to an overcast warm layer with no other active process.
* Homogeneous freezing: below -38 C every bit of cloud liquid freezes within the step, and the only temperature
  change is the latent heat of fusion, L_f q_l / c_p.
"""
import numpy as np

import oracle
from openifs_amd import cloudsc as C

DT = 900.0
# gas constant from the CODATA 1986 Boltzmann and Avogadro constants (the values the IFS uses), dry air 28.9644 g/mol
R_DRY = 1000.0 * 1.380658e-23 * 6.0221367e23 / 28.9644
CP_DRY = 3.5 * R_DRY


def _column(params, klev, t_of_sigma, rh, a, cond):
    a_h, b_h = C.hybrid_levels(klev)
    ps = 100000.0
    paph = a_h + b_h * ps
    pap = C.standard_full_levels(a_h, b_h, ps)
    arrs = C.allocate_block(1, klev, "dp")
    t = np.broadcast_to(t_of_sigma(pap / ps), (klev,)).astype(np.float64)
    es = np.minimum(oracle.thermo(params, "foeewm", t) / pap, 0.5)
    arrs["pt"][0] = t
    arrs["pq"][0] = rh * es / (1 - params.retv * es)
    arrs["pa"][0] = a
    arrs["pap"][0] = pap
    arrs["paph"][0] = paph
    for s in range(4):
        arrs["pclv"][0, :, s] = cond[s]
    return arrs


def test_snow_sedimentation_is_implicit_upwind_flux_form(params137):
    p = params137
    klev, k0 = 137, params137.ncldtop - 1
    # ice-saturated, cloud-free, 250 K (pure ice phase), snow only
    arrs = _column(p, klev, lambda s: 250.0 + 0.0 * s, 1.0, 0.0, [0.0, 0.0, 0.0, 1e-5])
    pap, paph, t = arrs["pap"][0].copy(), arrs["paph"][0].copy(), arrs["pt"][0].copy()
    qs0 = arrs["pclv"][0, :, 3].copy()
    oracle.cloudsc_block("dp", p, 1, 1, 1, klev, DT, arrs)
    g, rd = 9.80665, R_DRY  # standard gravity
    v = 1.0  # snow fall speed (m/s) of the default configuration
    flux_in = 0.0
    flux = np.zeros(klev + 1)
    qs_new = np.zeros(klev)
    for k in range(k0, klev):
        dp = paph[k + 1] - paph[k]
        rho = pap[k] / (rd * t[k])
        qs_new[k] = (qs0[k] + DT * g * flux_in / dp) / (1.0 + DT * g * rho * v / dp)
        flux_in = rho * v * qs_new[k]
        flux[k + 1] = flux_in
    np.testing.assert_allclose(arrs["pfplsn"][0], flux, rtol=1e-10, atol=1e-20)
    np.testing.assert_allclose(qs0[k0:] + DT * arrs["ptendency_loc_cld"][0, k0:, 3], qs_new[k0:], rtol=1e-10)
    assert np.all(arrs["pfplsl"][0] == 0.0)
    np.testing.assert_allclose(arrs["ptendency_loc_t"][0], 0.0, atol=1e-15)


def test_warm_rain_autoconversion_follows_kk2000(params137):
    p = params137
    klev, k0 = 137, params137.ncldtop - 1
    ql = 5e-4
    # overcast (no erosion, no clear-sky evaporation), slightly subsaturated (no condensation), warm, no rain
    arrs = _column(p, klev, lambda s: 288.0 + 0.0 * s, 0.99, 1.0, [ql, 0.0, 0.0, 0.0])
    arrs["plsm"][0] = 0.0  # sea: N_c = 50 cm^-3
    oracle.cloudsc_block("dp", p, 1, 1, 1, klev, DT, arrs)
    nc = 50.0
    k_rate = 1.5 * 1350.0 * ql ** (2.47 - 1.0) * nc ** -1.79  # inhomogeneity factor 1.5 (SUCLDP), per second
    ql_new = ql / (1.0 + k_rate * DT)
    np.testing.assert_allclose(ql + DT * arrs["ptendency_loc_cld"][0, k0:, 0], ql_new, rtol=1e-10)
    # the converted water is rain: total water is conserved level by level up to the rain that falls out
    assert arrs["pfplsl"][0, -1] > 0.0 and np.all(arrs["pfplsn"][0] == 0.0)
    assert np.all(arrs["ptendency_loc_cld"][0, :, 1] == 0.0) and np.all(arrs["ptendency_loc_cld"][0, :, 3] == 0.0)


def test_homogeneous_freezing_converts_all_liquid(params137):
    p = params137
    klev, k0 = 137, params137.ncldtop - 1
    ql = 2e-5
    # overcast, just ice-subsaturated, 230 K (below RTHOMO = 235.16 K)
    arrs = _column(p, klev, lambda s: 230.0 + 0.0 * s, 0.999, 1.0, [ql, 0.0, 0.0, 0.0])
    oracle.cloudsc_block("dp", p, 1, 1, 1, klev, DT, arrs)
    tl = arrs["ptendency_loc_cld"][0, k0:, 0]
    np.testing.assert_allclose(ql + DT * tl, 0.0, atol=1e-20)
    lf_over_cp = (2.8345e6 - 2.5008e6) / CP_DRY  # (L_s - L_v) / c_pd
    np.testing.assert_allclose(arrs["ptendency_loc_t"][0, k0:], lf_over_cp * ql / DT, rtol=1e-10)
    np.testing.assert_allclose(arrs["ptendency_loc_q"][0, k0:], 0.0, atol=1e-15)
