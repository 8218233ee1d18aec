"""Pins of the oracle's unit-level pieces: CLOUD_SUPERSATCHECK (invariants and
Newton consistency), the no-pivot 5x5 solve (numpy), the explicit-sink
truncation (closed-form and brute-force properties)."""
import numpy as np
import pytest

import oracle


def _state(p, n, rng, tmin, tmax, rhmin, rhmax):
    t = rng.uniform(tmin, tmax, n)
    pp = rng.uniform(20000.0, 100000.0, n)
    es = np.minimum(oracle.thermo(p, "foeewm", t) / pp, 0.5)
    qsmix = es / (1 - p.retv * es)
    q = rng.uniform(rhmin, rhmax, n) * qsmix
    a = rng.uniform(0.0, 1.0, n)
    return t, q, a, pp, np.full(n, 100000.0), qsmix


def test_supersat_conserves_water_and_enthalpy(params137):
    p = params137
    rng = np.random.default_rng(1)
    t, q, a, pp, ps, _ = _state(p, 4000, rng, 200.0, 310.0, 0.5, 1.5)
    ta, qa, aa, la, ia = oracle.supersatcheck(p, t, q, a, pp, ps)
    assert np.any(qa < 0)
    np.testing.assert_allclose(qa + la + ia, 0.0, atol=1e-18)
    np.testing.assert_allclose(p.rcpd * ta, p.rlvtt * la + p.rlstt * ia, rtol=1e-12, atol=1e-15)


def test_supersat_noop_when_subsaturated(params137):
    # q below mixed-phase saturation is never adjusted: the limit is >= qsmix at any cloud fraction
    p = params137
    rng = np.random.default_rng(2)
    t, q, a, pp, ps, _ = _state(p, 4000, rng, 170.0, 310.0, 0.0, 0.999)
    for x in oracle.supersatcheck(p, t, q, a, pp, ps):
        assert np.all(x == 0.0)


def test_supersat_warm_newton_step_is_first_order_exact(params137):
    # T > 0C: the limit is qsmix and the cloud becomes overcast; one linearised step
    # leaves a residual supersaturation of second order only
    p = params137
    rng = np.random.default_rng(3)
    t, q, a, pp, ps, qsmix = _state(p, 2000, rng, 275.0, 305.0, 1.001, 1.05)
    ta, qa, aa, la, ia = oracle.supersatcheck(p, t, q, a, pp, ps)
    np.testing.assert_allclose(a + aa, 1.0, rtol=1e-15)
    np.testing.assert_array_equal(ia, 0.0)
    t2, q2, a2 = t + ta, q + qa, a + aa
    ta2, qa2, aa2, la2, ia2 = oracle.supersatcheck(p, t2, q2, a2, pp, ps)
    ratio = np.abs(qa2) / np.abs(qa)
    assert np.max(ratio) < 0.02
    # a sign error or a dropped factor in the dq_s/dT correction would leave O(1) residuals
    assert np.median(ratio) < 5e-3


def test_supersat_cold_is_all_ice_with_timescale_cover(params137):
    p = params137
    rng = np.random.default_rng(4)
    t, q, a, pp, ps, _ = _state(p, 2000, rng, 190.0, p.rtice - 0.01, 1.3, 1.6)
    ta, qa, aa, la, ia = oracle.supersatcheck(p, t, q, a, pp, ps)
    hit = qa < 0
    assert hit.mean() > 0.5
    np.testing.assert_array_equal(la[hit], 0.0)
    np.testing.assert_allclose(aa[hit], (1 - a[hit]) * p.rssicefactor, rtol=1e-14)


def test_lu5_matches_numpy(params137):
    rng = np.random.default_rng(5)
    for _ in range(200):
        m = rng.normal(0, 1, (5, 5))
        m[np.diag_indices(5)] = np.abs(m).sum(axis=1) + rng.uniform(0.5, 2.0, 5)
        b = rng.normal(0, 1, 5)
        np.testing.assert_allclose(oracle.lu5(m, b), np.linalg.solve(m, b), rtol=1e-12, atol=1e-14)


def test_lu5_lower_triangular_cloudsc_structure():
    # the CLOUDSC default-path matrix: diag and (R,L), (S,L), (S,I) below it
    rng = np.random.default_rng(6)
    for _ in range(100):
        m = np.eye(5)
        m[np.diag_indices(5)] += rng.uniform(0, 3, 5)
        m[2, 0], m[3, 0], m[3, 1] = -rng.uniform(0, 1, 3)
        b = rng.uniform(0, 1e-3, 5)
        np.testing.assert_allclose(oracle.lu5(m, b), np.linalg.solve(m, b), rtol=1e-13)


def _random_qa(rng, scale):
    qa = np.zeros((5, 5))
    pairs = [(4, 0), (4, 1), (4, 2), (4, 3), (1, 0), (2, 1), (2, 3)]  # nonzero pairs of the default path
    for a, b in pairs:
        x = rng.normal(0, scale)
        qa[a, b], qa[b, a] = x, -x
    for d in range(4):
        qa[d, d] = abs(rng.normal(0, scale)) * rng.integers(0, 2)
    return qa


def test_truncation_noop_without_overshoot():
    rng = np.random.default_rng(7)
    for _ in range(500):
        qa = _random_qa(rng, 1e-6)
        qx = np.full(5, 1.0)  # far more than any sink
        np.testing.assert_array_equal(oracle.truncate_sinks(qa, qx), qa)


def test_truncation_single_overshoot_closed_form():
    # rain has 1e-5 and evaporation asks for 3e-5: the sink is cut to the rain present
    qa = np.zeros((5, 5))
    qa[4, 2], qa[2, 4] = 3e-5, -3e-5
    qx = np.array([0.0, 0.0, 1e-5, 0.0, 1e-2])
    out = oracle.truncate_sinks(qa, qx)
    assert out[4, 2] == pytest.approx(1e-5, rel=1e-15) and out[2, 4] == -out[4, 2]
    # a chain: ice melts into rain which evaporates; ice 2e-6 asked 4e-6 of melting
    qa = np.zeros((5, 5))
    qa[2, 1], qa[1, 2] = 4e-6, -4e-6
    qa[4, 2], qa[2, 4] = 5e-6, -5e-6
    qx = np.array([0.0, 2e-6, 1e-6, 0.0, 1e-2])
    out = oracle.truncate_sinks(qa, qx)
    # ice (ratio 0.5) is processed before rain (ratio 1e-6/(5e-6-4e-6) = 1): melt halves to 2e-6;
    # rain's net sink is then 5e-6 - 2e-6 = 3e-6 against 1e-6 present, so its sinks are scaled
    # by 1e-6/3e-6 (the rule scales sinks by amount/net-sink, cloudsc.F90:3517-3519)
    assert out[2, 1] == pytest.approx(2e-6, rel=1e-14)
    assert out[4, 2] == pytest.approx(5e-6 / 3.0, rel=1e-14)
    assert qx[2] + out[2].sum() >= 0.0


def test_truncation_properties_brute_force():
    rng = np.random.default_rng(8)
    for _ in range(2000):
        qa = _random_qa(rng, 1e-5)
        qx = np.where(rng.uniform(size=5) < 0.3, 0.0, rng.lognormal(np.log(1e-5), 1.5, 5))
        out = oracle.truncate_sinks(qa, qx)
        off = ~np.eye(5, dtype=bool)
        np.testing.assert_array_equal((out + out.T)[off], 0.0)  # antisymmetry kept exactly
        ratio = np.divide(out, qa, out=np.ones_like(qa), where=qa != 0)
        assert np.all(ratio <= 1.0 + 1e-15) and np.all(ratio >= 0.0)  # only scaled down
        # sinks of a species are never scaled below the amount that leaves it at -max(qx,ZEPSEC)
        # before later steps; later steps only shrink its sources, so no stronger bound holds
        # (the solver's final clip, cloudsc.F90:3633-3640, removes what remains negative)
