"""Whole-column checks of the oracle (FP64): water and enthalpy conservation
(the reference's own LCLDBUDGET check, cloudsc.F90:4146-4203), the reference
write set, closed-form fluxes above the cloud top and special columns."""
import numpy as np
import pytest

import oracle
from openifs_amd import cloudsc as C
from tests.helpers import column_budgets, columns, oracle_blocked

NG, NP, KL = 1024, 32, 137


@pytest.fixture(scope="module")
def run137(params137):
    inp = C.synth_host(params137, 2024, NG, NP, KL, "dp")
    out = oracle_blocked("dp", params137, inp, NG, NP, KL, nthreads=8)
    return inp, out


def test_water_and_enthalpy_conservation(params137, run137):
    inp, out = run137
    (wl, wr, ws), (hl, hr, hs) = column_budgets(params137, out, inp, NP, NG)
    assert np.max(np.abs(wl - wr) / ws) < 1e-10
    assert np.max(np.abs(hl - hr) / hs) < 1e-12
    # the budgets are not trivially zero
    assert np.median(np.abs(wr) / ws) > 1e-3


def test_write_set(params137, run137):
    inp, out = run137
    k0 = params137.ncldtop - 1
    tc = out["ptendency_loc_cld"]
    assert np.all(np.isnan(tc[:, 4]))  # species 5 (vapour) is never written
    assert not np.any(np.isnan(tc[:, :4]))
    cov = out["pcovptot"]
    assert np.all(np.isnan(cov[:, :k0])) and not np.any(np.isnan(cov[:, k0:]))
    for n in C.OUTPUT_NAMES:
        if n not in ("ptendency_loc_cld", "pcovptot"):
            assert not np.any(np.isnan(out[n])), n
    # PLUDE: untouched above NCLDTOP and at KLEV, rescaled or zeroed in between
    np.testing.assert_array_equal(out["plude"][:, :k0], inp["plude"][:, :k0])
    np.testing.assert_array_equal(out["plude"][:, -1], inp["plude"][:, -1])
    changed = out["plude"][:, k0:-1] != inp["plude"][:, k0:-1]
    assert changed.mean() > 0.3


def test_fluxes_closed_forms(params137, run137):
    p = params137
    inp, out = run137
    k0 = p.ncldtop - 1
    np.testing.assert_array_equal(out["pfhpsl"], -p.rlvtt * out["pfplsl"])
    np.testing.assert_array_equal(out["pfhpsn"], -p.rlstt * out["pfplsn"])
    for n in ("pfplsl", "pfplsn"):
        assert np.all(out[n][:, :k0 + 1] == 0.0)
    for n in ("pfsqlf", "pfsqif", "pfsqrf", "pfsqsf", "pfcqlng", "pfcqnng", "pfcqrng", "pfcqsng", "pfsqltur",
              "pfsqitur"):
        assert np.all(out[n][:, 0] == 0.0)
    # above NCLDTOP nothing is solved: the liquid DDH flux integrates -(PCLV + dt*TCML) dp/(g dt) + PLUDELI
    dt = 900.0
    dp = np.diff(inp["paph"], axis=1)
    ql0 = inp["pclv"][:, 0] + dt * inp["ptendency_cml_cld"][:, 0]
    incr = (-ql0 + inp["pvfl"] * dt) * (-dp / p.rg / dt) + inp["pludeli"][:, 0]
    ref = np.concatenate([np.zeros_like(incr[:, :1]), np.cumsum(incr, axis=1)], axis=1)
    np.testing.assert_allclose(out["pfsqlf"][:, :k0 + 1], ref[:, :k0 + 1], rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(out["pfsqltur"][:, -1], np.sum(inp["pvfl"] * dt * (-dp / p.rg / dt), axis=1),
                               rtol=1e-12, atol=1e-22)


def test_per_block_entry_equals_blocked(params137, run137):
    """Block 3 through the CLOUDSC argument list (per-block entry) gives the blocked result.
    A blocked array of block b in C order is the Fortran-ordered block array transposed."""
    inp, out = run137
    b = 3
    arrs = C.allocate_block(NP, KL, "dp")
    for n, kind, role in C.BLOCKED_ARRAYS:
        if role != "o":
            arrs[n][...] = inp[n][b].T
    oracle.cloudsc_block("dp", params137, 1, NP, NP, KL, 900.0, arrs)
    for n, kind, role in C.BLOCKED_ARRAYS:
        if role != "i":
            np.testing.assert_array_equal(arrs[n], out[n][b].T, err_msg=n)


def _column_inputs(params, klev, t_profile, rh, a, cond, ktype=0):
    """One-column block (KLON=1) with prescribed T profile, RH, cloud fraction and condensate."""
    p = params
    a_h, b_h = C.hybrid_levels(klev)
    ps = 100000.0
    paph = a_h + b_h * ps
    pap = C.standard_full_levels(a_h, b_h, ps)
    arrs = C.allocate_block(1, klev, "dp")
    t = t_profile(pap / ps)
    es = np.minimum(oracle.thermo(p, "foeewm", t) / pap, 0.5)
    arrs["pt"][0] = t
    arrs["pq"][0] = rh * es / (1 - p.retv * es)
    arrs["pa"][0] = a
    arrs["pap"][0] = pap
    arrs["paph"][0] = paph
    for s in range(4):
        arrs["pclv"][0, :, s] = cond[s]
    arrs["ktype"][0] = ktype
    arrs["ldcum"][0] = int(ktype > 0)
    return arrs


def test_dry_clear_column_is_inert(params137):
    arrs = _column_inputs(params137, 137, lambda s: np.maximum(216.65, 288.15 * s ** 0.19), 0.3, 0.0,
                          [0.0, 0.0, 0.0, 0.0])
    oracle.cloudsc_block("dp", params137, 1, 1, 1, 137, 900.0, arrs)
    for n in ("ptendency_loc_t", "ptendency_loc_q", "ptendency_loc_a", "pfplsl", "pfplsn", "pfsqlf", "pfsqrf"):
        assert np.all(arrs[n] == 0.0), n
    assert np.all(arrs["ptendency_loc_cld"][:, :, :4] == 0.0)
    assert np.all(arrs["pcovptot"][0, params137.ncldtop - 1:] == 0.0)


def test_warm_column_has_no_frozen_processes(params137):
    # T > 0C everywhere, liquid and rain only: ice and snow stay zero, no snow flux
    arrs = _column_inputs(params137, 137, lambda s: 280.0 + 15.0 * s, 0.95, 0.6, [2e-5, 0.0, 1e-5, 0.0])
    oracle.cloudsc_block("dp", params137, 1, 1, 1, 137, 900.0, arrs)
    assert np.all(arrs["ptendency_loc_cld"][:, :, 1] == 0.0) and np.all(arrs["ptendency_loc_cld"][:, :, 3] == 0.0)
    assert np.all(arrs["pfplsn"] == 0.0)
    assert arrs["pfplsl"][0, -1] > 0.0  # it rains at the surface


def test_supersaturated_warm_column_becomes_overcast(params137):
    p = params137
    arrs = _column_inputs(params137, 137, lambda s: 285.0 + 10.0 * s, 1.04, 0.3, [2e-5, 0.0, 0.0, 0.0])
    oracle.cloudsc_block("dp", p, 1, 1, 1, 137, 900.0, arrs)
    k0 = p.ncldtop - 1
    a_new = arrs["pa"][0] + 900.0 * arrs["ptendency_loc_a"][0]
    np.testing.assert_allclose(a_new[k0:], 1.0, rtol=1e-12)
    assert np.all(arrs["ptendency_loc_q"][0, k0:] < 0) and np.all(arrs["ptendency_loc_t"][0, k0:] > 0)
