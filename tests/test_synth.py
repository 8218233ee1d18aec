"""Synthetic atmospheres: Philox4x32-10 known answers, blocking independence,
and the input distributions of the benchmark configurations."""
import ctypes

import numpy as np
import pytest

from openifs_amd import cloudsc as C
from tests.helpers import columns


@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, expect):
    # Random123 kat_vectors (Salmon et al., SC'11)
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    C.lib().cloudsc_philox4x32_10(c, k, o)
    assert tuple(o) == expect


def test_blocking_independence(params137):
    a = C.synth_host(params137, 99, 300, 32, 137, "sp")
    b = C.synth_host(params137, 99, 300, 128, 137, "sp")
    for n in a:
        ca, cb = columns(a[n], 32)[:300], columns(b[n], 128)[:300]
        np.testing.assert_array_equal(ca, cb, err_msg=n)


def test_seed_changes_fields(params137):
    a = C.synth_host(params137, 1, 64, 32, 137, "dp")
    b = C.synth_host(params137, 2, 64, 32, 137, "dp")
    assert not np.array_equal(a["pt"], b["pt"])


def test_distributions(params137):
    p = params137
    d = C.synth_host(p, 5, 4096, 64, 137, "dp")
    g = lambda n: columns(d[n], 64)
    ps = g("paph")[:, -1]
    assert abs(ps.mean() - 100000.0) < 100.0 and abs(ps.std() - 1500.0) < 100.0
    t = g("pt")
    assert 200.0 < t.min() and t.max() < 310.0
    tt = t - 216.65
    assert np.all(np.abs(tt[:, :10]) < 12.0)  # isothermal stratosphere plus N(0, 2 K)
    pap = g("pap")
    q = g("pq")
    a = g("pa")
    import oracle
    esm = np.minimum(oracle.thermo(p, "foeewm", t.ravel()).reshape(t.shape) / pap, 0.5)
    rh = q / (esm / (1 - p.retv * esm))
    assert rh.min() >= 0.3 - 1e-9 and rh.max() <= 1.05 + 1e-9
    assert np.all(a[rh <= 0.8] == 0.0) and np.all((a >= 0) & (a < 1))
    clv = g("pclv")
    assert np.all(clv[:, 1][t >= 273.0] == 0) and np.all(clv[:, 3][t >= 273.0] == 0)
    ql = clv[:, 0]
    assert abs(np.median(ql) / 1e-5 - 1) < 0.05 and abs(np.log(ql).std() - 1.5) < 0.05
    kt = g("ktype")
    frac = np.bincount(kt, minlength=4) / kt.size
    np.testing.assert_allclose(frac, [0.5, 0.25, 0.15, 0.10], atol=0.03)
    assert np.all(g("ldcum") == (kt > 0))
    mf = g("pmfu")
    assert mf.min() >= 0 and mf.max() <= 0.05
    lude = g("plude")
    assert np.all(lude[kt == 0] == 0)
    ldl = g("pludeli")
    np.testing.assert_allclose(ldl[:, 0] + ldl[:, 1], lude, rtol=1e-14)
