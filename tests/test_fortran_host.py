"""The Fortran host path: CLOUDSC_HIP_MOD (ISO_C_BINDING) called from a flang-compiled
program with CLOUDSC's argument list."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from openifs_amd import cloudsc as C
from tests.helpers import relrms

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "cloudsc_fortran_driver")
needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")


def _write_common(d, klon, klev, kidia, kfdia, ptsphy=900.0):
    with open(os.path.join(d, "dims.bin"), "wb") as f:
        f.write(np.array([klon, klev, kidia, kfdia], dtype=np.int32).tobytes())
        f.write(np.array([ptsphy], dtype=np.float64).tobytes())
    a, b = C.hybrid_levels(klev)
    C.standard_full_levels(a, b).astype(np.float64).tofile(os.path.join(d, "stpre.bin"))


@needs_driver
def test_fortran_params_struct_matches_c(tmp_path):
    _write_common(str(tmp_path), 8, 137, 1, 8)
    subprocess.run([DRIVER, "params", str(tmp_path)], check=True)
    raw = open(tmp_path / "params_out.bin", "rb").read()
    ref = C.default_params(900.0, 137)
    assert raw == bytes(ref), "CLOUDSC_PARAMS_T (Fortran) and cloudsc_params_t (C) differ"


def block_inputs(params, prec, klon, klev, seed):
    d = C.synth_host(params, seed, klon, klon, klev, "dp" if prec == "dp" else "sp")
    arrs = C.allocate_block(klon, klev, "dp" if prec == "dp" else "sp")
    for n, kind, role in C.BLOCKED_ARRAYS:
        if role != "o":
            arrs[n][...] = d[n][0].T
    return arrs


def run_driver(tmp_path, arrs, prec, klon, klev, kidia, kfdia):
    d = str(tmp_path)
    _write_common(d, klon, klev, kidia, kfdia)
    for n, kind, role in C.CLOUDSC_ARRAYS:
        if role in ("i", "io", "o") and kind != "scalar":
            np.asfortranarray(arrs[n]).ravel(order="F").tofile(os.path.join(d, f"{n}.bin"))
    subprocess.run([DRIVER, "run", d, prec], check=True, timeout=120)
    out = {}
    for n, kind, role in C.CLOUDSC_ARRAYS:
        if role in ("o", "io"):
            raw = np.fromfile(os.path.join(d, f"{n}.out"), dtype=arrs[n].dtype)
            out[n] = raw.reshape(arrs[n].shape, order="F")
    return out


@needs_driver
@pytest.mark.gpu
def test_fortran_host_fp64_matches_oracle(tmp_path, params137):
    import oracle

    klon, klev = 32, 137
    arrs = block_inputs(params137, "dp", klon, klev, 77)
    out = run_driver(tmp_path, arrs, "dp", klon, klev, 1, klon)
    ref = {k: v.copy(order="F") for k, v in arrs.items()}
    oracle.cloudsc_block("dp", params137, 1, klon, klon, klev, 900.0, ref)
    for n in out:
        np.testing.assert_array_equal(np.isnan(out[n]), np.isnan(ref[n]), err_msg=n)
        assert relrms(out[n], ref[n]) <= 1e-12, n


@needs_driver
@pytest.mark.gpu
def test_fortran_host_fp32_equals_python_host(tmp_path, params137):
    klon, klev = 32, 137
    arrs = block_inputs(params137, "sp", klon, klev, 78)
    out = run_driver(tmp_path, arrs, "sp", klon, klev, 1, klon)
    ref = {k: v.copy(order="F") for k, v in arrs.items()}
    C.CLOUDSC(params137, 1, klon, klon, klev, 900.0, ref, precision="sp")
    for n in out:
        np.testing.assert_array_equal(out[n], ref[n], err_msg=n)


def run_satadj_driver(tmp_path, arrs, prec, klon, klev, kidia, kfdia):
    from openifs_amd import satadj as S

    d = str(tmp_path)
    _write_common(d, klon, klev, kidia, kfdia)
    for n, kind, role in S.SATADJ_ARRAYS:
        if role != "u":
            np.asfortranarray(arrs[n]).ravel(order="F").tofile(os.path.join(d, f"{n}.bin"))
    subprocess.run([DRIVER, "satadj", d, prec], check=True, timeout=120)
    return {n: np.fromfile(os.path.join(d, f"{n}.out"), dtype=arrs[n].dtype).reshape(arrs[n].shape, order="F")
            for n in S.SATADJ_OUTPUTS}


@needs_driver
@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["dp", "sp", "mixed"])
def test_fortran_host_satadj(tmp_path, params137, prec):
    """CLOUD_SATADJ through CLOUDSC_SATADJ_*_C from Fortran equals the Python host bit for bit, and the FP64
    variant equals the oracle to 1e-12."""
    import oracle
    from openifs_amd import satadj as S
    from tests.satadj_cases import block_arrays, blocked_case

    klon, klev = 48, 137
    arrs = block_arrays(blocked_case(params137, 31, 2 * klon, klon, klev, "dp"), 1, prec)
    out = run_satadj_driver(tmp_path, arrs, prec, klon, klev, 2, klon - 1)
    ref = {k: v.copy(order="F") for k, v in arrs.items()}
    S.CLOUD_SATADJ(params137, 2, klon - 1, klon, klev, 900.0, ref, precision=prec)
    for n in out:
        np.testing.assert_array_equal(out[n], ref[n], err_msg=n)
    if prec == "dp":
        orc = {k: v.copy(order="F") for k, v in arrs.items()}
        oracle.satadj_block("dp", params137, 2, klon - 1, klon, klev, 900.0, orc)
        for n in out:
            np.testing.assert_array_equal(np.isnan(out[n]), np.isnan(orc[n]), err_msg=n)
            assert relrms(out[n], orc[n]) <= 1e-12, n


@needs_driver
@pytest.mark.gpu
def test_fortran_block_loop_equals_step_on_fields(tmp_path, params137):
    """From Fortran: the per-block CLOUDSC_SP_C loop over NGPBLKS blocks (the EC_PHYS_DRV pattern) and one
    CLOUDSC_STEP_FIELDS_RUN on field wrappers around the same blocked arrays give identical results, including
    the partial last block."""
    ngptot, nproma, klev = 300, 128, 137
    ngpblks = (ngptot + nproma - 1) // nproma
    d = str(tmp_path)
    _write_common(d, nproma, klev, 1, nproma)
    np.array([ngptot, ngpblks], dtype=np.int32).tofile(os.path.join(d, "blocked.bin"))
    inp = C.synth_host(params137, 91, ngptot, nproma, klev, precision="sp")
    for n, kind, role in C.BLOCKED_ARRAYS:
        shape = C.blocked_shape(kind, nproma, klev, ngpblks)
        dt = np.int32 if kind == "1di" else np.float32
        a = inp[n] if n in inp else np.full(shape, np.nan, dtype=dt)
        np.ascontiguousarray(a, dtype=dt).tofile(os.path.join(d, f"{n}.bin"))
    subprocess.run([DRIVER, "fields", d, "sp"], check=True, timeout=120)
    for n, kind, role in C.BLOCKED_ARRAYS:
        if role in ("o", "io"):
            blk = np.fromfile(os.path.join(d, f"{n}.blk"), dtype=np.float32)
            fld = np.fromfile(os.path.join(d, f"{n}.fld"), dtype=np.float32)
            assert np.array_equal(blk, fld, equal_nan=True), n
            assert np.isfinite(fld).any(), n


def run_cloudsc2_driver(tmp_path, arrs, prec, klon, klev, kidia, kfdia, ktdia, ldrain1d, lphylin, levapls2):
    from openifs_amd import cloudsc2 as C2

    d = str(tmp_path)
    _write_common(d, klon, klev, kidia, kfdia)
    np.array([ktdia, int(ldrain1d), lphylin, levapls2], dtype=np.int32).tofile(os.path.join(d, "c2.bin"))
    C2.reference_eta(klev).astype(np.float64).tofile(os.path.join(d, "ceta.bin"))
    for n, kind, role in C2.CLOUDSC2_ARRAYS:
        np.asfortranarray(arrs[n]).ravel(order="F").tofile(os.path.join(d, f"{n}.bin"))
    subprocess.run([DRIVER, "cloudsc2", d, prec], check=True, timeout=120)
    return {n: np.fromfile(os.path.join(d, f"{n}.out"), dtype=arrs[n].dtype).reshape(arrs[n].shape, order="F")
            for n in C2.OUTPUT_NAMES}


@needs_driver
@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["dp", "sp"])
def test_fortran_host_cloudsc2(tmp_path, params137, prec):
    """CLOUDSC2 through CLOUDSC2_*_C from Fortran (CLOUDSC2_PARAMS_T filled by CLOUDSC2_PARAMS_DEFAULT, then
    LPHYLIN/LEVAPLS2 set) equals the Python host bit for bit; the FP64 variant equals the oracle to 1e-12."""
    import oracle
    from openifs_amd import cloudsc2 as C2

    klon, klev, ktdia = 40, 137, 3
    host = C2.synth_host(params137, 41, klon, klon, klev, prec)
    arrs = {k: np.asfortranarray(v[0].T) for k, v in host.items()}
    out = run_cloudsc2_driver(tmp_path, arrs, prec, klon, klev, 2, klon - 1, ktdia, False, 0, 1)
    q2 = C2.default_params2(params137)
    q2.levapls2 = 1
    ceta = C2.reference_eta(klev)
    ref = {k: v.copy(order="F") for k, v in arrs.items()}
    C2.CLOUDSC2(params137, q2, ceta, 2, klon - 1, klon, ktdia, klev, False, 900.0, ref, prec)
    for n in out:
        np.testing.assert_array_equal(out[n], ref[n], err_msg=n)
    if prec == "dp":
        orc = {k: v.copy(order="F") for k, v in arrs.items()}
        oracle.cloudsc2_block("dp", params137, q2, ceta, 2, klon - 1, klon, ktdia, klev, False, 900.0, orc)
        for n in out:
            np.testing.assert_array_equal(np.isnan(out[n]), np.isnan(orc[n]), err_msg=n)
            assert relrms(out[n], orc[n]) <= 1e-12, n


@needs_driver
@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["dp", "sp"])
def test_fortran_host_cloudst(tmp_path, params137, prec):
    """CLOUDST through CLOUDST_*_C from Fortran equals the Python host bit for bit; FP64 equals the oracle."""
    import oracle
    from openifs_amd import cloudsc2 as C2

    klon, klev, ktdia = 40, 137, 3
    c2 = C2.synth_host(params137, 43, klon, klon, klev, prec)
    dt = np.float64 if prec == "dp" else np.float32
    arrs = {}
    for n, kind, role in C2.CLOUDST_ARRAYS:
        shape = (klon, klev + 1) if kind == "h" else (klon, klev)
        arrs[n] = np.asfortranarray(c2[n][0].T.astype(dt)) if role == "i" else np.full(shape, np.nan, dt, order="F")
    d = str(tmp_path)
    _write_common(d, klon, klev, 2, klon - 1)
    np.array([ktdia, 0, 0, 1], dtype=np.int32).tofile(os.path.join(d, "c2.bin"))
    C2.reference_eta(klev).astype(np.float64).tofile(os.path.join(d, "ceta.bin"))
    for n in arrs:
        arrs[n].ravel(order="F").tofile(os.path.join(d, f"{n}.bin"))
    subprocess.run([DRIVER, "cloudst", d, prec], check=True, timeout=120)
    outs = [n for n, k, r in C2.CLOUDST_ARRAYS if r == "o"]
    out = {n: np.fromfile(os.path.join(d, f"{n}.out"), dtype=dt).reshape(arrs[n].shape, order="F") for n in outs}
    q2 = C2.default_params2(params137)
    q2.levapls2 = 1
    ceta = C2.reference_eta(klev)
    ref = {k: v.copy(order="F") for k, v in arrs.items()}
    C2.CLOUDST(params137, q2, ceta, 2, klon - 1, klon, ktdia, klev, False, 900.0, ref, prec)
    for n in outs:
        np.testing.assert_array_equal(out[n], ref[n], err_msg=n)
    if prec == "dp":
        orc = {k: v.copy(order="F") for k, v in arrs.items()}
        C2.CLOUDST(params137, q2, ceta, 2, klon - 1, klon, ktdia, klev, False, 900.0, orc, "dp",
                   func=oracle.lib().oracle_cloudst_dp)
        for n in outs:
            np.testing.assert_array_equal(np.isnan(out[n]), np.isnan(orc[n]), err_msg=n)
            assert relrms(out[n], orc[n]) <= 1e-12, n
