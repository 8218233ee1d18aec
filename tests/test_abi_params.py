"""The C-ABI library loads and exports every declared symbol; the default
configuration equals the reference's setup constants (no GPU needed)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from openifs_amd import cloudsc as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define(.*?\\\n)*.*?\n", "\n", text)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("cloudsc_dp_c", "cloudsc_sp_c", "cloudsc_mixed_c", "cloudsc_blocked_dp_c", "cloudsc_blocked_sp_c",
              "cloudsc_blocked_mixed_c", "cloudsc_params_default", "cloudsc_last_error"):
        assert n in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    lib = C.lib()
    assert hasattr(lib, name), f"{name} declared in include/ but not exported by libcloudsc_hip.so"


def test_params_struct_size():
    # 134 doubles + 8 int32 in include/cloudsc_hip.h
    assert ctypes.sizeof(C.CloudscParams) == 8 * len(C._PARAM_FIELDS) + 4 * len(C._PARAM_INTS)


def test_reference_constants(params137):
    p = params137
    # yomcst.F90:508-591, suphec.F90:231-254
    r = 6.0221367e23 * 1.380658e-23
    rd, rv = 1000.0 * r / 28.9644, 1000.0 * r / 18.0153
    expect = dict(rg=9.80665, rd=rd, rv=rv, rcpd=3.5 * rd, retv=rv / rd - 1.0, rtt=273.16, rlvtt=2.5008e6,
                  rlstt=2.8345e6, rlmlt=2.8345e6 - 2.5008e6, ra=6371229.0, rpi=math.pi, r2es=611.21 * rd / rv,
                  r3les=17.502, r3ies=22.587, r4les=32.19, r4ies=-0.7, r5les=17.502 * (273.16 - 32.19),
                  r5ies=22.587 * (273.16 + 0.7), rtice=273.16 - 23.0, rtwat=273.16, rtwat_rtice_r=1.0 / 23.0,
                  rkoop1=2.583, rkoop2=0.48116e-2)
    expect["r5alvcp"] = expect["r5les"] * 2.5008e6 / (3.5 * rd)
    expect["r5alscp"] = expect["r5ies"] * 2.8345e6 / (3.5 * rd)
    expect["ralvdcp"] = 2.5008e6 / (3.5 * rd)
    expect["ralsdcp"] = 2.8345e6 / (3.5 * rd)
    # sucldp.F90 defaults (RPLRG = RPLDARE = 1)
    expect.update(ramid=0.8, rprc1=100.0, ramin=1e-8, rlmin=1e-8, rcldiff=6e-6, rcldiff_convi=10.0,
                  rclcrit_sea=2.5e-4, rclcrit_land=5.5e-4, rkconv=1 / 6000.0, rcl_inhomogaut=1.5,
                  rcl_inhomogacc=3.0, rsnowlin1=1e-3, rsnowlin2=0.03, rlcritsnow=2e-5, rvice=0.13, rvrain=4.0,
                  rvsnow=1.0, rxfallspeedrain=1.0, rxfallspeedice=1.0, rnice=0.027, rccn=125.0, rcldtopcf=0.01,
                  rdepliqrefrate=0.5, rdepliqrefdepth=500.0, rcl_overlapliqice=0.65, rcl_effrime=1.0,
                  riceinit=1e-12, rthomo=273.16 - 38.0, rprecrhmax=0.7, rpecons=5.44e-4 / 9.80665,
                  rvrfactor=5.09e-3, rcovpmin=0.1, rtaumel=7200.0, rkooptau=10800.0, rssicefactor=900.0 / 10800.0,
                  rdenswat=1000.0, rdensref=1.0, rcl_kkaac=67.0, rcl_kkbac=1.15, rcl_kkaau=1350.0,
                  rcl_kkbauq=2.47, rcl_kkbaun=-1.79, rcl_kk_cloud_num_sea=50.0, rcl_kk_cloud_num_land=300.0,
                  rcl_fzrab=-0.66, rcl_apb1=7.14e11, rcl_apb2=1.16e8, rcl_apb3=2.416e2, rcl_cdenom1=5.57e11,
                  rcl_cdenom2=1.03e8, rcl_cdenom3=2.04e2, rcl_ka273=2.4e-2, rcl_eff_racw=0.7, rmfadvw=0.0,
                  rmfadvwdd=0.0, rcloud_frac_std=1.0, reisthsc=5.0)
    for k, v in expect.items():
        assert getattr(p, k) == pytest.approx(v, rel=1e-14, abs=0.0), k
    assert p.ncldtop == 15 and p.lmfdsnow == 1
    assert p.lcloud_inhomog == p.lrad_cloud_inhomog == p.laeraware == p.lspp == p.lcldbud == 0


def test_psd_constants_against_math_gamma(params137):
    """PSD constants of sucldp.F90:362-530 with the exact Gamma function.  The
    reference's Lanczos series carries single-precision coefficients, so the
    constants agree to ~1e-7 relative, not to double precision."""
    p = params137
    G = math.gamma
    ar, br, x1r, x2r, cr, dr = 1000.0 * math.pi / 6.0, 3.0, 0.22, 2.2, 386.8, 0.67
    ai = as_ = 0.069
    bi = bs = 2.0
    ci, di, cs, ds = 100.0, 1.0, 16.8, 0.527
    x1 = 2.0e6
    expect = dict(
        rcl_lam1r=ar * x1r * G(br + 1), rcl_lam2r=1 / (br + 1 - x2r),
        rcl_const1r=2 * math.pi * x1r,
        rcl_const2r=0.31 * 0.6 ** 0.3333 * G((dr + 5) / 2) * (cr / 1.717e-5) ** 0.5,
        rcl_const3r=(dr + 5) / 2 - x2r, rcl_const4r=2 - x2r, rcl_const7r=cr * G(dr + br + 1) / G(br + 1),
        rcl_const9r=math.pi * 0.25 * x1r * cr * G(3 + dr), rcl_const10r=3 + dr - x2r,
        rcl_const5r=math.pi / 6 * 200.0 * ar * x1r * G(4 + br), rcl_const6r=-(4 + br - x2r),
        rcl_const1i=1 / (ai * x1 * G(bi + 1)), rcl_const2i=math.pi * x1,
        rcl_const3i=0.44 * 0.6 ** (1 / 3) * ci ** 0.5 * G((di + 5) / 2) / 1.717e-5 ** 0.5,
        rcl_const4i=2 / (bi + 1), rcl_const5i=(5 + di) / (2 * (bi + 1)), rcl_const6i=G(2.0),
        rcl_const7i=ci * G(di + bi + 1) / G(bi + 1),
        rcl_lam1s=as_ * x1 * G(bs + 1), rcl_lam2s=1 / (bs + 1), rcl_const1s=1 / (as_ * x1 * G(bs + 1)),
        rcl_const2s=math.pi * x1, rcl_const3s=0.44 * 0.6 ** (1 / 3) * cs ** 0.5 * G((ds + 5) / 2) / 1.717e-5 ** 0.5,
        rcl_const4s=2 / (bs + 1), rcl_const5s=(5 + ds) / (2 * (bs + 1)), rcl_const6s=G(2.0),
        rcl_const7s=math.pi * 0.25 * x1 * cs * G(3 + ds), rcl_const8s=(3 + ds) / (bs + 1),
        rcl_dr=dr, rcl_x1r=x1r, rcl_x2r=x2r, rcl_x1s=x1)
    for k, v in expect.items():
        assert getattr(p, k) == pytest.approx(v, rel=3e-7), k


def test_ncldtop_levels(params91):
    # L137 standard atmosphere: level 14 ~ 97 Pa, level 15 ~ 119 Pa > RCLDTOPP = 100 Pa
    a, b = C.hybrid_levels(137)
    pf = C.standard_full_levels(a, b)
    assert pf[13] < 100.0 < pf[14]
    assert params91.ncldtop == 10


def test_levels_tables():
    for klev in (137, 91):
        a, b = C.hybrid_levels(klev)
        assert a.size == b.size == klev + 1
        assert a[0] == 0 and b[0] == 0 and a[-1] == 0 and b[-1] == 1
        for ps in (95000.0, 101325.0, 105000.0):
            assert np.all(np.diff(a + b * ps) > 0)


def test_params_default_rejects_bad_input():
    p = C.CloudscParams()
    st = np.ones(10)
    assert C.lib().cloudsc_params_default(ctypes.byref(p), 0.0, 10, st.ctypes.data) != 0
    assert C.lib().cloudsc_params_default(ctypes.byref(p), 900.0, 1, st.ctypes.data) != 0
