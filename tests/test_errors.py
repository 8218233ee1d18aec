"""Error paths of the C-ABI that are decided on the host, before any device call."""
import ctypes

import numpy as np
import pytest

from openifs_amd import cloudsc as C


def _blocked_null_call(params, ngptot=64, nproma=32, klev=137, ngpblks=2):
    f = C.lib().cloudsc_blocked_sp_c
    args = [ctypes.byref(params), ctypes.c_int(ngptot), ctypes.c_int(nproma), ctypes.c_int(klev),
            ctypes.c_int(ngpblks), ctypes.c_float(900.0)]
    args += [ctypes.c_void_p(0)] * len(C.BLOCKED_ARRAYS)
    args.append(ctypes.c_void_p(0))
    return f(*args)


def test_unsupported_switches_are_rejected():
    for sw in ("laeraware", "lspp"):
        p = C.default_params()
        setattr(p, sw, 1)
        assert _blocked_null_call(p) == -2  # CLOUDSC_EUNSUPPORTED
        assert b"not implemented" in C.lib().cloudsc_last_error()
    # LCLOUD_INHOMOG reads PGAW, which the plain blocked entry does not take
    p = C.default_params()
    p.lcloud_inhomog = 1
    assert _blocked_null_call(p) == -2
    assert b"PGAW" in C.lib().cloudsc_last_error()
    # LRAD_CLOUD_INHOMOG writes PFSD, which the plain blocked entry does not take
    p = C.default_params()
    p.lrad_cloud_inhomog = 1
    assert _blocked_null_call(p) == -2
    assert b"PFSD" in C.lib().cloudsc_last_error()
    # the budgets that write PEXTRA need an entry that takes PEXTRA
    for bits in (C.BUD_C, C.BUD_L, C.BUD_I, C.BUD_T, C.BUD_VERTINT):
        p = C.default_params()
        p.lcldbud = bits
        assert _blocked_null_call(p) == -2
        assert b"PEXTRA" in C.lib().cloudsc_last_error()
    p = C.default_params()
    p.lcldbud = 256
    assert _blocked_null_call(p) == -2
    # TIMEINT, LBUD23 and LCLDBUDGET alone write nothing: accepted (the NULL arrays are what fails)
    for bits in (C.BUD_TIMEINT, C.BUD_LBUD23, C.BUD_CHECK):
        p = C.default_params()
        p.lcldbud = bits
        assert _blocked_null_call(p) == -3


def test_bad_dimensions_are_rejected():
    p = C.default_params()
    assert _blocked_null_call(p, ngptot=64, nproma=32, ngpblks=3) == -1
    assert _blocked_null_call(p, klev=1) == -1
    p.ncldtop = 1
    assert _blocked_null_call(p) == -1


def test_null_arrays_are_rejected():
    p = C.default_params()
    assert _blocked_null_call(p) == -3  # CLOUDSC_ENULL


def test_per_block_window_checks():
    p = C.default_params()
    arrs = C.allocate_block(8, 137, "sp")
    for kidia, kfdia in ((0, 8), (5, 4), (1, 9), (3, 7)):  # KIDIA+KFDIA-1 must stay within KLON
        with pytest.raises(C.CloudscError) as e:
            C.CLOUDSC(p, kidia, kfdia, 8, 137, 900.0, arrs, precision="sp")
        assert e.value.code == -1


def test_python_layer_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(C, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(C, "_lib", None)
    with pytest.raises(C.CloudscLibraryError):
        C.lib()


def _satadj_blocked_null_call(params, ngptot=64, nproma=32, klev=137, ngpblks=2):
    from openifs_amd import satadj as S

    f = C.lib().cloudsc_satadj_blocked_sp_c
    args = [ctypes.byref(params), ctypes.c_int(ngptot), ctypes.c_int(nproma), ctypes.c_int(klev),
            ctypes.c_int(ngpblks), ctypes.c_float(900.0)]
    args += [ctypes.c_void_p(0)] * len(S.SATADJ_BLOCKED)
    args.append(ctypes.c_void_p(0))
    return f(*args)


def test_satadj_errors():
    from openifs_amd import satadj as S

    for sw in ("lspp", "lcldbud"):
        p = C.default_params()
        setattr(p, sw, 1)
        assert _satadj_blocked_null_call(p) == -2
        assert b"not implemented" in C.lib().cloudsc_last_error()
    p = C.default_params()
    assert _satadj_blocked_null_call(p, ngpblks=3) == -1
    assert _satadj_blocked_null_call(p, klev=1) == -1
    assert _satadj_blocked_null_call(p) == -3
    arrs = S.allocate_block(8, 137, "sp")
    for kidia, kfdia in ((0, 8), (5, 4), (1, 9)):
        with pytest.raises(C.CloudscError) as e:
            S.CLOUD_SATADJ(p, kidia, kfdia, 8, 137, 900.0, arrs, precision="sp")
        assert e.value.code == -1
