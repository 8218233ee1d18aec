"""Shared helpers of the test suite: oracle runs on synthetic blocked fields and error metrics."""
import numpy as np

import oracle
from openifs_amd import cloudsc as C


def real_dtype(prec):
    return np.float64 if prec == "dp" else np.float32


def oracle_blocked(prec, params, inputs, ngptot, nproma, klev, ptsphy=900.0, nthreads=0):
    """Run the oracle on a copy of ``inputs`` (blocked numpy arrays); outputs start as NaN."""
    dt = real_dtype(prec)
    ngpblks = (ngptot + nproma - 1) // nproma
    arr = {k: (v.astype(dt) if v.dtype.kind == "f" else v.copy()) for k, v in inputs.items()}
    for name, kind, role in C.BLOCKED_ARRAYS:
        if role == "o":
            arr[name] = np.full(C.blocked_shape(kind, nproma, klev, ngpblks), np.nan, dtype=dt)
    oracle.cloudsc_blocked(prec, params, ngptot, nproma, klev, ptsphy, arr, nthreads=nthreads)
    return arr


def columns(x, nproma):
    """(NGPBLKS,[n,]KLEV,NPROMA) blocked array -> (columns, [n,] levels)."""
    if x.ndim == 2:
        return x.reshape(-1)
    if x.ndim == 3:
        return x.transpose(0, 2, 1).reshape(-1, x.shape[1])
    return x.transpose(0, 3, 1, 2).reshape(-1, x.shape[1], x.shape[2])


def relrms(a, r):
    a = np.asarray(a, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    m = ~np.isnan(r)
    den = np.sqrt(np.sum(r[m] ** 2))
    num = np.sqrt(np.sum((a[m] - r[m]) ** 2))
    return num / den if den > 0 else num


def column_budgets(params, out, inp, nproma, ngptot, ptsphy=900.0):
    """Column-integrated water and enthalpy budgets of CLOUDSC outputs (FP64).

    water:    sum_k (tq + sum_m tcld_m) dp/g  =  sources - surface precipitation
    enthalpy: sum_k (cp tT - Lv (tl+tr) - Ls (ti+ts)) dp/g
                = -Lv (src_l+src_r) - Ls (src_i+src_s) + Lv PFPLSL(surf) + Ls PFPLSN(surf)
    sources are the detrained liquid/ice (PLUDELI where the detrainment branch ran,
    i.e. where the rescaled PLUDE is nonzero) and detrained rain/snow (PSNDE), on
    levels NCLDTOP..KLEV-1.  Returns (lhs, rhs, scale) per column for both budgets.
    """
    p = params
    g = lambda n: columns(np.asarray(out[n] if n in out else inp[n], dtype=np.float64), nproma)
    paph = g("paph")
    dp = np.diff(paph, axis=1)
    tq, tt = g("ptendency_loc_q"), g("ptendency_loc_t")
    tc = g("ptendency_loc_cld")[:, :4, :]
    klev = tq.shape[1]
    k0 = p.ncldtop - 1
    lev = np.zeros(klev, dtype=bool)
    lev[k0:klev - 1] = True
    detr = (g("plude") > 0) & lev[None, :]
    lde = g("pludeli")
    snde = g("psnde")
    # Above NCLDTOP only the tidy-up runs: it moves condensate to vapour in PTENDENCY_LOC_Q/T
    # but the condensate tendencies stay zero there (cloudsc.F90:1127-1175 vs :3742), so the
    # budgets are taken over the levels the scheme solves.
    tq, tt, tc = tq[:, k0:], tt[:, k0:], tc[:, :, k0:]
    dp = dp[:, k0:]
    src_l = np.where(detr, lde[:, 0, :], 0.0).sum(axis=1)
    src_i = np.where(detr, lde[:, 1, :], 0.0).sum(axis=1)
    src_r = np.where(lev[None, :], snde[:, 0, :], 0.0).sum(axis=1)
    src_s = np.where(lev[None, :], snde[:, 1, :], 0.0).sum(axis=1)
    pl, pn = g("pfplsl")[:, -1], g("pfplsn")[:, -1]
    w_lhs = ((tq + tc.sum(axis=1)) * dp).sum(axis=1) / p.rg
    w_rhs = src_l + src_i + src_r + src_s - pl - pn
    w_scale = (np.abs(tq * dp).sum(axis=1) + np.abs(tc * dp[:, None, :]).sum(axis=(1, 2))) / p.rg + np.abs(w_rhs) + pl + pn
    h_lhs = ((p.rcpd * tt - p.rlvtt * (tc[:, 0] + tc[:, 2]) - p.rlstt * (tc[:, 1] + tc[:, 3])) * dp).sum(axis=1) / p.rg
    h_rhs = -p.rlvtt * (src_l + src_r) - p.rlstt * (src_i + src_s) + p.rlvtt * pl + p.rlstt * pn
    h_scale = (np.abs(p.rcpd * tt * dp).sum(axis=1) + p.rlstt * np.abs(tc * dp[:, None, :]).sum(axis=(1, 2))) / p.rg
    h_scale = h_scale + np.abs(h_rhs)
    return (w_lhs, w_rhs, w_scale), (h_lhs, h_rhs, h_scale)
