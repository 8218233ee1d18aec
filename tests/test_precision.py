"""The precision policies of the oracle against its FP64 build, on identical FP32-representable inputs (CPU).

* sp (the ifs_sp build) is ~1e-2 away from FP64: a sink limited by the available amount (rain evaporation,
  melting, freezing, sink truncation) leaves a residue of FP32 size (~1e-13) in the solve's right-hand side, above
  the reference's ZEPSEC = 1E-14 clip, so the species survives where FP64 clears it (and ZCOVPTOT with it).
* mixed keeps the level's water budget (ZQXFG, ZSOLQA, the limited sinks, the truncation), the solve and
  CLOUD_SUPERSATCHECK in FP64: within 1e-4 relative RMS of FP64 on every output but the cloud-cover tendency,
  whose supersaturation-check ties (a humidity within a few 1e-7 of its limit) flip at ~1e-5 of the points.
The checks below fail if the FP64 budget is taken out of the mixed policy (the ~1e-2 of sp comes back).
"""
import numpy as np
import pytest

from openifs_amd import cloudsc as C
from tests.helpers import columns, oracle_blocked, relrms

FIELDS = ["ptendency_loc_t", "ptendency_loc_q", "ptendency_loc_cld", "pcovptot", "plude", "prainfrac_toprfz",
          "pfsqlf", "pfsqif", "pfcqnng", "pfcqlng", "pfsqrf", "pfsqsf", "pfcqrng", "pfcqsng", "pfsqltur",
          "pfsqitur", "pfplsl", "pfplsn", "pfhpsl", "pfhpsn"]


@pytest.fixture(scope="module", params=[137, 91])
def runs(request):
    klev = request.param
    ngptot, nproma = 2048, 64
    p = C.default_params(900.0, klev)
    inp = C.synth_host(p, 51, ngptot, nproma, klev, "sp")  # FP32 inputs, read exactly by every build
    return {prec: oracle_blocked(prec, p, inp, ngptot, nproma, klev, nthreads=8) for prec in ("dp", "sp", "mixed")}


def test_mixed_within_1e4_of_fp64(runs):
    r64, rm = runs["dp"], runs["mixed"]
    for n in FIELDS:
        e = relrms(columns(rm[n], 64), columns(r64[n], 64))
        assert e <= 1e-4, f"{n}: {e:.3e}"
    surf = lambda r: columns(r["pfplsl"], 64)[:, -1].astype(np.float64) + columns(r["pfplsn"], 64)[:, -1]
    assert relrms(surf(rm), surf(r64)) <= 1e-4


def test_mixed_cloud_cover_ties_are_rare(runs):
    g = columns(runs["mixed"]["ptendency_loc_a"], 64).astype(np.float64)
    r = columns(runs["dp"]["ptendency_loc_a"], 64)
    flip = np.abs(g - r) > 1e-4 * np.abs(r).max()
    assert flip.mean() <= 1e-4
    assert relrms(np.where(flip, r, g), r) <= 1e-4
    # a tie moves the cloud cover by a large part of the step's range, not by rounding
    if flip.any():
        assert np.abs(g - r)[flip].min() > 0.05 / 900.0


def test_sp_is_far_from_fp64_through_the_clip(runs):
    """The FP32 build's distance to FP64 is the clip mechanism, not accumulation: precipitation fraction and
    the rain flux are off by percent while the FP64-budget policy is not."""
    r64 = runs["dp"]
    e_sp = relrms(columns(runs["sp"]["pcovptot"], 64), columns(r64["pcovptot"], 64))
    e_mx = relrms(columns(runs["mixed"]["pcovptot"], 64), columns(r64["pcovptot"], 64))
    assert e_sp > 1e-3 and e_mx < 1e-5
