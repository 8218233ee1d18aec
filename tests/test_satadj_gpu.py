"""GPU parity of CLOUD_SATADJ: the gfx950 kernel against the CPU oracle on identical inputs.

FP64: relative RMS <= 1e-12 per field and the same write set.  FP32 and mixed: every point is independent (no
level-to-level recurrence), so the only FP32 departures are rounding and the points that rounding moves across
one of the scheme's switches (RLMIN, ZEPSEC, RAMIN, RTHOMO); the criterion is the FP32 oracle's own distance to
FP64, as for CLOUDSC: GPU relative RMS against the FP64 oracle <= 2 x the FP32 oracle's + 1e-6.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle
from openifs_amd import cloudsc as C
from openifs_amd import satadj as S
from tests.helpers import relrms
from tests.satadj_cases import block_arrays, blocked_case

pytestmark = pytest.mark.gpu

OUTS = ["ptendency_loc_t", "ptendency_loc_q", "ptendency_loc_a", "ptendency_loc_cld"]
INS = [n for n, k, r in S.SATADJ_BLOCKED if r == "i"]


def gpu_run(params, prec, ngptot, nproma, klev, seed, ptsphy=900.0):
    f = S.SatadjFields(ngptot, nproma, klev, prec, "cuda", fill=float("nan"))
    S.synth_device(f, params, seed)
    inp = {n: f[n].cpu().numpy().copy() for n in INS}
    S.satadj_blocked(f, params, ptsphy)
    torch.cuda.synchronize()
    out = {n: f[n].cpu().numpy() for n in OUTS}
    del f
    return inp, out


def oracle_run(prec, params, inp, ngptot, nproma, klev, ptsphy=900.0):
    dt = np.float64 if prec == "dp" else np.float32
    arr = {k: v.astype(dt) for k, v in inp.items()}
    ngpblks = (ngptot + nproma - 1) // nproma
    for n, kind, role in S.SATADJ_BLOCKED:
        if role == "o":
            arr[n] = np.full(C.blocked_shape(kind, nproma, klev, ngpblks), np.nan, dtype=dt)
    oracle.satadj_blocked(prec, params, ngptot, nproma, klev, ptsphy, arr, nthreads=16)
    return arr


def fields(out, ngptot, nproma):
    """(field name, columns x levels array) for the five tendencies, real columns only."""
    col = lambda x: x.transpose(0, 2, 1).reshape(-1, x.shape[1])[:ngptot]
    res = [(n, col(out[n])) for n in OUTS[:3]]
    res += [("ptendency_loc_l", col(out["ptendency_loc_cld"][:, 0])), ("ptendency_loc_i", col(out["ptendency_loc_cld"][:, 1]))]
    return res


@pytest.mark.parametrize("ngptot,nproma,klev,seed", [(2048, 64, 137, 1), (1000, 128, 137, 2), (777, 32, 91, 3)])
def test_fp64_matches_oracle(ngptot, nproma, klev, seed, params137, params91):
    p = params137 if klev == 137 else params91
    inp, out = gpu_run(p, "dp", ngptot, nproma, klev, seed)
    ref = oracle_run("dp", p, inp, ngptot, nproma, klev)
    for (n, g), (_, r) in zip(fields(out, ngptot, nproma), fields(ref, ngptot, nproma)):
        np.testing.assert_array_equal(np.isnan(g), np.isnan(r), err_msg=n)
        assert relrms(g, r) <= 1e-12, n
    # species 3..5 and the padding columns of the last block are not written
    assert np.all(np.isnan(out["ptendency_loc_cld"][:, 2:]))
    if ngptot % nproma:
        pad = out["ptendency_loc_t"][-1][:, ngptot % nproma:]
        assert np.all(np.isnan(pad))


@pytest.mark.parametrize("prec,klev", [("sp", 137), ("mixed", 137), ("sp", 91), ("mixed", 91)])
def test_fp32_as_accurate_as_fp32_oracle(prec, klev, params137, params91):
    p = params137 if klev == 137 else params91
    ngptot, nproma = 8192, 128
    inp, out = gpu_run(p, prec, ngptot, nproma, klev, 5)
    r64 = oracle_run("dp", p, inp, ngptot, nproma, klev)
    r32 = oracle_run(prec, p, inp, ngptot, nproma, klev)
    report = []
    for (n, g), (_, a), (_, b) in zip(*(fields(x, ngptot, nproma) for x in (out, r64, r32))):
        np.testing.assert_array_equal(np.isnan(g), np.isnan(a), err_msg=n)
        e_gpu, e_cpu = relrms(g, a), relrms(b, a)
        report.append((n, e_gpu, e_cpu))
        assert e_gpu <= 2.0 * e_cpu + 1e-6, f"{n}: GPU {e_gpu:.3e} vs FP32 oracle {e_cpu:.3e}"
    print("\n".join(f"{n:18s} gpu {a:.3e}  fp32-oracle {b:.3e}" for n, a, b in report))


@pytest.mark.parametrize("prec", ["dp", "sp", "mixed"])
def test_per_block_drop_in(prec, params137):
    """cloudsc_satadj_*_c (host arrays, one block, KIDIA:KFDIA window) equals the oracle of the same precision
    and leaves columns outside the window untouched."""
    p = params137
    arr = blocked_case(p, 9, 96, 48, 137, "dp")
    a_gpu, a_ref = block_arrays(arr, 1, prec), block_arrays(arr, 1, prec)
    S.CLOUD_SATADJ(p, 3, 45, 48, 137, 900.0, a_gpu, prec)
    oracle.satadj_block(prec, p, 3, 45, 48, 137, 900.0, a_ref)
    for n in S.SATADJ_OUTPUTS:
        g, r = a_gpu[n], a_ref[n]
        np.testing.assert_array_equal(np.isnan(g), np.isnan(r), err_msg=n)
        assert np.all(np.isnan(g[:2])) and np.all(np.isnan(g[45:])), n
        tol = 1e-12 if prec == "dp" else 1e-4
        assert relrms(g, r) <= tol, n


@pytest.mark.parametrize("prec", ["sp", "mixed"])
def test_full_size_properties(prec, params137):
    """655,360 columns at L137 (the benchmark workload): finite outputs on the written set, zeros above NCLDTOP,
    per-point water and enthalpy conservation to FP32 rounding of the process terms, untouched species 3..5, and a sample of blocks
    equal to the oracle by the FP32 criterion."""
    p = params137
    ngptot, nproma, klev = 655360, 128, 137
    f = S.SatadjFields(ngptot, nproma, klev, prec, "cuda", fill=float("nan"))
    S.synth_device(f, p, 77)
    S.satadj_blocked(f, p, 900.0)
    torch.cuda.synchronize()
    k0 = p.ncldtop - 1
    tt, tq, ta = f["ptendency_loc_t"].double(), f["ptendency_loc_q"].double(), f["ptendency_loc_a"].double()
    tl, ti = f["ptendency_loc_cld"][:, 0].double(), f["ptendency_loc_cld"][:, 1].double()
    for x in (tt, tq, ta, tl, ti):
        assert torch.isfinite(x).all()
        assert (x[:, :k0] == 0).all()
    assert torch.isnan(f["ptendency_loc_cld"][:, 2:]).all()
    w = (tq + tl + ti).abs() / (tq.abs() + tl.abs() + ti.abs() + 1e-30)
    h = (tt - p.ralvdcp * tl - p.ralsdcp * ti).abs() / (tt.abs() + p.ralsdcp * (tl.abs() + ti.abs()) + 1e-30)
    # each tendency is one FP32 sum of up to eight process terms, so a result that nearly cancels carries the
    # rounding of its largest term: the bound is on the worst point and on the mean
    assert float(w.max()) < 1e-3 and float(h.max()) < 1e-3
    assert float(w.mean()) < 1e-7 and float(h.mean()) < 1e-7
    # sampled parity: four blocks spread over the field
    blocks = [0, 1717, 3333, 5119]
    inp = {n: f[n][blocks].cpu().numpy().copy() for n in INS}
    out = {n: f[n][blocks].cpu().numpy() for n in OUTS}
    del f, tt, tq, ta, tl, ti, w, h
    torch.cuda.empty_cache()
    n = len(blocks) * nproma
    r64, r32 = oracle_run("dp", p, inp, n, nproma, klev), oracle_run(prec, p, inp, n, nproma, klev)
    for (name, g), (_, a), (_, b) in zip(*(fields(x, n, nproma) for x in (out, r64, r32))):
        assert relrms(g, a) <= 2.0 * relrms(b, a) + 1e-6, name


def test_synth_forcings(params137):
    """cloudsc_satadj_synth_*: deterministic, the documented moments, PLUDELI(:,:,3:4) tied to PLUDE."""
    p = params137
    f1 = S.SatadjFields(4096, 64, 137, "sp", "cuda")
    f2 = S.SatadjFields(4096, 128, 137, "sp", "cuda")
    S.synth_device(f1, p, 5)
    S.synth_device(f2, p, 5)
    for n in S.SATADJ_SYNTH:
        a = f1[n].permute(0, 2, 1).reshape(-1, 137)
        b = f2[n].permute(0, 2, 1).reshape(-1, 137)
        assert torch.equal(a, b), n  # the same columns whatever the blocking
    x = {n: f1[n].double() for n in S.SATADJ_SYNTH}
    assert 0.0 <= float(x["phrsw"].min()) and float(x["phrsw"].max()) <= 2e-5
    assert abs(float(x["phrsw"].mean()) - 1e-5) < 2e-7
    assert abs(float(x["phrlw"].mean()) + 2e-5) < 2e-7 and abs(float(x["phrlw"].std()) - 1e-5) < 2e-7
    assert abs(float(x["ptendency_vdf_t"].std()) - 1e-5) < 2e-7
    assert abs(float(x["ptendency_vdf_q"].std()) - 1e-8) < 2e-10
    assert abs(float(x["pvervel"].mean())) < 5e-3 and abs(float(x["pvervel"].std()) - 0.3) < 5e-3
    lde, lude = f1["pludeli"].double(), f1["plude"].double()
    torch.testing.assert_close(lde[:, 2], 1e-2 * lude, rtol=1e-6, atol=0)
    torch.testing.assert_close(lde[:, 3], 1e3 * lude, rtol=1e-6, atol=0)


# ---------------------------------------------------------------------------
# fused CALLPAR sequence (one launch) against the unfused sequence and the oracle
# ---------------------------------------------------------------------------
CML = ["ptendency_cml_t", "ptendency_cml_q", "ptendency_cml_a", "ptendency_cml_cld"]
SATOUT = ["ptendency_satadj_t", "ptendency_satadj_q", "ptendency_satadj_a", "ptendency_satadj_cld"]


def unfused_sequence(f, params, ptsphy=900.0):
    """CLOUD_SATADJ, STATE_INCREMENT of the CML tendencies (callpar.F90:1278-1279), STATE_COPY to
    TENDENCY_SATADJ (:1289) and CLOUDSC as separate steps on a FusedFields (the satadj tendencies go through
    TENDENCY_LOC, as in CALLPAR)."""
    S.satadj_blocked(f, params, ptsphy)
    torch.cuda.synchronize()
    sat = {"ptendency_satadj_t": f["ptendency_loc_t"].clone(), "ptendency_satadj_q": f["ptendency_loc_q"].clone(),
           "ptendency_satadj_a": f["ptendency_loc_a"].clone(),
           "ptendency_satadj_cld": f["ptendency_loc_cld"][:, :2].clone()}
    f["ptendency_cml_t"].add_(sat["ptendency_satadj_t"])
    f["ptendency_cml_q"].add_(sat["ptendency_satadj_q"])
    f["ptendency_cml_a"].add_(sat["ptendency_satadj_a"])
    f["ptendency_cml_cld"][:, :2].add_(sat["ptendency_satadj_cld"])
    C.cloudsc_blocked(f, params, ptsphy)
    torch.cuda.synchronize()
    return sat


@pytest.mark.parametrize("prec,ngptot,nproma,klev", [("dp", 2048, 64, 137), ("sp", 8192, 128, 137),
                                                     ("mixed", 8192, 128, 137), ("sp", 4096, 64, 91)])
def test_fused_equals_unfused_sequence(prec, ngptot, nproma, klev, params137, params91):
    """One fused launch reproduces CLOUD_SATADJ + increment + copy + CLOUDSC bit for bit: every CLOUDSC output,
    the incremented CML tendencies and TENDENCY_SATADJ."""
    p = params137 if klev == 137 else params91
    f = S.FusedFields(ngptot, nproma, klev, prec, "cuda", fill=float("nan"))
    S.synth_device(f, p, 41)
    g = S.FusedFields(ngptot, nproma, klev, prec, "cuda", fill=float("nan"))
    for n in f.t:
        g.t[n].copy_(f.t[n])
    S.satadj_cloudsc_blocked(f, p, 900.0)
    torch.cuda.synchronize()
    sat = unfused_sequence(g, p)
    for n in C.OUTPUT_NAMES + CML[:3]:
        assert torch.equal(torch.nan_to_num(f[n], nan=7.0), torch.nan_to_num(g[n], nan=7.0)), n
    assert torch.equal(f["ptendency_cml_cld"], g["ptendency_cml_cld"])
    for n in SATOUT[:3]:
        assert torch.equal(f[n], sat[n]), n
    assert torch.equal(f["ptendency_satadj_cld"][:, :2], sat["ptendency_satadj_cld"])
    assert torch.isnan(f["ptendency_satadj_cld"][:, 2:]).all()


def test_fused_fp64_matches_oracle_sequence(params137):
    """The fused FP64 launch against the oracle's CLOUD_SATADJ, the increment and the oracle's CLOUDSC."""
    from tests.helpers import columns, oracle_blocked

    p, ngptot, nproma, klev = params137, 1000, 64, 137
    f = S.FusedFields(ngptot, nproma, klev, "dp", "cuda", fill=float("nan"))
    S.synth_device(f, p, 43)
    inp = {n: f[n].cpu().numpy().copy() for n in f.t}
    S.satadj_cloudsc_blocked(f, p, 900.0)
    torch.cuda.synchronize()
    sref = oracle_run("dp", p, {n: inp[n] for n in INS}, ngptot, nproma, klev)
    cin = {n: inp[n].copy() for n in C.INPUT_NAMES}
    cin["ptendency_cml_t"] += np.nan_to_num(sref["ptendency_loc_t"])
    cin["ptendency_cml_q"] += np.nan_to_num(sref["ptendency_loc_q"])
    cin["ptendency_cml_a"] += np.nan_to_num(sref["ptendency_loc_a"])
    cin["ptendency_cml_cld"][:, :2] += np.nan_to_num(sref["ptendency_loc_cld"][:, :2])
    ref = oracle_blocked("dp", p, cin, ngptot, nproma, klev, nthreads=16)
    for n in C.OUTPUT_NAMES:
        g, r = columns(f[n].cpu().numpy(), nproma)[:ngptot], columns(ref[n], nproma)[:ngptot]
        np.testing.assert_array_equal(np.isnan(g), np.isnan(r), err_msg=n)
        assert relrms(g, r) <= 1e-12, n
    for n in CML:
        g, r = columns(f[n].cpu().numpy(), nproma)[:ngptot], columns(cin[n], nproma)[:ngptot]
        assert relrms(g, r) <= 1e-12, n


def test_fused_full_size_equals_unfused(params137):
    """655,360 columns x L137 FP32 (the benchmark workload): fused = unfused bit for bit."""
    p = params137
    f = S.FusedFields(655360, 128, 137, "sp", "cuda", satadj_out=False)
    S.synth_device(f, p, 45)
    snap = {n: f[n].clone() for n in C.INPUT_NAMES + S.SATADJ_EXTRA_INPUTS}
    S.satadj_cloudsc_blocked(f, p, 900.0)
    torch.cuda.synchronize()
    fused = {n: f[n].clone() for n in C.OUTPUT_NAMES + CML}
    for n, x in snap.items():
        f[n].copy_(x)
    del snap
    unfused_sequence(f, p)
    for n in C.OUTPUT_NAMES + CML:
        a, b = fused[n], f[n]
        assert torch.equal(torch.nan_to_num(a, nan=7.0), torch.nan_to_num(b, nan=7.0)), n
