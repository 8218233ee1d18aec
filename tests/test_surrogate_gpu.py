"""GPU tests of the BF16 MFMA MLP and of the CLOUDSC surrogate / hybrid dispatcher.

The products, gradients and Adam step are checked against the CPU reference oracle/mlp_ref.py (which emulates
the device's bfloat16 rounding points); the hybrid dispatcher is checked against the plain kernel and the
network-only path column by column; the ablation test trains the network and compares its held-out error with
the untrained network and with the climatological mean (the method without the learned part).
"""
import numpy as np
import pytest

from oracle import mlp_ref as R

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _torch():
    import torch

    assert torch.cuda.is_available(), "needs a GPU"
    return torch


# tile paths: 128x128 (small), 256x128 (M=32768, N=896), 256x256 with LDS filled straight from memory (M=16384, N=1024)
@pytest.mark.parametrize("M,N,K", [(128, 128, 32), (256, 384, 1920), (512, 896, 1024), (1024, 1024, 4096),
                                   (32768, 896, 32), (16384, 1024, 64), (16384, 1024, 1920), (65536, 1024, 128)])
def test_gemm_tn_matches_float64(M, N, K):
    torch = _torch()
    from openifs_amd import surrogate as S

    g = torch.Generator().manual_seed(M + N + K)
    a = torch.randn(M, K, generator=g).bfloat16()
    b = torch.randn(N, K, generator=g).bfloat16()
    c = S.gemm_tn(a.to(DEV), b.to(DEV))
    torch.cuda.synchronize()
    ref = a.double() @ b.double().T
    err = (c.cpu().double() - ref).abs().max().item()
    # FP32 accumulation of exact bf16 products: error ~ K * eps32 * |a||b|
    assert err <= 1e-5 * K ** 0.5 * 4 + 1e-6 * ref.abs().max().item()


def test_gemm_tn_rejects_bad_shapes():
    torch = _torch()
    from openifs_amd import cloudsc as C
    from openifs_amd import surrogate as S

    a = torch.zeros(100, 64, dtype=torch.bfloat16, device=DEV)
    b = torch.zeros(128, 64, dtype=torch.bfloat16, device=DEV)
    with pytest.raises(C.CloudscError):
        S.gemm_tn(a, b)


def _mlp_case(seed, n_in=256, hidden=128, layers=2, n_out=128, M=256):
    torch = _torch()
    from openifs_amd import surrogate as S

    m = S.MLP(n_in, hidden, layers, n_out, M, seed=seed)
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(M, n_in, generator=g).bfloat16()
    t = torch.randn(M, n_out, generator=g)
    w = torch.ones(n_out)
    w[n_out - 28:] = 0.0  # padded outputs carry no weight
    return m, x, t, w


# batch 4,096: the weight-gradient contractions are split in 4 stretches (few output tiles) and summed
@pytest.mark.parametrize("layers,M", [(1, 256), (3, 256), (2, 4096)])
def test_mlp_gradients_match_cpu_reference(layers, M):
    torch = _torch()
    m, x, t, w = _mlp_case(3 + layers, layers=layers, M=M)
    m.gradients(x.to(DEV), x.T.contiguous().to(DEV), t.to(DEV), w.to(DEV))
    torch.cuda.synchronize()
    ws = [m.weight(l) for l in range(m.n_layers)]
    bs = [m.bias(l) for l in range(m.n_layers)]
    loss, dws, dbs = R.loss_and_grads(ws, bs, x.float().numpy(), t.numpy(), w.numpy(), emulate=True)
    y, _, _ = R.forward(ws, bs, x.float().numpy(), emulate=True)
    assert abs(m.last_loss() - loss) <= 1e-4 * loss
    np.testing.assert_allclose(m.last_output(x.shape[0]), y, rtol=0, atol=2e-3 * np.abs(y).max())
    for l in range(m.n_layers):
        gw, gb = m.weight_grad(l), m.bias_grad(l)
        assert np.abs(gw - dws[l]).max() <= 1e-2 * np.abs(dws[l]).max(), f"dW layer {l}"
        assert np.abs(gb - dbs[l]).max() <= 1e-2 * np.abs(dbs[l]).max(), f"db layer {l}"
        # relative RMS far below the bf16 resolution of individual entries
        assert np.linalg.norm(gw - dws[l]) <= 2e-3 * np.linalg.norm(dws[l]), f"dW layer {l}"


def test_mlp_adam_step_matches_cpu_reference():
    torch = _torch()
    m, x, t, w = _mlp_case(11)
    w0 = [m.weight(l).astype(np.float64) for l in range(m.n_layers)]
    b0 = [m.bias(l).astype(np.float64) for l in range(m.n_layers)]
    m.gradients(x.to(DEV), x.T.contiguous().to(DEV), t.to(DEV), w.to(DEV))
    torch.cuda.synchronize()
    gws = [m.weight_grad(l).astype(np.float64) for l in range(m.n_layers)]
    gbs = [m.bias_grad(l).astype(np.float64) for l in range(m.n_layers)]
    m.adam(1e-3)
    torch.cuda.synchronize()
    for l in range(m.n_layers):
        wr, _, _ = R.adam(w0[l], 0.0, 0.0, gws[l], 1, 1e-3)
        br, _, _ = R.adam(b0[l], 0.0, 0.0, gbs[l], 1, 1e-3)
        np.testing.assert_allclose(m.weight(l), wr, rtol=0, atol=1e-6)
        np.testing.assert_allclose(m.bias(l), br, rtol=0, atol=1e-6)


def test_mlp_inference_forward_equals_training_forward():
    torch = _torch()
    m, x, t, w = _mlp_case(5, layers=3)
    xd = x.to(DEV)
    y = m.forward(xd)
    m.gradients(xd, x.T.contiguous().to(DEV), t.to(DEV), w.to(DEV))
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), m.last_output(x.shape[0]))


def test_mlp_large_batch_forward_equals_small_batches():
    """Large batches run on the 256-row tiles, small ones on the 128-row tiles: same numbers."""
    torch = _torch()
    from openifs_amd import surrogate as S

    m = S.MLP(256, 1024, 2, 896, 65536, seed=9)
    g = torch.Generator().manual_seed(9)
    x = torch.randn(65536, 256, generator=g).bfloat16().to(DEV)
    y_big = m.forward(x)
    y_small = torch.cat([m.forward(x[i:i + 1024]) for i in range(0, 65536, 1024)])
    torch.cuda.synchronize()
    d = (y_big - y_small).abs().max().item()
    assert d <= 1e-5 * y_small.abs().max().item(), d


def test_mlp_training_reduces_loss():
    torch = _torch()
    m, x, t, w = _mlp_case(7, layers=2)
    xd, xT, td, wd = x.to(DEV), x.T.contiguous().to(DEV), t.to(DEV), w.to(DEV)
    m.gradients(xd, xT, td, wd)
    torch.cuda.synchronize()
    first = m.last_loss()
    for _ in range(200):
        m.train_step(xd, xT, td, wd, 1e-3)
    m.gradients(xd, xT, td, wd)
    torch.cuda.synchronize()
    assert m.last_loss() < 0.2 * first


# ------------------------------------------------------------------------------------------------------------
# CLOUDSC surrogate
# ------------------------------------------------------------------------------------------------------------
def _fields(params, ngptot, seed, nproma=128, klev=137):
    from openifs_amd import cloudsc as C

    f = C.BlockedFields(ngptot, nproma, klev, "sp", DEV, fill=float("nan"))
    C.synth_device(f, params, seed)
    return f


def _clone(f):
    from openifs_amd import cloudsc as C

    g = C.BlockedFields(f.ngptot, f.nproma, f.klev, "sp", DEV)
    for k in f.t:
        g.t[k].copy_(f.t[k])
    return g


def _features_numpy(f, klev, params):
    """The 14 x KLEV raw features of every column, restated from the header's definition."""
    import oracle
    from tests.helpers import columns

    def col(n, s=None):
        a = f[n].cpu().numpy()
        c = columns(a, f.nproma)
        return c if s is None else c[:, s]

    lg = lambda v: np.log10(np.maximum(v, 0.0) + 1e-10)
    t, pap = col("pt").astype(np.float64), col("pap").astype(np.float64)
    es = np.minimum(oracle.thermo(params, "foeewm", t.reshape(-1)).reshape(t.shape) / pap, 0.5)
    rh = col("pq") / (es / (1.0 - params.retv * es))
    parts = [col("pt"), col("pq"), col("pa")] + [lg(col("pclv", s)) for s in range(4)] + \
            [col("ptendency_cml_t"), col("ptendency_cml_q"),
             col("ptendency_cml_a"), col("ptendency_cml_cld", 0), col("ptendency_cml_cld", 1), rh,
             lg(col("plude"))]
    return np.concatenate(parts, axis=1)[: f.ngptot]


def test_surrogate_prediction_equals_network_on_cpu_features(params137):
    """Feature extraction, normalisation and output writing, checked through the network itself."""
    torch = _torch()
    from openifs_amd import cloudsc as C
    from openifs_amd import surrogate as S
    from tests.helpers import columns

    klev = 137
    sur = S.Surrogate(klev, hidden=256, hidden_layers=2, max_batch=1024, seed=2)
    f = _fields(params137, 1000, 77)
    g = _clone(f)
    C.cloudsc_blocked(g, params137, 900.0)
    sur.fit_norm(g, params137)
    nrm = sur.norm()
    sur.predict(f, params137)
    torch.cuda.synchronize()
    X = np.clip((_features_numpy(f, klev, params137) - nrm["mu_x"]) / nrm["sd_x"], -10.0, 10.0)
    Xp = np.zeros((1024, sur.n_in), np.float32)
    Xp[:1000, : X.shape[1]] = X
    y = sur.mlp.forward(torch.from_numpy(Xp).bfloat16().to(DEV)).cpu().numpy()[:1000, : 6 * klev]
    y = y * nrm["sd_y"] + nrm["mu_y"]
    got = {n: columns(f[n].cpu().numpy(), f.nproma)[:1000] for n in ("ptendency_loc_t", "ptendency_loc_q",
                                                                      "ptendency_loc_a", "pfplsl", "pfplsn")}
    cld = columns(f["ptendency_loc_cld"].cpu().numpy(), f.nproma)[:1000]
    pr = np.maximum(y[:, 5 * klev:], 0.0)

    def close(a, b):
        scale = np.abs(b).max() + 1e-30
        # bf16 rounding of individual features may differ between numpy and the device
        return np.abs(a - b).max() <= 0.05 * scale

    assert close(got["ptendency_loc_t"], y[:, 0:klev])
    assert close(got["ptendency_loc_q"], y[:, klev:2 * klev])
    assert close(got["ptendency_loc_a"], y[:, 2 * klev:3 * klev])
    assert close(cld[:, 0], y[:, 3 * klev:4 * klev])
    assert close(cld[:, 1], y[:, 4 * klev:5 * klev])
    assert np.all(cld[:, 2] == 0) and np.all(cld[:, 3] == 0)
    assert close(got["pfplsl"][:, 1:] + got["pfplsn"][:, 1:], pr)
    assert np.all(got["pfplsl"][:, 0] == 0) and np.all(got["pfplsn"][:, 0] == 0)
    # phase split: liquid where the level above is warmer than RTT
    T = columns(f["pt"].cpu().numpy(), f.nproma)[:1000]
    warm = T > params137.rtt
    assert np.all(got["pfplsn"][:, 1:][warm] == 0) and np.all(got["pfplsl"][:, 1:][~warm] == 0)


def test_hybrid_selection_reproduces_plain_and_network(params137):
    torch = _torch()
    from openifs_amd import cloudsc as C
    from openifs_amd import surrogate as S
    from tests.helpers import columns

    sur = S.Surrogate(137, hidden=256, hidden_layers=2, max_batch=1024, seed=4)
    base = _fields(params137, 3000, 91)
    plain, net, all_plain, mix = _clone(base), _clone(base), _clone(base), _clone(base)
    C.cloudsc_blocked(plain, params137, 900.0)
    sur.predict(net, params137)
    n_all = sur.hybrid(all_plain, params137, rerun_fraction=1.0)
    n_mix = sur.hybrid(mix, params137, rerun_fraction=0.3, rerun_seed=5)
    n_none = sur.hybrid(_clone(base), params137, rerun_fraction=0.0)
    torch.cuda.synchronize()
    assert n_all == 3000 and n_none == 0
    assert 0.25 * 3000 < n_mix < 0.35 * 3000
    names = [n for n, _, r in C.BLOCKED_ARRAYS if r == "o"] + ["plude"]
    for n in names:
        a, p = all_plain[n].cpu().numpy(), plain[n].cpu().numpy()
        assert np.array_equal(a, p, equal_nan=True), n
    # the mixed run: each column equals either the plain or the network result in every output
    sel = None
    for n in ("ptendency_loc_t", "pfplsl", "ptendency_loc_cld", "pfsqlf"):
        m_, p_, q_ = (columns(x[n].cpu().numpy(), base.nproma)[:3000] for x in (mix, plain, net))
        ax = tuple(range(1, m_.ndim))
        is_p = np.all((m_ == p_) | (np.isnan(m_) & np.isnan(p_)), axis=ax)
        is_q = np.all((m_ == q_) | (np.isnan(m_) & np.isnan(q_)), axis=ax)
        assert np.all(is_p | is_q), n
        s = is_p & ~is_q
        sel = s if sel is None else (sel | s)
    assert abs(int(sel.sum()) - n_mix) <= 3  # columns where both paths agree exactly are ambiguous


def test_hybrid_is_reproducible_and_reruns_the_selected_columns(params137):
    """Two hybrid calls on the same inputs give the same outputs bit for bit (the rerun list is a prefix-sum
    compaction, not atomic appends), and exactly as many columns as reported equal the plain kernel in every
    output."""
    torch = _torch()
    from openifs_amd import cloudsc as C
    from openifs_amd import surrogate as S
    from tests.helpers import columns

    sur = S.Surrogate(137, hidden=256, hidden_layers=2, max_batch=1024, seed=4)
    base = _fields(params137, 4000, 92)
    plain, a, b = _clone(base), _clone(base), _clone(base)
    C.cloudsc_blocked(plain, params137, 900.0)
    na = sur.hybrid(a, params137, rerun_fraction=0.2, rerun_seed=11)
    nb = sur.hybrid(b, params137, rerun_fraction=0.2, rerun_seed=11)
    torch.cuda.synchronize()
    assert na == nb and 0.15 * 4000 < na < 0.25 * 4000
    names = [n for n, _, r in C.BLOCKED_ARRAYS if r == "o"] + ["plude"]
    for n in names:
        assert torch.equal(torch.nan_to_num(a[n], nan=7.0), torch.nan_to_num(b[n], nan=7.0)), n
    # the rerun columns are those equal to the plain kernel in every output; their count is the reported one
    same = None
    for n in names:
        x, p = columns(a[n].cpu().numpy(), base.nproma)[:4000], columns(plain[n].cpu().numpy(), base.nproma)[:4000]
        ax = tuple(range(1, x.ndim))
        eq = np.all((x == p) | (np.isnan(x) & np.isnan(p)), axis=ax)
        same = eq if same is None else (same & eq)
    assert int(same.sum()) == na


def test_gate_scores_flag_cold_supersaturation(params137):
    torch = _torch()
    from openifs_amd import surrogate as S
    from tests.helpers import columns

    sur = S.Surrogate(137, hidden=128, hidden_layers=1, max_batch=128, seed=1)
    f = _fields(params137, 512, 33)
    s = sur.gate_scores(f, params137).cpu().numpy()
    T = columns(f["pt"].cpu().numpy(), f.nproma)[:512]
    assert np.all(s >= 0) and np.all(np.isfinite(s))
    # columns without any level colder than 238 K score 0
    assert np.all(s[np.all(T >= 238.0, axis=1)] == 0)
    # RH ~ U(0.3, 1.05) by construction: scores cluster in that range
    assert 0.9 < np.median(s[s > 0]) < 1.1


def test_surrogate_ablation(params137):
    """What the trained network gains over the method without it (the mean state) and an untrained network."""
    torch = _torch()
    from openifs_amd import cloudsc as C
    from openifs_amd import surrogate as S
    from tests.helpers import columns

    klev = 137
    sur = S.Surrogate(klev, hidden=1024, hidden_layers=3, max_batch=4096, seed=6)
    held = _fields(params137, 8192, 424242)
    ref = _clone(held)
    C.cloudsc_blocked(ref, params137, 900.0)

    def errors(path):
        out = _clone(held)
        path(out)
        torch.cuda.synchronize()
        e = {}
        for n in ("ptendency_loc_t", "ptendency_loc_q"):
            e[n] = np.linalg.norm(columns(out[n].cpu().numpy(), 128) - columns(ref[n].cpu().numpy(), 128)) / \
                np.linalg.norm(columns(ref[n].cpu().numpy(), 128))
        sp = lambda x: x["pfplsl"].cpu().numpy()[:, -1, :].reshape(-1) + x["pfplsn"].cpu().numpy()[:, -1, :].reshape(-1)
        e["surface_precip"] = np.linalg.norm(sp(out) - sp(ref)) / np.linalg.norm(sp(ref))
        return e

    S.train_surrogate(sur, params137, 4096 * 16, chunk=4096 * 16, lr=3e-4, seed0=9000)  # fits statistics
    untrained = S.Surrogate(klev, hidden=1024, hidden_layers=3, max_batch=4096, seed=6)
    untrained.norm(sur.norm())
    e_untrained = errors(lambda o: untrained.predict(o, params137))
    st = sur.state()
    mean_only = S.Surrogate(klev, hidden=1024, hidden_layers=3, max_batch=4096, seed=6)
    zero = {k: (np.zeros_like(v) if k[0] in "wb" and k[1:].isdigit() else v) for k, v in st.items()}
    mean_only.load_state(zero)  # predicts the training mean of every target
    e_mean = errors(lambda o: mean_only.predict(o, params137))
    S.train_surrogate(sur, params137, 4096 * 512, chunk=4096 * 64, lr=3e-4, seed0=9100)
    e_trained = errors(lambda o: sur.predict(o, params137))
    e_hybrid = errors(lambda o: sur.hybrid(o, params137, rerun_fraction=0.01, rerun_seed=1))
    print("untrained", e_untrained, "\nmean", e_mean, "\ntrained", e_trained, "\nhybrid", e_hybrid)
    # the learned part must beat both the mean state and the untrained network on every metric
    for n in e_trained:
        assert e_trained[n] < 0.97 * e_mean[n], n
        assert e_trained[n] < 0.97 * e_untrained[n], n
        assert e_hybrid[n] <= e_trained[n] * 1.01, n
