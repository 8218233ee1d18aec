"""Pins of the CPU MLP reference (oracle/mlp_ref.py) that checks the device MLP kernels.

bfloat16 rounding is checked against torch's own float32 -> bfloat16 conversion, the analytic gradients against
central finite differences in float64, and Adam against torch.optim.Adam.
"""
import numpy as np
import pytest

from oracle import mlp_ref as R


def test_bf16_rounding_matches_torch():
    import torch

    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 10000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, 65504.0],
                                 dtype=np.float32)])
    ref = torch.from_numpy(x).bfloat16().float().numpy().astype(np.float64)
    assert np.array_equal(R.bf16(x), ref)
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert R.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert R.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def _tiny(seed=0, n_in=5, hidden=4, n_out=3, layers=2, M=7):
    rng = np.random.default_rng(seed)
    dims = [n_in] + [hidden] * layers + [n_out]
    ws = [rng.standard_normal((dims[l + 1], dims[l])) * 0.7 for l in range(len(dims) - 1)]
    bs = [rng.standard_normal(dims[l + 1]) * 0.3 for l in range(len(dims) - 1)]
    x = rng.standard_normal((M, n_in))
    t = rng.standard_normal((M, n_out))
    wgt = np.array([1.0, 0.5, 2.0])[:n_out]
    return ws, bs, x, t, wgt


@pytest.mark.parametrize("layers", [1, 2, 3])
def test_gradients_match_finite_differences(layers):
    ws, bs, x, t, wgt = _tiny(layers=layers)
    loss, dws, dbs = R.loss_and_grads(ws, bs, x, t, wgt, emulate=False)
    h = 1e-6
    for l in range(len(ws)):
        for (i, j) in [(0, 0), (1, 2), (ws[l].shape[0] - 1, ws[l].shape[1] - 1)]:
            wp = [w.copy() for w in ws]
            wm = [w.copy() for w in ws]
            wp[l][i, j] += h
            wm[l][i, j] -= h
            fd = (R.loss_and_grads(wp, bs, x, t, wgt, emulate=False)[0] -
                  R.loss_and_grads(wm, bs, x, t, wgt, emulate=False)[0]) / (2 * h)
            assert abs(fd - dws[l][i, j]) <= 1e-7 + 1e-6 * abs(fd)
        for i in [0, bs[l].shape[0] - 1]:
            bp = [b.copy() for b in bs]
            bm = [b.copy() for b in bs]
            bp[l][i] += h
            bm[l][i] -= h
            fd = (R.loss_and_grads(ws, bp, x, t, wgt, emulate=False)[0] -
                  R.loss_and_grads(ws, bm, x, t, wgt, emulate=False)[0]) / (2 * h)
            assert abs(fd - dbs[l][i]) <= 1e-7 + 1e-6 * abs(fd)


def test_forward_is_linear_silu_stack():
    ws, bs, x, t, wgt = _tiny(layers=1)
    y, zs, acts = R.forward(ws, bs, x, emulate=False)
    z = x @ ws[0].T + bs[0]
    np.testing.assert_allclose(zs[0], z)
    np.testing.assert_allclose(y, (z * (1 / (1 + np.exp(-z)))) @ ws[1].T + bs[1])


def test_adam_matches_torch():
    import torch

    rng = np.random.default_rng(1)
    w0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(5)]
    p = torch.nn.Parameter(torch.tensor(w0, dtype=torch.float64))
    opt = torch.optim.Adam([p], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    w, m, v = w0.copy(), np.zeros(50), np.zeros(50)
    for step, g in enumerate(grads, 1):
        p.grad = torch.tensor(g)
        opt.step()
        w, m, v = R.adam(w, m, v, g, step, 3e-3)
    np.testing.assert_allclose(w, p.detach().numpy(), rtol=1e-12, atol=1e-14)
