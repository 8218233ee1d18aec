"""CPU reference of the surrogate's MLP -- TEST INFRASTRUCTURE ONLY.

The checker of the BF16 MFMA forward, backward and Adam kernels of libcloudsc_hip.so (mlp.hip).  Only tests/
import it.  The network has no counterpart in the reference (OpenIFS has no learned microphysics); it is
x -> (Linear, SiLU) x H -> Linear -> y with the loss L = sum_{m,n} wgt[n] (y - t)^2 / M.

Two modes:
  emulate=True   rounds to bfloat16 (round to nearest even) at exactly the points the device kernels do --
                 the input, the weights, every hidden activation and every back-propagated gradient -- and
                 accumulates in float64, so the device must agree to FP32 summation-order error;
  emulate=False  plain float64 arithmetic, used to pin the analytic gradients against central finite
                 differences (the backward formulas are then fixed by calculus, not by a retyped formula).
"""
import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest, ties to even), returned as float64."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(ws, bs, x, emulate=True):
    """Returns (y, zs, acts): acts[l] is the input of layer l (acts[0] = x), zs[l] the pre-activation of hidden l."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    a = r(x)
    acts, zs = [a], []
    for l, (w, b) in enumerate(zip(ws, bs)):
        z = a @ r(w).T + np.asarray(b, np.float64)
        if l < len(ws) - 1:
            # the device keeps Z in float32 and rounds the activation to bfloat16
            z = z.astype(np.float32).astype(np.float64) if emulate else z
            zs.append(z)
            a = r(silu(z))
            acts.append(a)
        else:
            return z, zs, acts


def loss_and_grads(ws, bs, x, t, wgt, emulate=True):
    """(loss, [dW_l], [db_l]) of L = sum wgt[n] (y - t)^2 / M."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    M = x.shape[0]
    y, zs, acts = forward(ws, bs, x, emulate)
    d = y - np.asarray(t, np.float64)
    loss = float(np.sum(np.asarray(wgt, np.float64) * d * d) / M)
    g = r(2.0 * np.asarray(wgt, np.float64) * d / M)
    dws, dbs = [None] * len(ws), [None] * len(ws)
    for l in range(len(ws) - 1, -1, -1):
        dws[l] = g.T @ acts[l]
        dbs[l] = g.sum(axis=0)
        if l == 0:
            break
        da = g @ r(ws[l])
        g = r(da * dsilu(zs[l - 1]))
    return loss, dws, dbs


def adam(w, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected), float64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
