"""Python host of the learned CLOUDSC surrogate and the hybrid dispatcher (include/cloudsc_surrogate.h).

MLP        -- the BF16 MFMA multilayer perceptron of libcloudsc_hip.so (forward, backward, Adam on the device).
Surrogate  -- the CLOUDSC surrogate: per-column features (14 x KLEV) -> targets (6 x KLEV), its normalisation,
              training on plain-kernel outputs, network-only prediction and the hybrid entry.
train_surrogate -- the training loop: fresh seeded synthetic atmospheres, the plain FP32 kernel as teacher.

There is no CPU fallback: every call goes to libcloudsc_hip.so and raises CloudscError on failure.
"""
import ctypes
import math

import numpy as np

from . import cloudsc as C

NFEAT, NTARG = 14, 6
FEATURES = ["PT", "PQ", "PA", "log10 PCLV(l)", "log10 PCLV(i)", "log10 PCLV(r)", "log10 PCLV(s)",
            "PTENDENCY_CML_T", "PTENDENCY_CML_Q", "PTENDENCY_CML_A", "PTENDENCY_CML_CLD(l)", "PTENDENCY_CML_CLD(i)",
            "PQ/qsat_mix(PT,PAP)", "log10 PLUDE"]
TARGETS = ["PTENDENCY_LOC_T", "PTENDENCY_LOC_Q", "PTENDENCY_LOC_A", "PTENDENCY_LOC_CLD(l)", "PTENDENCY_LOC_CLD(i)",
           "PFPLSL+PFPLSN (lower half level)"]

_vp = ctypes.c_void_p


def _lib():
    L = C.lib()
    if not getattr(L, "_surrogate_typed", False):
        L.cloudsc_mlp_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_uint64]
        L.cloudsc_mlp_destroy.argtypes = [_vp]
        L.cloudsc_mlp_dims.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), _vp, _vp]
        L.cloudsc_mlp_forward.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp]
        L.cloudsc_mlp_gradients.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]
        L.cloudsc_mlp_adam.argtypes = [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp]
        L.cloudsc_mlp_train_step.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_float, _vp]
        L.cloudsc_mlp_param.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_size_t, ctypes.c_int]
        L.cloudsc_mlp_gemm_tn.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, _vp]
        L.cloudsc_surrogate_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_uint64]
        L.cloudsc_surrogate_destroy.argtypes = [_vp]
        L.cloudsc_surrogate_mlp.argtypes = [_vp]
        L.cloudsc_surrogate_mlp.restype = _vp
        L.cloudsc_surrogate_dims.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.cloudsc_surrogate_norm.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
        L._surrogate_typed = True
    return L


def _call(name, *args):
    C._check(name, getattr(_lib(), name)(*args))


def _stream(device):
    import torch

    return torch.cuda.current_stream(device).cuda_stream


class MLP:
    """x[n_in] -> (Linear, SiLU) x hidden_layers -> Linear -> y[n_out]; widths and batches multiples of 128."""

    def __init__(self, n_in, n_hidden, hidden_layers, n_out, max_batch, seed=1, handle=None):
        self._own = handle is None
        if handle is None:
            h = _vp()
            _call("cloudsc_mlp_create", ctypes.byref(h), n_in, n_hidden, hidden_layers, n_out, max_batch, seed)
            handle = h.value
        self.h = handle
        nl = ctypes.c_int()
        din = (ctypes.c_int * 16)()
        dout = (ctypes.c_int * 16)()
        _call("cloudsc_mlp_dims", self.h, ctypes.byref(nl), din, dout)
        self.n_layers = nl.value
        self.din = list(din[: nl.value])
        self.dout = list(dout[: nl.value])
        self.max_batch = max_batch

    def close(self):
        if self._own and self.h:
            _lib().cloudsc_mlp_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, x, y=None):
        """x: torch bf16 [M, n_in] on the device -> y float32 [M, n_out]."""
        import torch

        M = x.shape[0]
        if y is None:
            y = torch.empty((M, self.dout[-1]), dtype=torch.float32, device=x.device)
        _call("cloudsc_mlp_forward", self.h, x.data_ptr(), M, y.data_ptr(), _stream(x.device))
        return y

    def gradients(self, x, xT, t, wgt):
        _call("cloudsc_mlp_gradients", self.h, x.data_ptr(), xT.data_ptr(), t.data_ptr(), wgt.data_ptr(),
              x.shape[0], _stream(x.device))

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, device="cuda"):
        _call("cloudsc_mlp_adam", self.h, lr, beta1, beta2, eps, _stream(device))

    def train_step(self, x, xT, t, wgt, lr):
        _call("cloudsc_mlp_train_step", self.h, x.data_ptr(), xT.data_ptr(), t.data_ptr(), wgt.data_ptr(),
              x.shape[0], lr, _stream(x.device))

    def _param(self, layer, which, shape, value=None):
        n = int(np.prod(shape))
        if value is None:
            out = np.empty(n, dtype=np.float32)
            _call("cloudsc_mlp_param", self.h, layer, which, out.ctypes.data, n, 0)
            return out.reshape(shape)
        v = np.ascontiguousarray(value, dtype=np.float32).reshape(-1)
        _call("cloudsc_mlp_param", self.h, layer, which, v.ctypes.data, n, 1)

    def weight(self, layer, value=None):
        return self._param(layer, 0, (self.dout[layer], self.din[layer]), value)

    def bias(self, layer, value=None):
        return self._param(layer, 1, (self.dout[layer],), value)

    def weight_grad(self, layer):
        return self._param(layer, 2, (self.dout[layer], self.din[layer]))

    def bias_grad(self, layer):
        return self._param(layer, 3, (self.dout[layer],))

    def last_output(self, M):
        return self._param(self.n_layers - 1, 4, (M, self.dout[-1]))

    def last_loss(self):
        return float(self._param(0, 5, (1,))[0])


def gemm_tn(a, b):
    """C = A @ B.T on the matrix cores (A [M, K], B [N, K] bf16 torch tensors; M, N multiples of 128, K of 32)."""
    import torch

    M, K = a.shape
    N = b.shape[0]
    c = torch.empty((M, N), dtype=torch.float32, device=a.device)
    _call("cloudsc_mlp_gemm_tn", a.data_ptr(), a.stride(0), b.data_ptr(), b.stride(0), c.data_ptr(), c.stride(0), M,
          N, K, _stream(a.device))
    return c


def _blocked_call(name, fields, params, ptsphy, *extra):
    L = _lib()
    args = list(extra) + [ctypes.byref(params), ctypes.c_int(fields.ngptot), ctypes.c_int(fields.nproma),
                          ctypes.c_int(fields.klev), ctypes.c_int(fields.ngpblks), ctypes.c_float(ptsphy)]
    args += C.blocked_args(fields.t, lambda x: x.data_ptr())
    args.append(ctypes.c_void_p(_stream(fields.device)))
    C._check(name, getattr(L, name)(*args))


class Surrogate:
    """The CLOUDSC surrogate (BF16 MLP over 14 x KLEV features -> 6 x KLEV targets)."""

    def __init__(self, klev=137, hidden=1024, hidden_layers=3, max_batch=4096, seed=1):
        h = _vp()
        _call("cloudsc_surrogate_create", ctypes.byref(h), klev, hidden, hidden_layers, max_batch, seed)
        self.h = h.value
        k, nin, nout, mb = (ctypes.c_int() for _ in range(4))
        _call("cloudsc_surrogate_dims", self.h, ctypes.byref(k), ctypes.byref(nin), ctypes.byref(nout),
              ctypes.byref(mb))
        self.klev, self.n_in, self.n_out, self.max_batch = k.value, nin.value, nout.value, mb.value
        self.hidden, self.hidden_layers = hidden, hidden_layers
        self.fitted = False
        self.mlp = MLP(0, 0, 0, 0, mb.value, handle=_lib().cloudsc_surrogate_mlp(self.h))

    def close(self):
        if self.h:
            self.mlp.h = None
            _lib().cloudsc_surrogate_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- normalisation ------------------------------------------------------------------------------------------
    def norm(self, value=None):
        """Normalisation statistics and target weights (dict of numpy arrays); with ``value``, sets them."""
        nx, ny = NFEAT * self.klev, NTARG * self.klev
        if value is None:
            mx, sx = np.empty(nx, np.float32), np.empty(nx, np.float32)
            my, sy, wy = np.empty(ny, np.float32), np.empty(ny, np.float32), np.empty(ny, np.float32)
            _call("cloudsc_surrogate_norm", self.h, mx.ctypes.data, sx.ctypes.data, my.ctypes.data, sy.ctypes.data,
                  wy.ctypes.data, 1)
            return {"mu_x": mx, "sd_x": sx, "mu_y": my, "sd_y": sy, "wgt_y": wy}
        v = {k: np.ascontiguousarray(value[k], dtype=np.float32) for k in ("mu_x", "sd_x", "mu_y", "sd_y")}
        w = np.ascontiguousarray(value["wgt_y"], dtype=np.float32) if "wgt_y" in value else None
        _call("cloudsc_surrogate_norm", self.h, v["mu_x"].ctypes.data, v["sd_x"].ctypes.data, v["mu_y"].ctypes.data,
              v["sd_y"].ctypes.data, None if w is None else w.ctypes.data, 0)
        self.fitted = True

    def fit_norm(self, fields, params, ptsphy=900.0):
        """Statistics from a blocked field set that holds plain-kernel inputs and outputs."""
        _blocked_call("cloudsc_surrogate_fit_norm", fields, params, ptsphy, ctypes.c_void_p(self.h))
        self.fitted = True

    # -- training and inference ---------------------------------------------------------------------------------
    def train(self, fields, params, lr, batch=4096, ptsphy=900.0):
        """One pass over the columns of ``fields`` (plain-kernel inputs and outputs) in batches of ``batch``
        columns, one Adam step each; returns the mean loss."""
        loss = ctypes.c_float()
        _blocked_call("cloudsc_surrogate_train", fields, params, ptsphy, ctypes.c_void_p(self.h), ctypes.c_float(lr),
                      ctypes.c_int(batch), ctypes.byref(loss))
        return loss.value

    def predict(self, fields, params, ptsphy=900.0):
        _blocked_call("cloudsc_surrogate_predict_sp", fields, params, ptsphy, ctypes.c_void_p(self.h))

    def hybrid(self, fields, params, rerun_fraction=0.01, rerun_seed=12345, gate_threshold=float("inf"),
               ptsphy=900.0):
        """Network on all columns, plain kernel on the rerun and gated columns; returns the plain column count."""
        n = ctypes.c_int()
        _blocked_call("cloudsc_hybrid_sp_c", fields, params, ptsphy, ctypes.c_void_p(self.h),
                      ctypes.c_float(rerun_fraction), ctypes.c_uint64(rerun_seed), ctypes.c_float(gate_threshold),
                      ctypes.byref(n))
        return n.value

    def gate_scores(self, fields, params, ptsphy=900.0):
        import torch

        s = torch.empty(fields.ngptot, dtype=torch.float32, device=fields.device)
        _blocked_call("cloudsc_surrogate_gate_scores", fields, params, ptsphy, ctypes.c_void_p(self.h),
                      ctypes.c_void_p(s.data_ptr()))
        return s

    # -- persistence --------------------------------------------------------------------------------------------
    def state(self):
        st = {f"w{l}": self.mlp.weight(l) for l in range(self.mlp.n_layers)}
        st.update({f"b{l}": self.mlp.bias(l) for l in range(self.mlp.n_layers)})
        st.update(self.norm())
        st["dims"] = np.array([self.klev, self.hidden, self.hidden_layers], dtype=np.int64)
        return st

    def load_state(self, st):
        for l in range(self.mlp.n_layers):
            self.mlp.weight(l, st[f"w{l}"])
            self.mlp.bias(l, st[f"b{l}"])
        self.norm(st)
        self.fitted = True

    def save(self, path):
        np.savez(path, **self.state())

    @classmethod
    def load(cls, path, max_batch=4096):
        st = dict(np.load(path, allow_pickle=False))
        klev, hidden, layers = (int(x) for x in st["dims"])
        s = cls(klev, hidden, layers, max_batch)
        s.load_state(st)
        return s


def train_surrogate(sur, params, n_columns, chunk=262144, lr=3e-4, seed0=1000, nproma=128, batch=4096, log=None,
                    device="cuda", epochs=1, cosine=False, flux_weight=1.0, surface_weight=1.0):
    """Trains ``sur`` on ``n_columns`` synthetic columns (seeds seed0, seed0+1, ...) labelled by the plain FP32
    kernel, ``chunk`` columns per generated field set, ``epochs`` passes over the same seeds.  Statistics are
    fitted on the first chunk; then the loss weights of the precipitation-flux targets are multiplied by
    ``flux_weight`` and that of the surface flux by ``surface_weight`` as well (the surface precipitation is one
    of the 6 x KLEV targets).  With ``cosine`` the learning rate follows a cosine from ``lr`` down to lr/50 over
    all chunks.  Returns the list of per-chunk mean training losses."""
    import torch

    f = C.BlockedFields(chunk, nproma, sur.klev, "sp", device)
    plude = torch.empty_like(f["plude"])
    losses = []
    nchunks = max(1, (n_columns + chunk - 1) // chunk)
    total = nchunks * epochs
    for i in range(total):
        C.synth_device(f, params, seed0 + i % nchunks)
        plude.copy_(f["plude"])
        C.cloudsc_blocked(f, params, 900.0)
        f["plude"].copy_(plude)  # the features read the detrained liquid as CLOUDSC receives it
        if i == 0 and not sur.fitted:
            sur.fit_norm(f, params)
            if flux_weight != 1.0 or surface_weight != 1.0:
                nrm = sur.norm()
                k = sur.klev
                nrm["wgt_y"][5 * k:6 * k] *= flux_weight
                nrm["wgt_y"][6 * k - 1] *= surface_weight
                sur.norm(nrm)
        rate = lr
        if cosine:
            rate = max(lr * 0.5 * (1.0 + math.cos(math.pi * i / max(total - 1, 1))), lr / 50.0)
        losses.append(sur.train(f, params, rate, batch=batch))
        if log:
            log(f"columns {(i + 1) * chunk} (epoch {i // nchunks + 1}): loss {losses[-1]:.4f}, lr {rate:.2e}")
    del f, plude
    return losses
