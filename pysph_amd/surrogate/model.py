"""The edge-MLP network on the GPU (sph_mlp_* calls of the C ABI)."""
import ctypes

import numpy as np
import torch

from .. import _abi
from .._device import require_gpu


def param_count():
    return int(_abi.load().sph_mlp_param_count())


def init_params(seed=0):
    """He-style initial weights in the flat layout order (FP32): matrices
    ~ N(0, 2/fan_in), biases 0, the decoder scaled by 0.1."""
    H, NIN, NE, R = (_abi.SPH_MLP_HIDDEN, _abi.SPH_MLP_NIN, _abi.SPH_MLP_NE,
                     _abi.SPH_MLP_ROUNDS)
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []

    def mat(rows, cols, scale=1.0):
        parts.append((rng.normal(0.0, np.sqrt(2.0 / rows), rows * cols) * scale))

    def vec(n):
        parts.append(np.zeros(n))
    mat(NIN, H)
    vec(H)
    for _ in range(R):
        mat(H, 2 * H)
        mat(NE, H)
        vec(H)
        mat(2 * H, H)
        vec(H)
    # zero decoder: the untrained network is the plain SPH acceleration
    mat(H, 2, 0.0)
    # output bias: gravity (0, -1) in units of g, so that the untrained
    # network is the plain SPH acceleration
    parts.append(np.array([0.0, -1.0, 0.0, 0.0]))
    theta = np.concatenate(parts).astype(np.float32)
    assert len(theta) == param_count()
    return theta


class EdgeMLP:
    """Owns one model on the current GPU: FP32 parameters, BF16 copies,
    gradients and Adam moments."""

    def __init__(self, params=None, seed=0):
        h = _abi._vp()
        _abi.call("sph_mlp_create", ctypes.byref(h))
        self.handle = h
        self.adam_t = 0
        self.set_params(init_params(seed) if params is None else params)

    def close(self):
        if self.handle:
            _abi.call("sph_mlp_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters
    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        if theta.shape != (param_count(),):
            raise ValueError("expected %d parameters, got %s" % (param_count(), theta.shape))
        _abi.call("sph_mlp_set_params", self.handle, theta.ctypes.data)

    def params(self):
        out = np.empty(param_count(), dtype=np.float32)
        _abi.call("sph_mlp_get", self.handle, 0, out.ctypes.data)
        return out

    def grads(self):
        out = np.empty(param_count(), dtype=np.float32)
        _abi.call("sph_mlp_get", self.handle, 1, out.ctypes.data)
        return out

    def save(self, path, **meta):
        np.savez(path, params=self.params(), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as f:
            m = cls(params=f["params"])
            m.meta = {k: f[k] for k in f.files if k != "params"}
        return m

    def timing(self, enable=-1):
        """Switch forward-pass launch timing on (1) / off (0), or only read
        (-1).  Returns ({features, gemm, edges, decode}: ms total, calls)."""
        ms = (ctypes.c_double * 4)()
        calls = ctypes.c_int64(0)
        _abi.call("sph_mlp_timing", self.handle, int(enable), ms, ctypes.byref(calls))
        return dict(zip(("features", "gemm", "edges", "decode"), list(ms))), calls.value

    # ---- compute
    def forward(self, graph, out=None, infer=False, edge_fp32=False):
        """Predicted accelerations / g, [n x 2] FP32 device tensor.  With
        `infer` the pass keeps none of the activations a backward pass would
        read (loss_backward runs its own forward pass); `edge_fp32` takes the
        FP32 vector-unit edge aggregation instead of the matrix-core one."""
        n = graph.n
        if out is None:
            out = torch.empty((max(n, 1), 2), dtype=torch.float32, device=graph.device)
        graph.struct.flags = ((_abi.SPH_MLP_INFER if infer else 0) |
                              (_abi.SPH_MLP_EDGE_FP32 if edge_fp32 else 0))
        _abi.call("sph_mlp_forward", self.handle, require_gpu(out), ctypes.byref(graph.struct),
                  out.data_ptr())
        return out

    def loss_backward(self, graph, target, out=None):
        """Forward, MSE loss over fluid nodes and backward (gradients added
        to the model's buffer).  Returns the loss as a device scalar."""
        out = self.forward(graph, out)
        stream = require_gpu(out)
        loss = torch.empty(1, dtype=torch.float32, device=out.device)
        dout = torch.empty_like(out)
        _abi.call("sph_mlp_loss", self.handle, stream, ctypes.byref(graph.struct),
                  out.data_ptr(), target.data_ptr(), loss.data_ptr(), dout.data_ptr())
        _abi.call("sph_mlp_backward", self.handle, stream, ctypes.byref(graph.struct),
                  dout.data_ptr())
        return loss

    def zero_grad(self, stream):
        _abi.call("sph_mlp_zero_grad", self.handle, stream)

    def adam(self, stream, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.adam_t += 1
        _abi.call("sph_mlp_adam", self.handle, stream, float(lr), float(b1), float(b2),
                  float(eps), self.adam_t)
