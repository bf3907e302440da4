"""Device plumbing: where particle data lives and which stream calls use.

PyTorch-ROCm provides device memory and streams only.  Particle data lives
on the GPU whenever one is visible; without one, arrays live on the host so
that set-up code (operation grouping, integrator bookkeeping) can run, but
every compute call raises `GPURequired`.
"""
import torch

from . import _abi


class GPURequired(RuntimeError):
    """Raised by a compute call when no GPU is visible or data is on the host."""


_dev = None


def default_device():
    global _dev
    if _dev is None:
        _dev = torch.device("cuda", 0) if torch.cuda.is_available() else torch.device("cpu")
    return _dev


def set_default_device(dev):
    global _dev
    _dev = torch.device(dev)


def require_gpu(*tensors):
    """Check that compute can run: the library loads and every tensor is on
    a GPU.  Returns the current HIP stream handle."""
    _abi.load()
    if not torch.cuda.is_available():
        raise GPURequired("this operation runs on the GPU and no GPU is visible")
    for t in tensors:
        if t is not None and t.device.type != "cuda":
            raise GPURequired("tensor on %s; particle data must be on the GPU" % t.device)
    return stream_handle()


def stream_handle():
    return torch.cuda.current_stream().cuda_stream


def dtype_code(dt):
    if dt == torch.float32:
        return _abi.SPH_F32
    if dt == torch.float64:
        return _abi.SPH_F64
    raise TypeError("real fields must be float32 or float64, got %s" % dt)


def ptr(t):
    return t.data_ptr() if t is not None else None
