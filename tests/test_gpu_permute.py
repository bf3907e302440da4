"""sph_permute on the GPU against torch indexing: runs of consecutive
sources (copied 16 bytes at a time), scattered gathers, a ragged tail,
4- and 8-byte fields, the element-wise path taken for unaligned outputs,
and the aliasing check."""
import numpy as np
import pytest
import torch

from pysph_amd import _abi
from pysph_amd._device import require_gpu

pytestmark = pytest.mark.gpu


def _permute(perm, srcs, dsts):
    stream = require_gpu(perm)
    esz = (_abi.ctypes.c_int32 * len(srcs))(*[s.element_size() for s in srcs])
    _abi.call("sph_permute", stream, perm.numel(), perm.data_ptr(), len(srcs),
              _abi.ptr_array([s.data_ptr() for s in srcs]),
              _abi.ptr_array([d.data_ptr() for d in dsts]), esz)
    torch.cuda.synchronize()


def _fields(n, rng):
    dev = "cuda"
    return [torch.tensor(rng.standard_normal(n), dtype=torch.float32, device=dev),
            torch.tensor(rng.integers(-2**40, 2**40, n), dtype=torch.int64, device=dev),
            torch.tensor(rng.integers(-2**30, 2**30, n), dtype=torch.int32, device=dev),
            torch.tensor(rng.standard_normal(n), dtype=torch.float64, device=dev),
            torch.tensor(rng.standard_normal(n), dtype=torch.float32, device=dev)]


@pytest.mark.parametrize("kind", ["identity", "mostly_runs", "random"])
@pytest.mark.parametrize("n", [1, 7, 4099, 100003])
def test_permute_matches_indexing(kind, n):
    rng = np.random.Generator(np.random.PCG64(n))
    if kind == "identity":
        p = np.arange(n)
    elif kind == "mostly_runs":
        p = np.arange(n)
        k = max(1, n // 50)
        a, b = rng.integers(0, n, k), rng.integers(0, n, k)
        for x, y in zip(a, b):
            p[x], p[y] = p[y], p[x]
    else:
        p = rng.permutation(n)
    perm = torch.tensor(p, dtype=torch.int32, device="cuda")
    srcs = _fields(n, rng)
    dsts = [torch.empty_like(s) for s in srcs]
    _permute(perm, srcs, dsts)
    for s, d in zip(srcs, dsts):
        assert torch.equal(d, s[perm.long()])


def test_permute_many_fields_and_unaligned_outputs():
    n = 50001
    rng = np.random.Generator(np.random.PCG64(3))
    p = np.arange(n)
    p[100:200] = p[100:200][::-1]
    perm = torch.tensor(p, dtype=torch.int32, device="cuda")
    # 30 fields: more than one launch
    srcs = []
    for k in range(30):
        srcs += [_fields(n, rng)[k % 5]]
    dsts = [torch.empty_like(s) for s in srcs]
    _permute(perm, srcs, dsts)
    for s, d in zip(srcs, dsts):
        assert torch.equal(d, s[perm.long()])
    # outputs one element past an aligned base: the element-wise path
    bases = [torch.empty(n + 1, dtype=s.dtype, device="cuda") for s in srcs[:5]]
    dsts = [b[1:] for b in bases]
    assert all(d.data_ptr() % 16 != 0 for d in dsts)
    _permute(perm, srcs[:5], dsts)
    for s, d in zip(srcs[:5], dsts):
        assert torch.equal(d, s[perm.long()])


def test_permute_rejects_aliasing():
    n = 64
    rng = np.random.Generator(np.random.PCG64(5))
    perm = torch.arange(n, dtype=torch.int32, device="cuda")
    a, b = _fields(n, rng)[:2]
    with pytest.raises(_abi.SphError, match="aliases"):
        _permute(perm, [a, b], [torch.empty_like(a), a])
