"""sph_scan_counts_noalloc (no device allocation; scratch in a fixed device
array of the library) against sph_scan_counts, through the C ABI: two calls
in a row, the second with a smaller input, bit-exact offsets and totals,
and the size limit rejected with an error."""
import ctypes

import numpy as np
import pytest
import torch

from pysph_amd import _abi

pytestmark = pytest.mark.gpu


def _scan(name, counts):
    n = counts.numel()
    off = torch.full((n + 1,), -7, dtype=torch.int64, device=counts.device)
    total = ctypes.c_int64(-1)
    _abi.call(name, torch.cuda.current_stream().cuda_stream, n, counts.data_ptr(),
              off.data_ptr(), ctypes.byref(total))
    return off.cpu().numpy(), int(total.value)


@pytest.mark.parametrize("seed", [0, 1])
def test_noalloc_scan_matches_allocating_scan(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    # a large input first, then a smaller one: the second call must not see
    # the first call's partials
    for n in (3_000_001, 70_001, 1):
        counts = torch.as_tensor(rng.integers(0, 120, n), dtype=torch.int64, device="cuda")
        ref, tref = _scan("sph_scan_counts", counts)
        got, tgot = _scan("sph_scan_counts_noalloc", counts)
        assert np.array_equal(got, ref), n
        assert tgot == tref == int(counts.sum().item())
        print("n=%d total=%d" % (n, tgot))


def test_noalloc_scan_without_host_total_and_limits():
    counts = torch.arange(5000, dtype=torch.int64, device="cuda")
    off = torch.empty(5001, dtype=torch.int64, device="cuda")
    _abi.call("sph_scan_counts_noalloc", torch.cuda.current_stream().cuda_stream, 5000,
              counts.data_ptr(), off.data_ptr(), None)
    torch.cuda.synchronize()
    assert int(off[-1].item()) == 4999 * 5000 // 2
    # empty input: offsets[0] = 0
    _abi.call("sph_scan_counts_noalloc", torch.cuda.current_stream().cuda_stream, 0, None,
              off.data_ptr(), None)
    torch.cuda.synchronize()
    assert int(off[0].item()) == 0
    # beyond the scratch capacity: rejected before any launch
    with pytest.raises(_abi.SphError, match="exceed the maximum"):
        _abi.call("sph_scan_counts_noalloc", torch.cuda.current_stream().cuda_stream,
                  (1 << 27) + 1, counts.data_ptr(), off.data_ptr(), None)
