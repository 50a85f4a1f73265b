"""The ray-queue radix sort (rx_radix_sort_pairs) against numpy's stable
argsort: the length is read on the device, clamped to the capacity, and
queues shorter than the minimum are left in place."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sort(keys, count, cap, key_bits, min_n=0):
    import torch

    from roxy_dcc_amd import _native as nat

    k = torch.zeros(cap, dtype=torch.int32, device="cuda")
    k[: len(keys)] = torch.from_numpy(np.asarray(keys, np.uint32).view(np.int32)).cuda()
    v = torch.arange(cap, dtype=torch.int32, device="cuda")
    k2, v2 = torch.empty_like(k), torch.empty_like(v)
    cnt = torch.tensor([count], dtype=torch.int32, device="cuda")
    sb = C.c_size_t()
    nat.check(nat.lib().rx_radix_sort_scratch_size(cap, C.byref(sb)))
    scratch = torch.empty(sb.value, dtype=torch.uint8, device="cuda")
    nat.check(nat.lib().rx_radix_sort_pairs(k.data_ptr(), v.data_ptr(), k2.data_ptr(), v2.data_ptr(), cnt.data_ptr(),
                                            cap, min_n, key_bits, scratch.data_ptr(), sb.value, nat.stream_ptr()))
    torch.cuda.synchronize()
    return k.cpu().numpy().view(np.uint32), v.cpu().numpy()


@pytest.mark.parametrize("n,cap,bits,seed", [(1, 1, 30, 0), (100, 100, 8, 1), (4096, 5000, 30, 2),
                                             (4097, 4097, 16, 3), (123457, 200000, 30, 4), (1 << 20, 1 << 20, 24, 5),
                                             (3000000, 3000000, 30, 6)])
def test_radix_sort_matches_stable_argsort(n, cap, bits, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << bits, size=cap, dtype=np.uint64).astype(np.uint32)
    # long runs of equal keys: stability is visible in the values
    keys[: cap // 3] = (rng.integers(0, 5, size=cap // 3) << max(bits - 3, 0)).astype(np.uint32)
    print(f"sort n={n} cap={cap} bits={bits}", flush=True)
    k, v = _sort(keys, n, cap, bits)
    order = np.argsort(keys[:n], kind="stable")
    np.testing.assert_array_equal(v[:n], order)
    np.testing.assert_array_equal(k[:n], keys[:n][order])
    np.testing.assert_array_equal(v[n:], np.arange(n, cap))      # past the length: untouched


def test_radix_sort_clamps_and_skips():
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 1 << 30, size=10000, dtype=np.uint64).astype(np.uint32)
    # a length above the capacity sorts the capacity
    k, v = _sort(keys, 50000, 10000, 30)
    np.testing.assert_array_equal(v, np.argsort(keys, kind="stable"))
    # below min_n nothing moves; zero and negative lengths sort nothing
    for count, min_n in ((5000, 8192), (0, 0), (-3, 0)):
        k, v = _sort(keys, count, 10000, 30, min_n=min_n)
        np.testing.assert_array_equal(v, np.arange(10000))
        np.testing.assert_array_equal(k, keys)
