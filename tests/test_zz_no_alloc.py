"""The no-allocation radix sort (svo_k_radix_sort_noalloc) against the
allocating one (svo_k_radix_sort): the same keys sorted in place in device
memory the caller owns, through the C ABI, on two calls in a row whose second
input is smaller (the fixed scratch keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 22


def _noalloc_sort(ctx, keys: np.ndarray, bits: int) -> np.ndarray:
    import svo_amd
    L = svo_amd.lib()
    nbytes = keys.nbytes
    p = C.c_void_p()
    svo_amd._check(L.svo_device_alloc(ctx.handle, nbytes, C.byref(p)), "device alloc")
    try:
        svo_amd._check(L.svo_device_copy(ctx.handle, p, keys.ctypes.data, nbytes, 0), "upload")
        st = L.svo_k_radix_sort_noalloc(0, p, keys.size, keys.dtype.itemsize, bits, None)
        svo_amd._check(st, "noalloc sort")
        out = np.empty_like(keys)
        svo_amd._check(L.svo_device_copy(ctx.handle, out.ctypes.data, p, nbytes, 1), "download")
        return out
    finally:
        L.svo_device_free(ctx.handle, p, nbytes)


def _keys(n, dtype, bits, seed):
    rng = np.random.default_rng(seed)
    hi = (1 << bits) if bits < 64 else None
    if hi is None:
        return rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    a = rng.integers(0, hi, n, dtype=np.uint64)
    a[: n // 8] = a[n // 8: 2 * (n // 8)]  # runs of repeated keys
    return a.astype(dtype)


# (dtype, bits): 4 and 8 passes end in the caller's buffer, 3 and 5 in the
# scratch buffer (copied back)
@pytest.mark.parametrize("dtype,bits", [(np.uint32, 30), (np.uint32, 24), (np.uint64, 33),
                                        (np.uint64, 64)])
def test_noalloc_sort_equals_allocating_sort(gpu_ctx, dtype, bits):
    import svo_amd
    for call, n in enumerate((3_000_001, 70_007)):
        keys = _keys(n, dtype, bits, seed=100 * bits + call)
        print(f"noalloc sort: dtype {np.dtype(dtype).name} bits {bits} n {n} call {call}",
              flush=True)
        got = _noalloc_sort(gpu_ctx, keys, bits)
        want = svo_amd.k_radix_sort(keys, bits)
        assert np.array_equal(got, want), f"call {call}: differs from svo_k_radix_sort"
        assert np.array_equal(got, np.sort(keys, kind="stable"))


def test_noalloc_sort_limits(gpu_ctx):
    import svo_amd
    L = svo_amd.lib()
    # rejected on the host before any launch
    assert L.svo_k_radix_sort_noalloc(0, C.c_void_p(1), MAX_KEYS + 1, 4, 32, None) == \
        svo_amd.E_TOO_LARGE
    assert L.svo_k_radix_sort_noalloc(0, C.c_void_p(1), 16, 2, 16, None) == svo_amd.E_INVALID
    assert L.svo_k_radix_sort_noalloc(0, None, 0, 4, 32, None) == svo_amd.OK
    # the largest accepted size
    keys = _keys(MAX_KEYS, np.uint64, 40, seed=7)
    print(f"noalloc sort: largest size n {MAX_KEYS}", flush=True)
    got = _noalloc_sort(gpu_ctx, keys, 40)
    assert np.array_equal(got, svo_amd.k_radix_sort(keys, 40))
