"""The allocation-free tridag_array entry point (lg_tridag_array_device_noalloc,
scratch in the library's static device array) against the allocating one
(lg_tridag_array_device, the same kernels with per-call hipMalloc scratch),
through the C ABI: two calls in a row, the second on a smaller system set,
each bit-equal to the allocating entry point's result; the entries the
reference leaves alone keep the caller's values; a zero pivot after a good
call is still reported (the status words are cleared per call), and a size
beyond the static scratch is refused with LG_ERR_ARG."""
import numpy as np
import pytest

from test_gpu_tridag import SENTINEL, systems

pytestmark = pytest.mark.gpu


def _solve(entry, lh, ny, n, a, b, c, r):
    from hip_util import DeviceArray
    dev = [DeviceArray(x) for x in (a, b, c, r)]
    du = DeviceArray(shape=r.shape, fill=SENTINEL)
    rc = entry(lh, ny, n, *[x.ptr for x in dev], du.ptr, None)
    out = du.get() if rc == 0 else None
    for x in dev + [du]:
        x.free()
    return rc, out


def test_noalloc_equals_allocating_entry_point_twice():
    from lesgo_amd import _lib
    L = _lib.lib()
    for (lh, ny, n, kind, seed) in [(129, 256, 194, "poisson", 3), (33, 64, 40, "random", 4)]:
        a, b, c, r = systems(lh, ny, n, kind, seed=seed)
        rc0, ref = _solve(L.lg_tridag_array_device, lh, ny, n, a, b, c, r)
        rc1, got = _solve(L.lg_tridag_array_device_noalloc, lh, ny, n, a, b, c, r)
        print(f"  {kind} {lh}x{ny}x{n}: rc {rc0} {rc1}", flush=True)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(got, ref)
        assert (got[n - 1, ny // 2, 2:] == SENTINEL).all()


def test_noalloc_pivot_after_good_call_and_size_limit():
    from lesgo_amd import _lib
    L = _lib.lib()
    lh, ny, n = 17, 16, 10
    a, b, c, r = systems(lh, ny, n, "random", seed=5)
    assert _solve(L.lg_tridag_array_device_noalloc, lh, ny, n, a, b, c, r)[0] == 0
    b[0, 1, 2] = 0.0
    rc, _ = _solve(L.lg_tridag_array_device_noalloc, lh, ny, n, a, b, c, r)
    assert rc == 6                          # LG_ERR_PIVOT
    assert "3, 2, 1" in L.lg_last_error().decode()
    b[0, 1, 2] = 3.0
    assert _solve(L.lg_tridag_array_device_noalloc, lh, ny, n, a, b, c, r)[0] == 0
    # 2^23 + 1 points: refused before any device access (null pointers would
    # be refused too, so pass a real one)
    from hip_util import DeviceArray
    d = DeviceArray(shape=(8,), fill=0.0)
    rc = L.lg_tridag_array_device_noalloc(3, 2, (1 << 23) // 6 + 1, d.ptr, d.ptr, d.ptr, d.ptr,
                                          d.ptr, None)
    d.free()
    assert rc == 1                          # LG_ERR_ARG
    assert "exceeds" in L.lg_last_error().decode()
