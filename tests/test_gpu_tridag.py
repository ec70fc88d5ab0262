"""tridag_array through the C ABI (lg_tridag_array with host arrays,
lg_tridag_array_device with device arrays) against the oracle's
tridag_array (tridag_array.f90:165-246), which a dense linear solve pins
(tests/test_oracle_ops.py::test_tridag_dense): the systems
press_stag_array builds (Neumann rows at both ends, -k^2 - 2/dz^2 on the
diagonal) and random diagonally dominant ones, at a small size and at the
256 x 256 x 192 wind farm's (lh = 129, ny = 256, nz+1 = 194); the entries
the reference skips keep the caller's values; a zero pivot is reported with
its indices.  FP64 Thomas in both: 1e-13 relative."""
import numpy as np
import pytest

from oracle import lesgo_oracle as o

pytestmark = pytest.mark.gpu

SENTINEL = 12345.0


def systems(lh, ny, n, kind, seed=0):
    rng = np.random.default_rng(seed)
    shp = (n, ny, lh)
    if kind == "poisson":
        dz = 1.0 / (n - 2)
        kx = np.arange(lh)[None, None, :] * 1.0
        ky = np.fft.fftfreq(ny, 1.0 / ny)[None, :, None] * 2.0
        a = np.full(shp, 1.0 / dz ** 2)
        c = np.full(shp, 1.0 / dz ** 2)
        b = np.broadcast_to(-(kx ** 2 + ky ** 2) - 2.0 / dz ** 2, shp).copy()
        a[0], b[0], c[0] = o.BOGUS, -1.0, 1.0
        a[-1], b[-1], c[-1] = -1.0, 1.0, o.BOGUS
    else:
        a = rng.uniform(-1.0, 1.0, shp)
        c = rng.uniform(-1.0, 1.0, shp)
        b = 2.5 + rng.uniform(0.0, 1.0, shp)
    r = rng.standard_normal((n, ny, 2 * lh))
    return a, b, c, r


def oracle_solve(lh, ny, n, a, b, c, r):
    p = o.Params(2 * (lh - 1), ny, n - 2)
    rc = r[..., 0::2] + 1j * r[..., 1::2]
    uc = o.tridag_array(p, a, b, c, rc)
    u = np.empty_like(r)
    u[..., 0::2], u[..., 1::2] = uc.real, uc.imag
    return u


def written_mask(lh, ny, n):
    """Entries of u (in the interleaved layout) that tridag_array writes."""
    m = np.zeros((n, ny, 2 * lh), dtype=bool)
    m[0, :, :2 * (lh - 1)] = True
    m[1:, :, :2 * (lh - 1)] = True
    m[1:, ny // 2, :] = False
    m[1:, 0, 0:2] = False
    return m


def host_call(lh, ny, n, a, b, c, r, u):
    from lesgo_amd import _lib
    D = _lib._D
    args = [np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c, r)]
    return _lib.lib().lg_tridag_array(lh, ny, n, *[x.ctypes.data_as(D) for x in args],
                                      u.ctypes.data_as(D))


@pytest.mark.parametrize("lh,ny,n", [(17, 16, 10), (129, 256, 194)])
@pytest.mark.parametrize("kind", ["poisson", "random"])
def test_tridag_array_host_matches_oracle(lh, ny, n, kind):
    a, b, c, r = systems(lh, ny, n, kind)
    u = np.full_like(r, SENTINEL)
    assert host_call(lh, ny, n, a, b, c, r, u) == 0
    ref = oracle_solve(lh, ny, n, a, b, c, r)
    m = written_mask(lh, ny, n)
    err = np.linalg.norm(u[m] - ref[m]) / np.linalg.norm(ref[m])
    print(f"  {kind} {lh}x{ny}x{n}: rel L2 {err:.2e}")
    assert err <= 1e-13
    assert (u[~m] == SENTINEL).all()


def test_tridag_array_device_matches_host():
    from hip_util import DeviceArray
    from lesgo_amd import _lib
    lh, ny, n = 129, 256, 194
    a, b, c, r = systems(lh, ny, n, "poisson", seed=3)
    u_host = np.full_like(r, SENTINEL)
    assert host_call(lh, ny, n, a, b, c, r, u_host) == 0
    dev = [DeviceArray(x) for x in (a, b, c, r)]
    du = DeviceArray(shape=r.shape, fill=SENTINEL)
    rc = _lib.lib().lg_tridag_array_device(lh, ny, n, *[x.ptr for x in dev], du.ptr, None)
    assert rc == 0
    assert np.array_equal(du.get(), u_host)


@pytest.mark.parametrize("row", [1, 2])
def test_tridag_array_zero_pivot_is_reported(row):
    from lesgo_amd import _lib
    lh, ny, n = 17, 16, 10
    a, b, c, r = systems(lh, ny, n, "random", seed=5)
    jx, jy = 3, 2                       # the reference's 1-based indices
    if row == 1:
        b[0, jy - 1, jx - 1] = 0.0
    else:
        # bet = b(2) - a(2) c(1) / b(1) = 2 - 4 * 1 / 2 = 0 exactly
        b[0, jy - 1, jx - 1], c[0, jy - 1, jx - 1] = 2.0, 1.0
        a[1, jy - 1, jx - 1], b[1, jy - 1, jx - 1] = 4.0, 2.0
    u = np.zeros_like(r)
    rc = host_call(lh, ny, n, a, b, c, r, u)
    assert rc == 6                       # LG_ERR_PIVOT
    msg = _lib.lib().lg_last_error().decode()
    assert f"{jx}, {jy}, {row}" in msg, msg
