"""Pins the oracle's FFT layer (fft.f90, derivatives.f90, test_filtermodule.f90)
against the DFT definition and closed forms, independently of its code."""
import math

import numpy as np
import pytest

from oracle import lesgo_oracle as o


def brute_r2c(f, nx, ny):
    """Unnormalised forward DFT, FFTW sign (-i), x halved: F[ky, kx]."""
    j = np.arange(ny)[:, None]
    i = np.arange(nx)[None, :]
    out = np.zeros((ny, nx // 2 + 1), dtype=complex)
    for ky in range(ny):
        for kx in range(nx // 2 + 1):
            out[ky, kx] = np.sum(f[:, :nx] * np.exp(-2j * np.pi * (kx * i / nx + ky * j / ny)))
    return out


@pytest.mark.parametrize("nx,ny", [(8, 6), (12, 10), (6, 12)])
def test_r2c_c2r_match_dft(nx, ny):
    p = o.Params(nx, ny, 4)
    sp = o.Spectral(p)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((ny, p.ld))
    F = sp.r2c(f)
    np.testing.assert_allclose(F, brute_r2c(f, nx, ny), rtol=0, atol=1e-12)
    # c2r is unnormalised: c2r(r2c(f)) = nx*ny*f
    np.testing.assert_allclose(sp.c2r(F)[:, :nx], nx * ny * f[:, :nx], atol=1e-11)


def trig_field(p, modes, shift=0.3):
    """sum_a c_a cos(kx x + ky y + phase)."""
    x = np.arange(p.nx) * p.dx
    y = np.arange(p.ny) * p.dy
    X, Y = np.meshgrid(x, y)
    f = np.zeros((p.ny, p.ld))
    fx = np.zeros_like(f)
    fy = np.zeros_like(f)
    for n, (mx, my, c) in enumerate(modes):
        kx = 2 * math.pi / p.lx * mx
        ky = 2 * math.pi / p.ly * my
        ph = kx * X + ky * Y + shift * (n + 1)
        f[:, :p.nx] += c * np.cos(ph)
        fx[:, :p.nx] += -c * kx * np.sin(ph)
        fy[:, :p.nx] += -c * ky * np.sin(ph)
    return f, fx, fy


def test_filt_da_trig():
    p = o.Params(16, 12, 4)
    sp = o.Spectral(p)
    f, fx, fy = trig_field(p, [(0, 0, 2.0), (1, 0, 0.5), (3, -2, 0.25), (7, 5, 0.1)])
    ff, dfdx, dfdy = o.filt_da(sp, f)
    np.testing.assert_allclose(ff[:, :p.nx], f[:, :p.nx], atol=1e-13)
    np.testing.assert_allclose(dfdx[:, :p.nx], fx[:, :p.nx], atol=1e-12)
    np.testing.assert_allclose(dfdy[:, :p.nx], fy[:, :p.nx], atol=1e-12)
    # Nyquist modes are killed (derivatives.f90:190-193)
    g, _, _ = trig_field(p, [(8, 1, 1.0), (2, 6, 1.0)])
    gg, gx, gy = o.filt_da(sp, g)
    assert np.abs(gg[:, :p.nx]).max() < 1e-13
    assert np.abs(gx[:, :p.nx]).max() < 1e-12


def test_ddx_ddy_ddxy_trig():
    p = o.Params(12, 16, 4, lx=3.0, ly=5.0)
    sp = o.Spectral(p)
    f, fx, fy = trig_field(p, [(2, 3, 1.0), (5, -7, 0.3)])
    np.testing.assert_allclose(o.ddx(sp, f)[:, :p.nx], fx[:, :p.nx], atol=1e-12)
    np.testing.assert_allclose(o.ddy(sp, f)[:, :p.nx], fy[:, :p.nx], atol=1e-12)
    a, b = o.ddxy(sp, f)
    np.testing.assert_allclose(a[:, :p.nx], fx[:, :p.nx], atol=1e-12)
    np.testing.assert_allclose(b[:, :p.nx], fy[:, :p.nx], atol=1e-12)


def test_test_filters_cutoff():
    """Sharp cutoff at k^2 < (pi/(alpha sqrt(dx dy)))^2 (test_filtermodule.f90:56-66)."""
    p = o.Params(32, 32, 4)
    sp = o.Spectral(p)
    kc2 = (math.pi / (2.0 * math.sqrt(p.dx * p.dy))) ** 2
    kc2b = (math.pi / (4.0 * math.sqrt(p.dx * p.dy))) ** 2
    for mx, my in [(0, 0), (1, 0), (3, 2), (5, 1), (7, 0), (8, 0), (3, 3), (12, 3)]:
        f, _, _ = trig_field(p, [(mx, my, 1.0)])
        k2 = (2 * math.pi / p.lx * mx) ** 2 + (2 * math.pi / p.ly * my) ** 2
        keep = 1.0 if k2 < kc2 else 0.0
        keep2 = 1.0 if k2 < kc2b else 0.0
        np.testing.assert_allclose(sp.test_filter(f)[:, :p.nx], keep * f[:, :p.nx], atol=1e-13)
        np.testing.assert_allclose(sp.test_test_filter(f)[:, :p.nx], keep2 * f[:, :p.nx], atol=1e-13)


def test_padd_3half_rule_exact_product():
    """padd -> c2r(big) -> product -> r2c(big) -> unpadd equals the exact
    spectral convolution truncated to |kx|<nx/2, |ky|<ny/2 (Canuto 7.2)."""
    p = o.Params(8, 8, 4)
    sp = o.Spectral(p)
    rng = np.random.default_rng(2)
    a = rng.standard_normal((p.ny, p.ld))
    b = rng.standard_normal((p.ny, p.ld))
    c = 1.0 / (p.nx * p.ny)
    A = sp.zero_oddballs(sp.r2c(c * a))
    B = sp.zero_oddballs(sp.r2c(c * b))
    ab = sp.c2r(sp.padd(A)) * sp.c2r(sp.padd(B))
    got = sp.unpadd(np.fft.rfft2(ab[:, :p.nx2]) / (p.nx2 * p.ny2))
    # brute force: full Hermitian spectra, direct convolution, truncate
    def full(H):
        F = np.zeros((p.ny, p.nx), dtype=complex)
        F[:, :p.lh] = H
        for ky in range(p.ny):
            for kx in range(1, p.nx // 2):
                F[(-ky) % p.ny, p.nx - kx] = np.conj(H[ky, kx])
        return F
    FA, FB = full(A), full(B)
    kyv = [ky if ky < p.ny // 2 else ky - p.ny for ky in range(p.ny)]
    kxv = [kx if kx < p.nx // 2 else kx - p.nx for kx in range(p.nx)]
    want = np.zeros((p.ny, p.lh), dtype=complex)
    for iy1 in range(p.ny):
        for ix1 in range(p.nx):
            for iy2 in range(p.ny):
                for ix2 in range(p.nx):
                    ky, kx = kyv[iy1] + kyv[iy2], kxv[ix1] + kxv[ix2]
                    if 0 <= kx < p.nx // 2 and -p.ny // 2 < ky < p.ny // 2:
                        want[ky % p.ny, kx] += FA[iy1, ix1] * FB[iy2, ix2]
    np.testing.assert_allclose(got, want, atol=1e-13)


def test_threaded_fft_backend_agrees_with_numpy():
    """set_fft_workers(n > 1) routes the oracle's transforms through scipy.fft
    (threaded pocketfft, the bench's multi-core CPU baseline); two steps of a
    fresh start with the dynamic model agree with the numpy backend to
    rounding (1e-13 relative L2 on u, v, w, p) and the switch restores
    numpy."""
    from oracle import lesgo_oracle as o
    runs = []
    for workers in (1, 3):
        o.set_fft_workers(workers)
        try:
            p = o.Params(32, 32, 16, initu=False, dyn_init=2)
            sp = o.Spectral(p)
            s = o.initial_condition(p, 4)
            for _ in range(2):
                o.step(sp, s)
            runs.append({n: s.f[n][1:p.nz, :, :p.nx].copy() for n in ("u", "v", "w", "p")})
        finally:
            o.set_fft_workers(1)
    assert o._FFT_WORKERS == 1
    for n in runs[0]:
        a, b = runs[0][n], runs[1][n]
        e = np.linalg.norm(a - b) / np.linalg.norm(a)
        assert e <= 1e-13, (n, e)
