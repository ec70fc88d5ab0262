"""CPU oracle for one lesgo LES time step (numpy, float64).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it.  The product (``lesgo_amd``) never routes through it.

It restates, in vectorised numpy, the body of the reference time loop
(main.f90:130-468) as the reference executes it in its one-rank ``PPMPI``
build (``lbz = 0``, ``nz = Nz + 1``, every plane exchange a no-op because the
1-D Cartesian communicator is non-periodic, mpi_defs.f90:77-85), with
``PPSAFETYMODE`` BOGUS fills, ``sgs_model = 5``, ``lbc_mom = 2``,
``ubc_mom = 0``, either start schedule of initial.f90:96-110 (file start
``initu = .true.`` or fresh start ``inilag = .true.``), optionally
``PPTURBINES``, and the running time averages of time_average.f90.

Array layout mirrors the reference exactly: a 3-D field is a C-ordered
``(nz+1, ny, ld)`` float64 array indexed ``[k, j, i]`` for Fortran
``f(i+1, j+1, k)`` with ``k = 0..nz``; x is contiguous and columns
``nx..ld-1`` are the in-place FFT padding.  A spectral plane is the same
memory viewed as ``(ny, lh)`` complex128 (FFTW's interleaved half-complex
layout, fft.f90:102-127).

FFTW (not vendored, no version pinned by the reference, CMakeLists.txt:112-155)
is replaced by numpy's pocketfft with FFTW's conventions: unnormalised r2c and
c2r, caller-side 1/(nx*ny) scaling, x the halved dimension.

Parity status of each function is recorded in its docstring ("pinned by"
names the test that fixes it independently of this code).
"""
from __future__ import annotations

import math
import sys
from dataclasses import dataclass, field

import numpy as np

# 2-D real transforms: numpy.fft (one thread; what the pins run) unless
# set_fft_workers(n > 1) routes them through scipy.fft's threaded pocketfft
# (the bench's multi-core CPU baseline).  Both are pocketfft with the same
# conventions; results agree to rounding (tests/test_oracle_fft.py).
_FFT_WORKERS = 1


def set_fft_workers(n: int) -> None:
    """Threads for the oracle's 2-D FFTs: 1 = numpy.fft, n > 1 = scipy.fft."""
    global _FFT_WORKERS
    _FFT_WORKERS = max(1, int(n))


def _rfft2(a, axes=(-2, -1)):
    if _FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.rfft2(a, axes=axes, workers=_FFT_WORKERS)
    return np.fft.rfft2(a, axes=axes)


def _irfft2(a, s, axes=(-2, -1), norm=None):
    if _FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.irfft2(a, s=s, axes=axes, norm=norm, workers=_FFT_WORKERS)
    return np.fft.irfft2(a, s=s, axes=axes, norm=norm)


BOGUS = -1234567890.0                      # param.f90:93
PI = 3.1415926535897932384626433           # param.f90:94 (== math.pi in double)
ZERO_LAG = 1.0e-24                         # lagrange_Sdep.f90:68


# --------------------------------------------------------------------------
# Parameters (param.f90, input_util.f90:196-235,403-409)
# --------------------------------------------------------------------------
@dataclass
class Params:
    """Run parameters and derived sizes of the one-rank PPMPI build.

    Derived sizes follow input_util.f90:196-235: ``nz = Nz/nproc + 1``,
    ``lh = nx/2+1``, ``ld = 2 lh``, ``nx2 = 3nx/2``, ``ld_big = 2(nx2/2+1)``,
    ``dz = Lz/(nz_tot-1)``.  AB2 coefficients from input_util.f90:403-409.
    """
    nx: int
    ny: int
    Nz: int                        # number of uv levels in the domain
    lx: float = 2.0 * PI
    ly: float = PI
    lz: float = 1.0
    dt: float = 2.0e-4
    zo: float = 1.0e-4
    vonk: float = 0.4
    cs_count: int = 5
    use_mean_p_force: bool = True
    mean_p_force_x: float = 1.0
    mean_p_force_y: float = 0.0
    nu: float = 0.0                # molec = .false. -> nu = 0 (sgs_param.f90:187-192)
    # initu = .true.: the state comes from a file (initial.f90:100-102) and the
    # dynamic model runs from the first cs_count step; .false.: a fresh start
    # (inilag = .true.): Euler first step, Cs_opt2 = 0.03 until DYN_init, F_*
    # initialised from MM, NN at the first dynamic call (param.f90:123,171-173)
    initu: bool = True
    dyn_init: int = 100
    # use_cfl_dt = .true.: dt from the velocity each step (cfl_util.f90:73-111,
    # main.f90:135-144); dt above is then only the value before the first step
    use_cfl_dt: bool = False
    cfl: float = 0.05

    def __post_init__(self):
        assert self.nx % 2 == 0 and self.ny % 2 == 0
        self.nz = self.Nz + 1
        self.nz_tot = self.nz
        self.lh = self.nx // 2 + 1
        self.ld = 2 * self.lh
        self.nx2 = 3 * self.nx // 2
        self.ny2 = 3 * self.ny // 2
        self.lh_big = self.nx2 // 2 + 1
        self.ld_big = 2 * self.lh_big
        self.dx = self.lx / self.nx
        self.dy = self.ly / self.ny
        self.dz = self.lz / (self.nz_tot - 1)
        self.tadv1 = 1.5
        self.tadv2 = 1.0 - self.tadv1
        self.delta = 1 * (self.dx * self.dy * self.dz) ** (1.0 / 3.0)
        self.lagran_dt = self.cs_count * self.dt


# --------------------------------------------------------------------------
# FFT layer (fft.f90, emul_complex.f90)
# --------------------------------------------------------------------------
class Spectral:
    """Wavenumbers, test-filter kernels and FFTW-convention transforms.

    kx/ky/k2 follow fft.f90:130-160 (Nyquist column and row set to 0,
    scaled by 2*pi/L).  G_test/G_test_test follow
    test_filtermodule.f90:38-123 with ``ifilter = 1`` (sharp cutoff,
    1/(nx*ny) normalisation folded in, Nyquist zeroed).
    Pinned by: tests/test_oracle_fft.py (brute-force DFT, spectral
    derivative of trig polynomials, filter passes/kills chosen modes).
    """

    def __init__(self, p: Params):
        self.p = p
        nx, ny, lh = p.nx, p.ny, p.lh
        jx = np.arange(lh, dtype=np.float64)
        kx = np.tile(jx, (ny, 1))
        kx[:, lh - 1] = 0.0
        jy = np.arange(1, ny + 1)
        kyv = (np.mod(jy - 1 + ny // 2, ny) - ny // 2).astype(np.float64)
        ky = np.tile(kyv[:, None], (1, lh))
        ky[:, lh - 1] = 0.0
        kx[ny // 2, :] = 0.0
        ky[ny // 2, :] = 0.0
        self.kx = 2.0 * PI / p.lx * kx
        self.ky = 2.0 * PI / p.ly * ky
        self.k2 = self.kx * self.kx + self.ky * self.ky

        def cutoff(alpha):
            g = np.full((ny, lh), 1.0 / (nx * ny))
            delta_t = alpha * math.sqrt(p.dx * p.dy)
            kc2 = (PI / delta_t) ** 2
            g[self.k2 >= kc2] = 0.0
            g[:, lh - 1] = 0.0
            g[ny // 2, :] = 0.0
            return g

        self.G_test = cutoff(2.0)
        self.G_test_test = cutoff(4.0)

    # -- transforms -------------------------------------------------------
    def r2c(self, f):
        """dfftw_execute_dft_r2c on (..., ny, ld) planes -> (..., ny, lh)."""
        return _rfft2(f[..., : self.p.nx], axes=(-2, -1))

    def c2r(self, fh, ld=None):
        """dfftw_execute_dft_c2r (unnormalised) -> (..., ny, ld), pads 0."""
        p = self.p
        ny = fh.shape[-2]
        nx = 2 * (fh.shape[-1] - 1)
        out = np.zeros(fh.shape[:-2] + (ny, nx + 2))
        out[..., :nx] = _irfft2(fh, s=(ny, nx), axes=(-2, -1), norm="forward")
        return out

    def zero_oddballs(self, fh):
        """f(ld-1:ld,:) = 0 ; f(:,ny/2+1) = 0 (e.g. derivatives.f90:50-52)."""
        fh[..., :, self.p.lh - 1] = 0.0
        fh[..., self.p.ny // 2, :] = 0.0
        return fh

    @staticmethod
    def muli(fh, k):
        """.MULI. (emul_complex.f90:154-206): (a_r, a_i) * i k = (-a_i k, a_r k)."""
        out = np.empty_like(fh)
        out.real = -fh.imag * k
        out.imag = fh.real * k
        return out

    # -- padding for the 3/2 rule (fft.f90:43-99) --------------------------
    def padd(self, fh):
        """(..., ny, lh) -> (..., ny2, lh_big); copies kx < nx/2, |ky| < ny/2."""
        p = self.p
        nyh = p.ny // 2
        big = np.zeros(fh.shape[:-2] + (p.ny2, p.lh_big), dtype=np.complex128)
        big[..., :nyh, : p.nx // 2] = fh[..., :nyh, : p.nx // 2]
        big[..., p.ny2 - nyh + 1:, : p.nx // 2] = fh[..., nyh + 1:, : p.nx // 2]
        return big

    def unpadd(self, big):
        """(..., ny2, lh_big) -> (..., ny, lh) with oddballs zeroed."""
        p = self.p
        nyh = p.ny // 2
        fh = np.zeros(big.shape[:-2] + (p.ny, p.lh), dtype=np.complex128)
        fh[..., :nyh, : p.nx // 2] = big[..., :nyh, : p.nx // 2]
        fh[..., nyh + 1:, : p.nx // 2] = big[..., p.ny2 - nyh + 1:, : p.nx // 2]
        return fh

    # -- filters ------------------------------------------------------------
    def test_filter(self, f):
        """test_filter (test_filtermodule.f90:126-146): r2c, *G_test, c2r."""
        return self.c2r(self.r2c(f) * self.G_test)

    def test_test_filter(self, f):
        """test_test_filter (test_filtermodule.f90:149-169) with G_test_test."""
        return self.c2r(self.r2c(f) * self.G_test_test)


# --------------------------------------------------------------------------
# derivatives.f90
# --------------------------------------------------------------------------
def filt_da(sp: Spectral, f):
    """filt_da (derivatives.f90:166-211): returns (f_filtered, dfdx, dfdy).

    Per level: f <- const f, r2c, zero oddballs, dfdx = c2r(i kx f^),
    dfdy = c2r(i ky f^), f = c2r(f^).  Applied to every level 0..nz.
    Pinned by: tests/test_oracle_fft.py::test_filt_da_trig.
    """
    p = sp.p
    const = 1.0 / (p.nx * p.ny)
    fh = sp.zero_oddballs(sp.r2c(const * f))
    return sp.c2r(fh), sp.c2r(sp.muli(fh, sp.kx)), sp.c2r(sp.muli(fh, sp.ky))


def ddx(sp: Spectral, f):
    """ddx (derivatives.f90:37-72)."""
    p = sp.p
    fh = sp.zero_oddballs(sp.r2c((1.0 / (p.nx * p.ny)) * f))
    return sp.c2r(sp.muli(fh, sp.kx))


def ddy(sp: Spectral, f):
    """ddy (derivatives.f90:75-110)."""
    p = sp.p
    fh = sp.zero_oddballs(sp.r2c((1.0 / (p.nx * p.ny)) * f))
    return sp.c2r(sp.muli(fh, sp.ky))


def ddxy(sp: Spectral, f):
    """ddxy (derivatives.f90:113-152): returns (dfdx, dfdy)."""
    p = sp.p
    fh = sp.zero_oddballs(sp.r2c((1.0 / (p.nx * p.ny)) * f))
    return sp.c2r(sp.muli(fh, sp.kx)), sp.c2r(sp.muli(fh, sp.ky))


def ddz_uv(p: Params, f, dfdz=None):
    """ddz_uv (derivatives.f90:214-264): uv -> w nodes, PPMPI+SAFETYMODE.

    dfdz(k) = (f(k)-f(k-1))/dz for k=1..nz on columns 1..nx, then BOGUS at
    k = 0, 1 (bottom rank) and nz (top rank).  Padding columns untouched.
    Pinned by: tests/test_oracle_ops.py::test_ddz_linear.
    """
    out = np.zeros_like(f) if dfdz is None else dfdz.copy()
    nx, nz = p.nx, p.nz
    const = 1.0 / p.dz
    out[1:nz + 1, :, :nx] = const * (f[1:nz + 1, :, :nx] - f[0:nz, :, :nx])
    out[0] = BOGUS
    out[1] = BOGUS
    out[nz] = BOGUS
    return out


def ddz_w(p: Params, f, dfdz=None):
    """ddz_w (derivatives.f90:267-311): w -> uv nodes, PPMPI+SAFETYMODE.

    dfdz(k) = (f(k+1)-f(k))/dz for k=0..nz-1, BOGUS at k=0 and k=nz.
    Pinned by: tests/test_oracle_ops.py::test_ddz_linear.
    """
    out = np.zeros_like(f) if dfdz is None else dfdz.copy()
    nx, nz = p.nx, p.nz
    const = 1.0 / p.dz
    out[0:nz, :, :nx] = const * (f[1:nz + 1, :, :nx] - f[0:nz, :, :nx])
    out[0] = BOGUS
    out[nz] = BOGUS
    return out


# --------------------------------------------------------------------------
# State
# --------------------------------------------------------------------------
FIELDS = ("u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz",
          "dwdx", "dwdy", "dwdz", "RHSx", "RHSy", "RHSz", "RHSx_f", "RHSy_f",
          "RHSz_f", "dpdx", "dpdy", "dpdz", "txx", "txy", "tyy", "txz", "tyz",
          "tzz", "divtx", "divty", "divtz", "p", "fxa", "fya", "fza",
          "S11", "S12", "S13", "S22", "S23", "S33", "Nu_t", "Cs_opt2",
          "F_LM", "F_MM", "F_QN", "F_NN", "Beta", "Tn_all")


@dataclass
class State:
    """sim_param + sgs_param module state (sim_param.f90:31-93,
    sgs_param.f90:29-184); every field (nz+1, ny, ld), levels 0..nz."""
    p: Params
    f: dict = field(default_factory=dict)
    jt: int = 0
    jt_total: int = 0
    turbines: "Turbines | None" = None
    ustar_lbc: np.ndarray | None = None
    lm_init: bool = False          # F_LM_MM_init (lagrange_Sdep.f90:266-277)
    qn_init: bool = False          # F_QN_NN_init (lagrange_Sdep.f90:326-337)
    # the time-stepping clock (param: dt, tadv1, tadv2; sgs_param: lagran_dt;
    # total_time); clock_set: dt holds a previous step's value (with
    # use_cfl_dt, initialize.f90:191-200 forces an Euler step otherwise)
    dt: float = 0.0
    tadv1: float = 1.5
    tadv2: float = -0.5
    lagran_dt: float = 0.0
    total_time: float = 0.0
    clock_set: bool = False

    @classmethod
    def zeros(cls, p: Params):
        s = cls(p)
        s.dt, s.tadv1, s.tadv2 = p.dt, p.tadv1, p.tadv2
        s.lagran_dt = 0.0 if p.use_cfl_dt else p.lagran_dt
        for n in FIELDS:
            s.f[n] = np.zeros((p.nz + 1, p.ny, p.ld))
        return s

    def __getattr__(self, name):
        f = self.__dict__.get("f")
        if f is not None and name in f:
            return f[name]
        raise AttributeError(name)

    def copy(self):
        s = State(self.p, {k: v.copy() for k, v in self.f.items()}, self.jt,
                  self.jt_total, self.turbines, None, self.lm_init, self.qn_init,
                  self.dt, self.tadv1, self.tadv2, self.lagran_dt, self.total_time,
                  self.clock_set)
        if self.turbines is not None:
            s.turbines = self.turbines.copy()
        return s


# --------------------------------------------------------------------------
# Seeded initial condition (replaces the /dev/urandom ic_les, initial.f90:554)
# --------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform01(seed, fid, i, j, k):
    """Counter-based U[0,1) keyed by (seed, field, i, j, k): splitmix64 of a
    packed key, top 53 bits.  Shared bit-for-bit with lesgo_amd's IC writer."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)) & _M64
        key = _splitmix64(key ^ np.uint64(fid))
        key = _splitmix64(key ^ (np.asarray(k, np.uint64) << np.uint64(42)
                                 ^ np.asarray(j, np.uint64) << np.uint64(21)
                                 ^ np.asarray(i, np.uint64)))
    return (key >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def initial_condition(p: Params, seed: int, noise: float = 0.1):
    """Seeded vel.out-equivalent start (the ic_file path, initial.f90:226-239).

    u = (1/vonk) ln(z_k/z0) + U(-noise, noise) * ubar(z_k) on levels 1..nz-1;
    v, w get the same-amplitude noise (ubar(z_k) of the uv level k); w(1) =
    w(nz) = 0; u(nz) = u(nz-1), v(nz) = v(nz-1); padding 0; RHS = 0;
    Cs_opt2 = 0.03; F_* = 0.  Level 0 is BOGUS (initial.f90:202-207).
    """
    s = State.zeros(p)
    nx, ny, nz = p.nx, p.ny, p.nz
    k = np.arange(1, nz)[:, None, None]
    j = np.arange(ny)[None, :, None]
    i = np.arange(nx)[None, None, :]
    zk = (k - 0.5) * p.dz
    ubar = (1.0 / p.vonk) * np.log(zk / p.zo)
    amp = noise * ubar
    for fid, name in enumerate(("u", "v", "w")):
        r = uniform01(seed, fid, i, j, k)
        pert = amp * (2.0 * r - 1.0)
        base = ubar if name == "u" else 0.0
        s.f[name][1:nz, :, :nx] = base + pert
    s.w[1] = 0.0
    s.w[nz] = 0.0
    s.u[nz] = s.u[nz - 1]
    s.v[nz] = s.v[nz - 1]
    s.Cs_opt2[1:nz + 1] = 0.03
    for name in ("u", "v", "w"):
        s.f[name][0] = BOGUS
    s.jt = 0
    s.jt_total = 0
    return s


# --------------------------------------------------------------------------
# wallstress.f90
# --------------------------------------------------------------------------
def wallstress(sp: Spectral, s: State):
    """wallstress (wallstress.f90:21-257): lbc_mom=2 at k=1, ubc_mom=0 at nz.

    ws_equilibrium_lbc (171-219): u1,v1 = test_filter(u,v at k=1);
    U = sqrt(u1^2+v1^2); u* = U vonk / ln(0.5 dz/z0); txz = -(u*^2)/U u1;
    dudz = u*/(0.5 dz vonk) u/U (0 where u == 0).  ws_free_ubc (119-128).
    Pinned by: tests/test_oracle_ops.py::test_wallstress_loglaw.
    """
    p = sp.p
    nx, nz = p.nx, p.nz
    u1 = sp.test_filter(s.u[1])
    v1 = sp.test_filter(s.v[1])
    denom = math.log(0.5 * p.dz / p.zo)
    u_avg = np.sqrt(u1[:, :nx] ** 2 + v1[:, :nx] ** 2)
    ustar = u_avg * p.vonk / denom
    s.ustar_lbc = ustar
    const = -(ustar ** 2) / u_avg
    s.txz[1, :, :nx] = const * u1[:, :nx]
    s.tyz[1, :, :nx] = const * v1[:, :nx]
    ub = s.u[1, :, :nx]
    vb = s.v[1, :, :nx]
    dudz = ustar / (0.5 * p.dz * p.vonk) * ub / u_avg
    dvdz = ustar / (0.5 * p.dz * p.vonk) * vb / u_avg
    s.dudz[1, :, :nx] = np.where(ub == 0.0, 0.0, dudz)
    s.dvdz[1, :, :nx] = np.where(vb == 0.0, 0.0, dvdz)
    s.txz[nz] = 0.0
    s.tyz[nz] = 0.0
    s.dudz[nz] = 0.0
    s.dvdz[nz] = 0.0


# --------------------------------------------------------------------------
# sgs_stag_util.f90
# --------------------------------------------------------------------------
def calc_Sij(p: Params, s: State):
    """calc_Sij (sgs_stag_util.f90:467-634), one rank: wall (lbc>0) level 1
    on uvp nodes, stress-free top level nz, w-node averages in 2..nz-1.
    Pinned by: tests/test_oracle_ops.py::test_sij_linear_shear."""
    nx, nz = p.nx, p.nz
    X = np.s_[:, :nx]
    S11, S12, S13, S22, S23, S33 = (s.S11, s.S12, s.S13, s.S22, s.S23, s.S33)
    dudx, dudy, dudz = s.dudx, s.dudy, s.dudz
    dvdx, dvdy, dvdz = s.dvdx, s.dvdy, s.dvdz
    dwdx, dwdy, dwdz = s.dwdx, s.dwdy, s.dwdz
    # k = 1, wall
    S11[1][X] = dudx[1][X]
    S12[1][X] = 0.5 * (dudy[1][X] + dvdx[1][X])
    wx = 0.5 * (dwdx[1][X] + dwdx[2][X])
    S13[1][X] = 0.5 * (dudz[1][X] + wx)
    S22[1][X] = dvdy[1][X]
    wy = 0.5 * (dwdy[1][X] + dwdy[2][X])
    S23[1][X] = 0.5 * (dvdz[1][X] + wy)
    S33[1][X] = dwdz[1][X]
    # k = nz, stress free
    S11[nz][X] = dudx[nz - 1][X]
    S12[nz][X] = 0.5 * (dudy[nz - 1][X] + dvdx[nz - 1][X])
    S13[nz][X] = 0.5 * (dudz[nz][X] + dwdx[nz][X])
    S22[nz][X] = dvdy[nz - 1][X]
    S23[nz][X] = 0.5 * (dvdz[nz][X] + dwdy[nz][X])
    S33[nz][X] = 0.5 * (dwdz[nz - 1][X] + 0.0)
    # k = 2..nz-1
    K = np.s_[2:nz, :, :nx]
    Km = np.s_[1:nz - 1, :, :nx]
    S11[K] = 0.5 * (dudx[K] + dudx[Km])
    uy = dudy[K] + dudy[Km]
    vx = dvdx[K] + dvdx[Km]
    S12[K] = 0.25 * (uy + vx)
    S13[K] = 0.5 * (dudz[K] + dwdx[K])
    S22[K] = 0.5 * (dvdy[K] + dvdy[Km])
    S23[K] = 0.5 * (dvdz[K] + dwdy[K])
    S33[K] = 0.5 * (dwdz[K] + dwdz[Km])


def smag_norm(S11, S12, S13, S22, S23, S33):
    """|S| = sqrt(2(S11^2+S22^2+S33^2+2(S12^2+S13^2+S23^2)))."""
    return np.sqrt(2.0 * (S11 ** 2 + S22 ** 2 + S33 ** 2
                          + 2.0 * (S12 ** 2 + S13 ** 2 + S23 ** 2)))


def sgs_stag(sp: Spectral, s: State):
    """sgs_stag (sgs_stag_util.f90:33-464) for sgs_model = 5, one rank.

    calc_Sij; lagran_dt = cs_count dt; Cs_opt2 = 0.03 at jt = 1 of a fresh
    start (inilag, sgs_stag_util.f90:187-189); lagrange_Sdep when (jt >=
    DYN_init or initu) and mod(jt_total, cs_count) == 0 (192-193);
    Nu_t = |S| Cs^2 delta^2; tau_ij with the wall row (lbc 1:) at k=1 and the
    stress-free top row at nz-1; BOGUS at 0 and nz.
    Pinned by: tests/test_oracle_ops.py::test_sgs_constant_cs.
    """
    p = sp.p
    nx, nz = p.nx, p.nz
    calc_Sij(p, s)
    if p.use_cfl_dt:
        # displacement time since the last dynamic call (sgs_stag_util.f90:74-79)
        if s.jt >= p.dyn_init - p.cs_count + 1 or p.initu:
            s.lagran_dt = s.lagran_dt + s.dt
    else:
        s.lagran_dt = p.cs_count * s.dt
    if s.jt == 1 and not p.initu:
        s.Cs_opt2[:] = 0.03
    elif (s.jt >= p.dyn_init or p.initu) and s.jt_total % p.cs_count == 0:
        lagrange_Sdep(sp, s)
    X = np.s_[:, :nx]
    l2 = p.delta ** 2
    K = np.s_[1:nz + 1, :, :nx]
    S = smag_norm(s.S11[K], s.S12[K], s.S13[K], s.S22[K], s.S23[K], s.S33[K])
    s.Nu_t[K] = S * s.Cs_opt2[K] * l2
    Nu_t = s.Nu_t
    nu = p.nu
    # bottom: wall (lbc_mom >= 1), Sij on uvp nodes at k=1
    const = -2.0 * (Nu_t[1][X] + nu)
    s.txx[1][X] = const * s.S11[1][X]
    s.txy[1][X] = const * s.S12[1][X]
    s.tyy[1][X] = const * s.S22[1][X]
    s.tzz[1][X] = const * s.S33[1][X]
    # top: stress free (ubc_mom = 0)
    t = nz - 1
    const = 0.5 * (Nu_t[t][X] + Nu_t[nz][X]) + nu
    const2 = 2.0 * (Nu_t[t][X] + nu)
    s.txx[t][X] = -const * (s.S11[t][X] + s.S11[nz][X])
    s.txy[t][X] = -const * (s.S12[t][X] + s.S12[nz][X])
    s.tyy[t][X] = -const * (s.S22[t][X] + s.S22[nz][X])
    s.tzz[t][X] = -const * (s.S33[t][X] + s.S33[nz][X])
    s.txz[t][X] = -const2 * s.S13[t][X]
    s.tyz[t][X] = -const2 * s.S23[t][X]
    # interior 2..nz-2
    if nz - 2 >= 2:
        K = np.s_[2:nz - 1, :, :nx]
        Kp = np.s_[3:nz, :, :nx]
        const3 = -2.0 * nu * 0.5
        const4 = -2.0 * nu
        c = -0.5 * (Nu_t[K] + Nu_t[Kp])
        c2 = -2.0 * Nu_t[K]
        s.txx[K] = (c + const3) * (s.S11[K] + s.S11[Kp])
        s.txy[K] = (c + const3) * (s.S12[K] + s.S12[Kp])
        s.tyy[K] = (c + const3) * (s.S22[K] + s.S22[Kp])
        s.tzz[K] = (c + const3) * (s.S33[K] + s.S33[Kp])
        s.txz[K] = (c2 + const4) * s.S13[K]
        s.tyz[K] = (c2 + const4) * s.S23[K]
    for n in ("txx", "txy", "txz", "tyy", "tyz", "tzz"):
        s.f[n][0] = BOGUS
    for n in ("txx", "txy", "tyy", "tzz"):
        s.f[n][nz] = BOGUS


# --------------------------------------------------------------------------
# interpolag_Sdep.f90 + functions.f90:188-265,349-474
# --------------------------------------------------------------------------
def _fmodulo(a, pdiv):
    """Fortran MODULO for reals: result has the sign of the divisor."""
    r = np.fmod(a, pdiv)
    return np.where((r != 0.0) & ((r < 0.0) != (pdiv < 0.0)), r + pdiv, r)


def trilinear_interp_w(p: Params, var, px, py, pz):
    """trilinear_interp_w (functions.f90:349-474) + cell_indx_w (188-265),
    vectorised; ``var`` is the (nz+1, ny, nx) w-grid array (levels 0..nz).

    x/y are wrapped with MODULO (cell_indx_w modifies its argument, so the
    wrapped value is the one used for xdiff), the 1e-9 end thresholds pick
    index 1 or N; near the wall (lbc_mom > 0, pz < zw(2)) the uvp level 1 is
    used with a zero-gradient clamp below z(1); otherwise kstart =
    floor((pz - zw(1))/dz) + 1 with the nz-1 clamp at zw(nz).
    Pinned by: tests/test_oracle_ops.py::test_interp_linear_exact (interior
    cells), tests/test_oracle_lagrange.py::test_interp_near_wall_cells_linear_in_z
    (zero-gradient and half cells at the wall, the lid point) and
    ::test_interp_periodic_seam_matches_numpy_interp (the seam, MODULO wrap).
    """
    nx, ny, nz = p.nx, p.ny, p.nz
    dx, dy, dz = p.dx, p.dy, p.dz
    thresh = 1.0e-9
    px = _fmodulo(px, p.lx)
    py = _fmodulo(py, p.ly)
    i0 = np.where(np.abs(px) / p.lx < thresh, 1,
                  np.where(np.abs(px - p.lx) / p.lx < thresh, nx,
                           np.floor(px / dx).astype(np.int64) + 1))
    j0 = np.where(np.abs(py) / p.ly < thresh, 1,
                  np.where(np.abs(py - p.ly) / p.ly < thresh, ny,
                           np.floor(py / dy).astype(np.int64) + 1))
    i1 = np.where(i0 + 1 > nx, 1, i0 + 1)
    j1 = np.where(j0 + 1 > ny, 1, j0 + 1)
    xg = (i0 - 1) * dx
    yg = (j0 - 1) * dy
    xdiff = px - xg
    ydiff = py - yg
    z1 = (1 - 0.5) * dz
    zw = lambda k: ((k - 0.5) * dz) - dz / 2.0
    near = pz < zw(2)
    below = pz < z1
    kgen = np.where(np.abs(pz - zw(nz)) / p.lz < thresh, nz - 1,
                    np.floor((pz - zw(1)) / dz).astype(np.int64) + 1)
    k0 = np.where(near, 1, kgen)
    k1 = np.where(near, np.where(below, 1, 2), kgen + 1)
    zdiff = np.where(near, np.where(below, 0.0, 2.0 * (pz - z1)),
                     pz - zw(np.where(near, 1, kgen)))
    # guard the gather (the reference would read outside the array here)
    k0c = np.clip(k0, 0, nz)
    k1c = np.clip(k1, 0, nz)
    i0c, i1c, j0c, j1c = i0 - 1, i1 - 1, j0 - 1, j1 - 1
    V = lambda kk, jj, ii: var[kk, jj, ii]
    u1 = V(k0c, j0c, i0c) + xdiff * (V(k0c, j0c, i1c) - V(k0c, j0c, i0c)) / dx
    u2 = V(k0c, j1c, i0c) + xdiff * (V(k0c, j1c, i1c) - V(k0c, j1c, i0c)) / dx
    u3 = V(k1c, j0c, i0c) + xdiff * (V(k1c, j0c, i1c) - V(k1c, j0c, i0c)) / dx
    u4 = V(k1c, j1c, i0c) + xdiff * (V(k1c, j1c, i1c) - V(k1c, j1c, i0c)) / dx
    u5 = u1 + ydiff * (u2 - u1) / dy
    u6 = u3 + ydiff * (u4 - u3) / dy
    return u5 + zdiff * (u6 - u5) / dz


def interpolag_Sdep(p: Params, s: State):
    """interpolag_Sdep (interpolag_Sdep.f90:21-270), one rank, lbc_mom>0,
    ubc_mom=0: departure points x - u lagran_dt (u,v averaged to w nodes,
    0.25 w(2) at the uvp level 1, the w=0 lid at nz) and trilinear gathers
    of F_LM, F_MM, F_QN, F_NN from their pre-step copies."""
    nx, ny, nz = p.nx, p.ny, p.nz
    ldt = s.lagran_dt
    x = (np.arange(nx) * p.dx)[None, None, :]
    y = (np.arange(ny) * p.dy)[None, :, None]
    u, v, w = s.u[:, :, :nx], s.v[:, :, :nx], s.w[:, :, :nx]
    zk = lambda k: (k - 0.5) * p.dz
    zwk = lambda k: zk(k) - p.dz / 2.0
    PX = np.empty((nz, ny, nx))
    PY = np.empty((nz, ny, nx))
    PZ = np.empty((nz, ny, nx))
    # k = 1 (uvp grid)
    PX[0] = x[0] - u[1] * ldt
    PY[0] = y[0] - v[1] * ldt
    PZ[0] = zk(1) - 0.25 * w[2] * ldt
    # 2..nz-1
    if nz > 2:
        PX[1:nz - 1] = x - 0.5 * (u[1:nz - 1] + u[2:nz]) * ldt
        PY[1:nz - 1] = y - 0.5 * (v[1:nz - 1] + v[2:nz]) * ldt
        PZ[1:nz - 1] = zwk(np.arange(2, nz))[:, None, None] - w[2:nz] * ldt
    # k = nz (stress-free lid, w-grid)
    PX[nz - 1] = x[0] - u[nz - 1] * ldt
    PY[nz - 1] = y[0] - v[nz - 1] * ldt
    PZ[nz - 1] = zwk(nz)
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        tmp = s.f[n][:, :, :nx].copy()
        s.f[n][1:nz + 1, :, :nx] = trilinear_interp_w(p, tmp, PX, PY, PZ)


# --------------------------------------------------------------------------
# lagrange_Sdep.f90
# --------------------------------------------------------------------------
def lagrange_Sdep(sp: Spectral, s: State):
    """lagrange_Sdep (lagrange_Sdep.f90:21-432), inflow_type = 0, no
    PPDYN_TN: test-filtered L,Q,M,N at 2 and 4 delta, Lagrangian relaxation
    of F_LM/MM/QN/NN, beta, Cs_opt2.  On a fresh start (inilag) the first
    call at jt = cs_count or DYN_init sets F_MM = MM, F_LM = 0.03 MM (and
    F_NN, F_QN alike), padding columns 1 (266-277, 326-337).  All levels
    1..nz are processed at once (each level's work is independent).
    Pinned by tests/test_oracle_lagrange.py: single-mode fields whose
    filtered products, L, Q, M, N and contractions are known in closed form,
    the u_bar/w_bar staggering, the clips, beta, the relaxation weights
    (epsi = 1 and 1/2) and the first-call initialisation; plus translation
    invariance (tests/test_oracle_ops.py)."""
    p = sp.p
    nz = p.nz
    opftdelta = 1.5 * p.delta
    powcoeff = -1.0 / 8.0
    const = 2.0 * (p.delta ** 2)
    tf1, tf2 = 2.0, 4.0
    tf1_2, tf2_2 = tf1 ** 2, tf2 ** 2
    interpolag_Sdep(p, s)
    u, v, w = s.u, s.v, s.w
    ub = np.empty((nz, p.ny, p.ld))
    vb = np.empty_like(ub)
    wb = np.empty_like(ub)
    ub[0] = u[1]
    vb[0] = v[1]
    wb[0] = 0.25 * w[2]
    ub[1:nz - 1] = 0.5 * (u[2:nz] + u[1:nz - 1])
    vb[1:nz - 1] = 0.5 * (v[2:nz] + v[1:nz - 1])
    wb[1:nz - 1] = w[2:nz]
    ub[nz - 1] = u[nz - 1]
    vb[nz - 1] = v[nz - 1]
    wb[nz - 1] = 0.0
    TF, TTF = sp.test_filter, sp.test_test_filter
    prods = lambda a, b, c: (a * a, a * b, a * c, b * b, b * c, c * c)
    L = prods(ub, vb, wb)
    fb = [TF(q) for q in (ub, vb, wb)]
    fh = [TTF(q) for q in (ub, vb, wb)]
    Lf = [TF(q) for q in L]
    Qf = [TTF(q) for q in L]
    pb = prods(*fb)
    ph = prods(*fh)
    L = [a - b for a, b in zip(Lf, pb)]
    Q = [a - b for a, b in zip(Qf, ph)]
    K = np.s_[1:nz + 1]
    Sij = [s.S11[K], s.S12[K], s.S13[K], s.S22[K], s.S23[K], s.S33[K]]
    S = smag_norm(*Sij)
    Sb = [TF(q) for q in Sij]
    Sh = [TTF(q) for q in Sij]
    S_bar = smag_norm(*Sb)
    S_hat = smag_norm(*Sh)
    SSb = [TF(S * q) for q in Sij]
    SSh = [TTF(S * q) for q in Sij]
    # order of tensors: 11, 12, 13, 22, 23, 33
    M = [const * (a - tf1_2 * S_bar * b) for a, b in zip(SSb, Sb)]
    N = [const * (a - tf2_2 * S_hat * b) for a, b in zip(SSh, Sh)]

    def contract(A, B):
        return (A[0] * B[0] + A[3] * B[3] + A[5] * B[5]
                + 2.0 * (A[1] * B[1] + A[2] * B[2] + A[4] * B[4]))

    def sq(A):
        return (A[0] ** 2 + A[3] ** 2 + A[5] ** 2
                + 2.0 * (A[1] ** 2 + A[2] ** 2 + A[4] ** 2))

    LM, MM, QN, NN = contract(L, M), sq(M), contract(Q, N), sq(N)
    F_LM, F_MM, F_QN, F_NN = (s.F_LM[K], s.F_MM[K], s.F_QN[K], s.F_NN[K])
    first = (not p.initu) and (s.jt == p.cs_count or s.jt == p.dyn_init)
    if first and not s.lm_init:
        F_MM = MM.copy()
        F_LM = 0.03 * MM
        F_MM[:, :, p.ld - 2:] = 1.0
        F_LM[:, :, p.ld - 2:] = 1.0
        s.lm_init = True
    if first and not s.qn_init:
        F_NN = NN.copy()
        F_QN = 0.03 * NN
        F_NN[:, :, p.ld - 2:] = 1.0
        F_QN[:, :, p.ld - 2:] = 1.0
        s.qn_init = True
    zero = ZERO_LAG
    Tn = np.maximum(F_LM * F_MM, zero)
    Tn = opftdelta * (Tn ** powcoeff)
    Tn = np.maximum(zero, Tn)
    dumfac = s.lagran_dt / Tn
    epsi = dumfac / (1.0 + dumfac)
    F_LM = epsi * LM + (1.0 - epsi) * F_LM
    F_MM = epsi * MM + (1.0 - epsi) * F_MM
    F_LM = np.maximum(zero, F_LM)
    cs2d = F_LM / (F_MM + zero)
    cs2d[:, :, p.ld - 2:] = zero
    cs2d = np.maximum(zero, cs2d)
    Tn = np.maximum(F_QN * F_NN, zero)
    Tn = opftdelta * (Tn ** powcoeff)
    Tn = np.maximum(zero, Tn)
    dumfac = s.lagran_dt / Tn
    epsi = dumfac / (1.0 + dumfac)
    F_QN = epsi * QN + (1.0 - epsi) * F_QN
    F_NN = epsi * NN + (1.0 - epsi) * F_NN
    F_QN = np.maximum(zero, F_QN)
    cs4d = F_QN / (F_NN + zero)
    cs4d[:, :, p.ld - 2:] = zero
    cs4d = np.maximum(zero, cs4d)
    expo = math.log(tf1) / (math.log(tf2) - math.log(tf1))
    beta = (cs4d / cs2d) ** expo
    beta[nz - 1] = 1.0
    betaclip = np.maximum(beta, 1.0 / (tf1 * tf2))
    cs = cs2d / betaclip
    cs[:, :, p.ld - 2:] = zero
    cs = np.maximum(zero, cs)
    s.F_LM[K], s.F_MM[K], s.F_QN[K], s.F_NN[K] = F_LM, F_MM, F_QN, F_NN
    s.Cs_opt2[K] = cs
    s.Beta[K] = beta
    s.Tn_all[K] = Tn
    if p.use_cfl_dt:
        s.lagran_dt = 0.0                  # lagrange_Sdep.f90:430


# --------------------------------------------------------------------------
# divstress_uv.f90, divstress_w.f90
# --------------------------------------------------------------------------
def divstress_uv(sp: Spectral, s: State):
    """divstress_uv (divstress_uv.f90:21-85): divtx = ddx txx + ddy txy +
    ddz_w txz, divty = ddx txy + ddy tyy + ddz_w tyz on 1..nz-1, padding 0,
    BOGUS at 0 and nz.  Pinned by: tests/test_oracle_ops.py::test_divstress_trig."""
    p = sp.p
    nz = p.nz
    dtxdx = ddx(sp, s.txx)
    dtzdz = ddz_w(p, s.txz)
    dtydy2 = ddy(sp, s.tyy)
    dtzdz2 = ddz_w(p, s.tyz)
    dtxdx2, dtydy = ddxy(sp, s.txy)
    K = np.s_[1:nz]
    s.divtx[K] = dtxdx[K] + dtydy[K] + dtzdz[K]
    s.divtx[K][:, :, p.ld - 2:] = 0.0
    s.divtx[0] = BOGUS
    s.divtx[nz] = BOGUS
    s.divty[K] = dtxdx2[K] + dtydy2[K] + dtzdz2[K]
    s.divty[K][:, :, p.ld - 2:] = 0.0
    s.divty[0] = BOGUS
    s.divty[nz] = BOGUS


def divstress_w(sp: Spectral, s: State):
    """divstress_w (divstress_w.f90:21-116): divtz = ddx txz + ddy tyz +
    ddz_uv tzz, the z term dropped at k=1 and k=nz; padding of 1..nz-1 = 0."""
    p = sp.p
    nx, nz = p.nx, p.nz
    dtxdx = ddx(sp, s.txz)
    dtydy = ddy(sp, s.tyz)
    dtzdz = ddz_uv(p, s.tzz)
    d = s.divtz
    d[0] = BOGUS
    d[1, :, :nx] = dtxdx[1, :, :nx] + dtydy[1, :, :nx]
    d[nz, :, :nx] = dtxdx[nz, :, :nx] + dtydy[nz, :, :nx]
    d[2:nz, :, :nx] = dtxdx[2:nz, :, :nx] + dtydy[2:nz, :, :nx] + dtzdz[2:nz, :, :nx]
    d[1:nz, :, p.ld - 2:] = 0.0


# --------------------------------------------------------------------------
# convec.f90
# --------------------------------------------------------------------------
def convec(sp: Spectral, s: State):
    """convec (convec.f90:21-334): dealiased u x omega on the 3/2 grid.

    Writes the (positive) convective term into RHSx/y/z on 1..nz-1 (z: 1..nz),
    BOGUS at 0 and at nz for x,y.  sgs = .true. so jzLo = 2; one rank so
    both the wall (k=1) and lid (k=nz-1 / nz) special rows apply.
    Pinned by: tests/test_oracle_ops.py::test_convec_band_limited.
    """
    p = sp.p
    nz = p.nz
    const = 1.0 / (p.nx * p.ny)

    def to_big(f):
        return sp.c2r(sp.padd(sp.r2c(f)))

    u_big = to_big(const * s.u)           # levels 0..nz
    v_big = to_big(const * s.v)
    w_big = to_big(const * s.w)
    v1 = np.empty((nz + 1, p.ny, p.ld))
    v2 = np.empty_like(v1)
    v1[1:] = const * (s.dwdy[1:] - s.dvdz[1:])
    v2[1:] = const * (s.dudz[1:] - s.dwdx[1:])
    v1[1] = const * (0.5 * (s.dwdy[1] + s.dwdy[2]) - s.dvdz[1])
    v2[1] = const * (s.dudz[1] - 0.5 * (s.dwdx[1] + s.dwdx[2]))
    v3 = const * (s.dvdx - s.dudy)
    v1[0] = 0.0
    v2[0] = 0.0
    vort1 = to_big(v1)
    vort2 = to_big(v2)
    vort3 = to_big(v3)
    c = 1.0 / (p.nx2 * p.ny2)
    ccx = np.zeros((nz + 1, p.ny2, p.ld_big))
    ccy = np.zeros_like(ccx)
    ccz = np.zeros_like(ccx)
    ccx[1] = c * (v_big[1] * (-vort3[1]) + 0.5 * w_big[2] * vort2[2])
    ccx[nz - 1] = c * (v_big[nz - 1] * (-vort3[nz - 1])
                       + 0.5 * w_big[nz - 1] * vort2[nz - 1])
    ccy[1] = c * (u_big[1] * vort3[1] + 0.5 * w_big[2] * (-vort1[2]))
    ccy[nz - 1] = c * (u_big[nz - 1] * vort3[nz - 1]
                       + 0.5 * w_big[nz - 1] * (-vort1[nz - 1]))
    if nz - 2 >= 2:
        K = np.s_[2:nz - 1]
        Kp = np.s_[3:nz]
        ccx[K] = c * (v_big[K] * (-vort3[K])
                      + 0.5 * (w_big[Kp] * vort2[Kp] + w_big[K] * vort2[K]))
        ccy[K] = c * (u_big[K] * vort3[K]
                      + 0.5 * (w_big[Kp] * (-vort1[Kp]) + w_big[K] * (-vort1[K])))
    K = np.s_[2:nz]
    Km = np.s_[1:nz - 1]
    ccz[K] = c * 0.5 * ((u_big[K] + u_big[Km]) * (-vort2[K])
                        + (v_big[K] + v_big[Km]) * vort1[K])

    def from_big(cc):
        return sp.c2r(sp.unpadd(_rfft2(cc[..., : p.nx2], axes=(-2, -1))))

    s.RHSx[1:nz] = from_big(ccx[1:nz])
    s.RHSy[1:nz] = from_big(ccy[1:nz])
    s.RHSz[1:nz + 1] = from_big(ccz[1:nz + 1])
    for n in ("RHSx", "RHSy", "RHSz"):
        s.f[n][0] = BOGUS
    s.RHSx[nz] = BOGUS
    s.RHSy[nz] = BOGUS


# --------------------------------------------------------------------------
# press_stag_array.f90 + tridag_array.f90
# --------------------------------------------------------------------------
def tridag_array(p: Params, a, b, c, r):
    """tridag_array (tridag_array.f90:165-246, equal to the PPMPI variant
    at one rank): batched Thomas over (kx, ky) with real a,b,c (nz+1, ny, lh)
    and complex r (nz+1, ny, lh); skips kx = nx/2, ky = ny/2 and (0,0).
    Returns u (complex, rows 0..nz), untouched entries set to 0.
    Pinned by: tests/test_oracle_ops.py::test_tridag_dense."""
    n = p.nz + 1
    lh1 = p.lh - 1
    u = np.zeros_like(r)
    mask = np.ones((p.ny, lh1), dtype=bool)
    mask[p.ny // 2, :] = False
    mask[0, 0] = False
    A = a[:, :, :lh1]
    B = b[:, :, :lh1]
    C = c[:, :, :lh1]
    R = r[:, :, :lh1]
    U = u[:, :, :lh1]
    U[0] = R[0] / B[0]
    bet = B[0].copy()
    gam = np.zeros((n, p.ny, lh1))
    for j in range(1, n):
        g = np.where(mask, C[j - 1] / np.where(mask, bet, 1.0), 0.0)
        gam[j] = g
        bet = np.where(mask, B[j] - A[j] * g, bet)
        U[j] = np.where(mask, (R[j] - A[j] * U[j - 1]) / np.where(mask, bet, 1.0), 0.0)
    for j in range(n - 2, -1, -1):
        U[j] = np.where(mask, U[j] - gam[j + 1] * U[j + 1], U[j])
    return u


def press_stag_array(sp: Spectral, s: State):
    """press_stag_array (press_stag_array.f90:21-290), one rank.

    RHS of the pressure Poisson equation in (kx, ky) space from u*, v*, w*
    scaled by 1/(nx ny tadv1 dt); Neumann rows from divtz(1), divtz(nz);
    Thomas per wavenumber; the (0,0) mode integrated upward from p(0) = 0;
    p, dpdx, dpdy back in physical space, dpdz by differences.
    Pinned by: tests/test_oracle_ops.py::test_projection_divergence_free.
    """
    p = sp.p
    nz, ny, lh = p.nz, p.ny, p.lh
    const = 1.0 / (p.nx * p.ny)
    const2 = const / s.tadv1 / s.dt
    const3 = 1.0 / (p.dz ** 2)
    const4 = 1.0 / p.dz
    Hx = np.zeros((nz + 1, ny, lh), dtype=np.complex128)
    Hy = np.zeros_like(Hx)
    Hz = np.zeros_like(Hx)
    Hx[1:nz] = sp.r2c(const2 * s.u[1:nz])
    Hy[1:nz] = sp.r2c(const2 * s.v[1:nz])
    Hz[1:nz + 1] = sp.r2c(const2 * s.w[1:nz + 1])
    rb = sp.zero_oddballs(sp.r2c(const * s.divtz[1]))
    rt = sp.zero_oddballs(sp.r2c(const * s.divtz[nz]))
    for H in (Hx, Hy, Hz):
        sp.zero_oddballs(H[1:nz])
    sp.zero_oddballs(Hz[nz])
    a = np.zeros((nz + 1, ny, lh))
    b = np.zeros_like(a)
    c = np.zeros_like(a)
    r = np.zeros((nz + 1, ny, lh), dtype=np.complex128)
    a[0] = BOGUS
    b[0] = -1.0
    c[0] = 1.0
    r[0] = -p.dz * rb
    a[nz] = -1.0
    b[nz] = 1.0
    c[nz] = BOGUS
    r[nz] = -p.dz * rt
    kx, ky = sp.kx, sp.ky
    a[1:nz] = const3
    b[1:nz] = -(kx ** 2 + ky ** 2 + 2.0 * const3)
    c[1:nz] = const3
    # row m (p level m, m = 1..nz-1) uses uv level m and w levels m, m+1
    ahx = sp.muli(Hx[1:nz], kx)
    ahy = sp.muli(Hy[1:nz], ky)
    dHz = Hz[2:nz + 1] - Hz[1:nz]
    r[1:nz] = (ahx + ahy) + dHz * const4
    ph = tridag_array(p, a, b, c, r)
    # (0,0) mode (press_stag_array.f90:226-234)
    ph[0, 0, 0] = 0.0
    ph[1, 0, 0] = ph[0, 0, 0] - p.dz * rb[0, 0]
    for k in range(2, nz + 1):
        ph[k, 0, 0] = ph[k - 1, 0, 0] + Hz[k, 0, 0] * p.dz
    sp.zero_oddballs(ph)
    s.f["p"][:] = sp.c2r(ph)
    s.dpdx[1:nz] = sp.c2r(sp.muli(ph[1:nz], kx))
    s.dpdy[1:nz] = sp.c2r(sp.muli(ph[1:nz], ky))
    s.dpdx[nz] = BOGUS
    s.dpdy[nz] = BOGUS
    X = np.s_[:, :, : p.nx]
    s.dpdz[1:nz][X] = (s.f["p"][1:nz][X] - s.f["p"][0:nz - 1][X]) / p.dz
    s.dpdz[nz][:, : p.nx] = (s.f["p"][nz][:, : p.nx] - s.f["p"][nz - 1][:, : p.nx]) / p.dz
    return ph


# --------------------------------------------------------------------------
# forcing.f90::project
# --------------------------------------------------------------------------
def project(p: Params, s: State):
    """project (forcing.f90:149-244): u = u* - dt tadv1 grad p on 1..nz-1
    (w on 2..nz-1), then u(nz) = u(nz-1), v(nz) = v(nz-1), w(nz) = 0,
    w(1) = 0."""
    nx, nz = p.nx, p.nz
    K = np.s_[1:nz, :, :nx]
    s.u[K] = s.u[K] + s.dt * (-s.tadv1 * s.dpdx[K])
    s.v[K] = s.v[K] + s.dt * (-s.tadv1 * s.dpdy[K])
    K = np.s_[2:nz, :, :nx]
    s.w[K] = s.w[K] + s.dt * (-s.tadv1 * s.dpdz[K])
    s.u[nz] = s.u[nz - 1]
    s.v[nz] = s.v[nz - 1]
    s.w[nz] = 0.0
    s.w[1] = 0.0


# --------------------------------------------------------------------------
# Diagnostics (rmsdiv.f90, io.f90::energy, cfl_util.f90, functions.f90)
# --------------------------------------------------------------------------
def rmsdiv(p: Params, s: State):
    """rmsdiv (rmsdiv.f90:21-60): mean |dudx + dvdy + dwdz| over 1..nz-1."""
    K = np.s_[1:p.nz, :, :p.nx]
    return float(np.sum(np.abs(s.dudx[K] + s.dvdy[K] + s.dwdz[K]))
                 / (p.nx * p.ny * (p.nz - 1)))


def energy(p: Params, s: State):
    """energy (io.f90:99-141)."""
    nx, nz = p.nx, p.nz
    tw = 0.5 * (s.w[1:nz, :, :nx] + s.w[2:nz + 1, :, :nx])
    ke = np.sum(s.u[1:nz, :, :nx] ** 2 + s.v[1:nz, :, :nx] ** 2 + tw ** 2)
    return float(ke * 0.5 / (nx * p.ny * (nz - 1)))


def max_cfl(p: Params, s: State):
    """get_max_cfl (cfl_util.f90:35-70)."""
    K = np.s_[1:p.nz, :, :p.nx]
    return s.dt * max(np.max(np.abs(s.u[K])) / p.dx,
                      np.max(np.abs(s.v[K])) / p.dy,
                      np.max(np.abs(s.w[K])) / p.dz)


def get_cfl_dt(p: Params, s: State):
    """get_cfl_dt (cfl_util.f90:73-111): cfl / max(max|u|/dx, max|v|/dy,
    max|w|/dz) over i < nx, levels 1..nz-1."""
    K = np.s_[1:p.nz, :, :p.nx]
    return p.cfl / max(float(np.max(np.abs(s.u[K]))) / p.dx,
                       float(np.max(np.abs(s.v[K]))) / p.dy,
                       float(np.max(np.abs(s.w[K]))) / p.dz)


def cfl_clock(p: Params, s: State):
    """The use_cfl_dt head of the time loop (main.f90:135-144): dt_f = dt,
    dt = get_cfl_dt(), tadv1 = 1 + dt/(2 dt_f), tadv2 = 1 - tadv1.  A clock
    that holds no previous step first gets dt = get_cfl_dt() * huge
    (initialize.f90:191-200), which makes the step Euler (tadv1 = 1,
    tadv2 = 0)."""
    if not s.clock_set:
        s.dt = get_cfl_dt(p, s) * sys.float_info.max
        s.clock_set = True
    dt_f = s.dt
    s.dt = get_cfl_dt(p, s)
    s.tadv1 = 1.0 + 0.5 * s.dt / dt_f
    s.tadv2 = 1.0 - s.tadv1


def tau_wall_bot(p: Params, s: State):
    """get_tau_wall_bot (functions.f90:882-905)."""
    tx = np.sum(s.txz[1, :, :p.nx]) / (p.nx * p.ny)
    ty = np.sum(s.tyz[1, :, :p.nx]) / (p.nx * p.ny)
    return math.sqrt(tx ** 2 + ty ** 2)


def domain_output(p: Params, s: State):
    """inst_write(2) quantities (io.f90:786-884): u, v, w_uv and
    p - 0.5(u^2 + w_uv^2 + v^2) on 1..nz, columns 0..nx-1."""
    nx, nz = p.nx, p.nz
    w = s.w[:, :, :nx]
    w_uv = np.empty((nz + 1, p.ny, nx))
    w_uv[1:nz] = 0.5 * (w[2:nz + 1] + w[1:nz])
    w_uv[nz] = w_uv[nz - 1]
    u = s.u[:, :, :nx]
    v = s.v[:, :, :nx]
    pr = s.f["p"][:, :, :nx] - 0.5 * (u ** 2 + w_uv ** 2 + v ** 2)
    return u[1:], v[1:], w_uv[1:], pr[1:]


# --------------------------------------------------------------------------
# time_average.f90 (tavg) and streamwise spectra
# --------------------------------------------------------------------------
TAVG_NAMES = ("u", "v", "w", "w_uv", "u_w", "v_w", "u2", "v2", "w2", "uv", "uw", "vw",
              "txx", "tyy", "tzz", "txy", "txz", "tyz", "p", "fx", "fy", "fz", "cs_opt2",
              "vortx", "vorty", "vortz")


class Stats:
    """Running time sums of time_average.f90 (``tavg%compute``, 176-319) on
    the nx x ny x (nz+1) grid, driven as io.f90:558-582: between nstart and
    nend every step adds dt to the pending weight; at nstart and every nskip
    steps after it the sums are updated (the first such call only starts the
    summation, as ``tavg%init`` does).  Levels the reference leaves
    unassigned before summing (w_uv, u_w, v_w, vortz at k = 0) are 0 here.

    Spectra (project-defined; the reference's spectra output is not part of
    the build): for u, v, w on levels 1..nz-1 the y-summed power of the
    normalised x-y transform, E(kx, z) = sum_ky |u_hat(kx, ky, z)|^2 for
    kx < nx/2, time-weighted like the sums; ``spectrum`` doubles kx > 0 so
    that sum_kx E = plane mean of u^2 (Parseval; the Nyquist mode excluded).
    Pinned by tests/test_oracle_stats.py."""

    def __init__(self, p: Params, nstart: int, nend: int, nskip: int):
        self.p = p
        self.nstart, self.nend, self.nskip = nstart, nend, nskip
        shp = (p.nz + 1, p.ny, p.nx)
        self.sum = {n: np.zeros(shp) for n in TAVG_NAMES}
        self.spec = np.zeros((3, p.nz + 1, p.nx // 2))
        self.total_time = 0.0
        self.dt = 0.0
        self.initialized = False

    def after_step(self, sp: Spectral, s: State):
        jt = s.jt_total
        if self.nstart <= jt <= self.nend:
            self.dt += s.dt
            if (jt - self.nstart) % self.nskip == 0:
                if not self.initialized:
                    self.initialized = True
                    self.dt = 0.0
                else:
                    self.compute(sp, s)

    def compute(self, sp: Spectral, s: State):
        p = self.p
        nx, nz = p.nx, p.nz
        X = np.s_[:, :, :nx]
        u, v, w = s.u[X], s.v[X], s.w[X]
        w_uv = np.zeros_like(u)
        w_uv[1:nz] = 0.5 * (w[2:nz + 1] + w[1:nz])
        w_uv[nz] = w_uv[nz - 1]
        u_w = np.zeros_like(u)
        v_w = np.zeros_like(u)
        u_w[1:nz + 1] = 0.5 * (u[0:nz] + u[1:nz + 1])
        v_w[1:nz + 1] = 0.5 * (v[0:nz] + v[1:nz + 1])
        u_w[1] = 0.0
        v_w[1] = 0.0
        fz = s.fza[X]
        fza_uv = np.zeros_like(u)
        fza_uv[1:nz] = 0.5 * (fz[2:nz + 1] + fz[1:nz])
        fza_uv[nz] = fza_uv[nz - 1]
        pres = s.f["p"][X] - 0.5 * (u ** 2 + w_uv ** 2 + v ** 2)
        vt = s.dvdx[X] - s.dudy[X]
        vortz = np.zeros_like(u)
        vortz[1:nz + 1] = 0.5 * (vt[0:nz] + vt[1:nz + 1])
        vortz[1] = 0.0
        vortx = s.dwdy[X] - s.dvdz[X]
        vorty = s.dudz[X] - s.dwdx[X]
        dt = self.dt
        q = {"u": u, "v": v, "w": w, "w_uv": w_uv, "u_w": u_w, "v_w": v_w,
             "u2": u * u, "v2": v * v, "w2": w * w, "uv": u * v, "uw": u_w * w,
             "vw": v_w * w, "txx": s.txx[X], "tyy": s.tyy[X], "tzz": s.tzz[X],
             "txy": s.txy[X], "txz": s.txz[X], "tyz": s.tyz[X], "p": pres,
             "vortx": vortx, "vorty": vorty, "vortz": vortz}
        for n, a in q.items():
            self.sum[n] += a * dt
        K = np.s_[1:nz + 1]
        self.sum["fx"][K] += s.fxa[X][K] * dt
        self.sum["fy"][K] += s.fya[X][K] * dt
        self.sum["fz"][K] += fza_uv[K] * dt
        self.sum["cs_opt2"][K] += s.Cs_opt2[X][K] * dt
        c = 1.0 / (nx * p.ny)
        for comp, f in enumerate((s.u, s.v, s.w)):
            F = sp.r2c(f[1:nz] * c)[:, :, :nx // 2].copy()
            F[:, p.ny // 2, :] = 0.0
            self.spec[comp, 1:nz] += dt * np.sum(np.abs(F) ** 2, axis=1)
        self.total_time += dt
        self.dt = 0.0

    def mean(self, name):
        return self.sum[name] / self.total_time

    def profile(self, name):
        return self.mean(name).mean(axis=(1, 2))

    def spectrum(self, comp):
        E = self.spec[comp] / self.total_time
        E[:, 1:] *= 2.0
        return E


# --------------------------------------------------------------------------
# Actuator disks (turbines.f90, turbine_indicator.f90)
# --------------------------------------------------------------------------
@dataclass
class TurbineSpec:
    xloc: float
    yloc: float
    height: float
    dia: float
    thk: float
    theta1: float = 0.0
    theta2: float = 0.0
    ct_prime: float = 1.33


def _linear_interp_table(xt, vt, xq):
    """linear_interp_sa (functions.f90:711-760) with binary_search (811-860),
    vectorised over xq: clamp to the ends, else linear between brackets."""
    N = len(xt)
    xq = np.asarray(xq, dtype=np.float64)
    low = np.clip(np.searchsorted(xt, xq, side="right") - 1, 0, N - 1)
    out = np.empty_like(xq)
    lo = xq < xt[0]
    hi = xq > xt[N - 1]
    mid = ~(lo | hi)
    i = np.minimum(low, N - 2)
    out[lo] = vt[0]
    out[hi] = vt[N - 1]
    im = i[mid]
    xqm = xq[mid]
    out[mid] = vt[im] + (xqm - xt[im]) * (vt[im + 1] - vt[im]) / (xt[im + 1] - xt[im])
    return out


class IndicatorFunction:
    """turb_ind_func_t (turbine_indicator.f90:43-169): Gaussian-filtered disk.

    init: R2(r) is row 0 of the circular convolution (2048^2 r2c/c2r) of a
    Gaussian of width delta2 with the unit-disk indicator, normalised so
    that int 2 pi r R2 dr = 1; M = pi int 2 pi r R2^2 dr.  val(r, x) = R1(x)
    R2(r) / ell^3 with R1 the erf-smoothed slab of thickness thk.
    Pinned by: tests/test_oracle_turbines.py (the table against the exact
    Hankel-transform profile sampled as the reference's circular convolution
    samples it, M against the sampled and the continuum integrals and its
    delta -> 0 limit, the slab's integral and second moment)."""

    def __init__(self, delta1, delta2, thk, dia, N=2048):
        self.ell = 0.5 * dia
        self.delta1 = delta1 / self.ell
        self.delta2 = delta2 / self.ell
        self.thk = thk / self.ell
        self.s6d1 = math.sqrt(6.0) / self.delta1
        L = 1 + 10 * self.delta2 / math.sqrt(12.0)
        d = 2 * L / N
        yz = -L + d * (np.arange(1, N + 1) - 0.5)
        Y2 = yz[:, None] ** 2 + yz[None, :] ** 2
        g = 6.0 / (PI * self.delta2 ** 2) * np.exp(-6 * Y2 / self.delta2 ** 2)
        h = np.where(np.sqrt(Y2) < 1, 1.0 / PI, 0.0)
        fx = np.fft.irfft2(np.fft.rfft2(g) * np.fft.rfft2(h) * d ** 2, s=(N, N),
                           norm="forward")
        self.r = yz[: N // 2] + L
        R2 = fx[0, : N // 2] / N ** 2
        dr = self.r[1] - self.r[0]
        self.R2 = R2 / np.sum(2 * PI * self.r * R2 * dr)
        self.M = float(np.sum(2 * PI * self.r * self.R2 ** 2 * dr) * PI)

    def val(self, r, x):
        rr = np.asarray(r) / self.ell
        xx = np.asarray(x) / self.ell
        R2 = _linear_interp_table(self.r, self.R2, rr)
        from scipy.special import erf
        R1 = 0.5 / self.thk * (erf(self.s6d1 * (xx + 0.5 * self.thk))
                               - erf(self.s6d1 * (xx - 0.5 * self.thk)))
        return R1 * R2 / self.ell ** 3


class Turbines:
    """Actuator-disk model state (turbines.f90:123-638): node lists with
    normalised indicator weights, and u_d_T per turbine.
    Pinned by: tests/test_oracle_turbines.py (node weights against the exact
    indicator at the nodes' coordinates, disk velocity, thrust and total
    force in uniform flow with yaw and tilt, the time filter's weight)."""

    def __init__(self, p: Params, specs, alpha1=1.5, alpha2=1.5,
                 filter_cutoff=1e-2, adm_correction=True, N=2048, T_avg_dim=-1.0,
                 z_i=1.0, u_star=1.0):
        self.p = p
        self.T_avg_dim, self.z_i, self.u_star = T_avg_dim, z_i, u_star
        self.specs = [TurbineSpec(**vars(t)) for t in specs]
        self.adm_correction = adm_correction
        dx, dy, dz = p.dx, p.dy, p.dz
        for t in self.specs:
            t.thk = max(t.thk, dx * 1.01)
        delta1 = alpha1 * math.sqrt(dx ** 2 + dy ** 2 + dz ** 2)
        delta2 = alpha2 * math.sqrt(dx ** 2 + dy ** 2 + dz ** 2)
        cache = {}
        self.ind_funcs = []
        for t in self.specs:
            key = (t.thk, t.dia)
            if key not in cache:
                cache[key] = IndicatorFunction(delta1, delta2, t.thk, t.dia, N)
            self.ind_funcs.append(cache[key])
        self.nodes = []       # list of (i0, j0, k) index arrays (0-based x,y)
        self.ind = []
        self.nhat = []
        nx, ny = p.nx, p.ny
        nz_tot = p.nz_tot
        z_tot = (np.arange(1, nz_tot + 1) - 0.5) * dz
        x = np.arange(nx + 1) * dx
        y = np.arange(ny + 1) * dy
        for t, itf in zip(self.specs, self.ind_funcs):
            search_rad = 0.5 * t.dia + 3 * max(alpha1, alpha2) * math.sqrt(dx ** 2 + dy ** 2 + dz ** 2)
            imax = min(int(search_rad / dx + 2), nx // 2)
            jmax = min(int(search_rad / dy + 2), ny // 2)
            n1 = -math.cos(PI * t.theta1 / 180.) * math.cos(PI * t.theta2 / 180.)
            n2 = -math.sin(PI * t.theta1 / 180.) * math.cos(PI * t.theta2 / 180.)
            n3 = math.sin(PI * t.theta2 / 180.)
            icp = _nint(t.xloc / dx)
            jcp = _nint(t.yloc / dy)
            filt_max = float(itf.val(0.0, 0.0))
            I = np.arange(icp - imax + 1, icp + imax + 1)
            J = np.arange(jcp - jmax + 1, jcp + jmax + 1)
            Kk = np.arange(1, p.nz)                       # k_start..k_end
            i2 = np.mod(I + nx - 1, nx) + 1
            rx = np.where(I < 1, (x[i2 - 1] - p.lx) - t.xloc,
                          np.where(I > nx, (p.lx + x[i2 - 1]) - t.xloc, x[np.clip(I, 1, nx) - 1] - t.xloc))
            j2 = np.mod(J + ny - 1, ny) + 1
            ry = np.where(J < 1, (y[j2 - 1] - p.ly) - t.yloc,
                          np.where(J > ny, (p.ly + y[j2 - 1]) - t.yloc, y[np.clip(J, 1, ny) - 1] - t.yloc))
            rz = z_tot[Kk - 1] - t.height
            RZ, RY, RX = np.meshgrid(rz, ry, rx, indexing="ij")
            KK, JJ, II = np.meshgrid(Kk, j2, i2, indexing="ij")
            r = np.sqrt(RX * RX + RY * RY + RZ * RZ)
            r_norm = np.abs(RX * n1 + RY * n2 + RZ * n3)
            r_disk = np.sqrt(r * r - r_norm * r_norm)
            filt = itf.val(r_disk, r_norm)
            sel = filt > filter_cutoff * filt_max
            sumA = _seq_sum(filt[sel] * dx * dy * dz)
            self.nodes.append((II[sel] - 1, JJ[sel] - 1, KK[sel]))
            self.ind.append(filt[sel] / sumA)
            self.nhat.append((n1, n2, n3))
        self.u_d_T = np.full(len(self.specs), -1.0)
        self.u_d = np.zeros(len(self.specs))
        self.f_n = np.zeros(len(self.specs))
        # disk centre cells (turbines.f90:196-198), 0-based x, y; an index
        # past the grid is taken periodically, the level kept within 1..nz
        self.centers = []
        for t in self.specs:
            ic = (_nint(t.xloc / dx) - 1) % nx
            jc = (_nint(t.yloc / dy) - 1) % ny
            kc = min(max(_nint(t.height / dz + 0.5), 1), p.nz)
            self.centers.append((ic, jc, kc))
        self.center_vel = np.zeros((len(self.specs), 3))

    def copy(self):
        t = object.__new__(Turbines)
        t.__dict__.update(self.__dict__)
        t.u_d_T = self.u_d_T.copy()
        t.u_d = self.u_d.copy()
        t.f_n = self.f_n.copy()
        t.center_vel = self.center_vel.copy()
        return t

    def forcing(self, s: State):
        """turbines_forcing (turbines.f90:465-638), no rotation; the u_d_T
        filter eps = (dt_dim/T)/(1 + dt_dim/T), dt_dim = dt z_i/u_star, for
        T = T_avg_dim > 0, else 1 (561-566).  Ct' is each spec's current one
        (a dynamic schedule sets it between steps, 504-515)."""
        p = self.p
        nz = p.nz
        w_uv = np.empty_like(s.w)
        w_uv[1:nz] = 0.5 * (s.w[2:nz + 1] + s.w[1:nz])
        w_uv[nz] = w_uv[nz - 1]
        s.fxa[:] = 0.0
        s.fya[:] = 0.0
        fza = np.zeros_like(s.fza)
        if self.T_avg_dim > 0.0:
            dt_dim = s.dt * self.z_i / self.u_star
            eps = (dt_dim / self.T_avg_dim) / (1. + dt_dim / self.T_avg_dim)
        else:
            eps = 1.0
        dV = p.dx * p.dy * p.dz
        for n, t in enumerate(self.specs):
            I, J, K = self.nodes[n]
            n1, n2, n3 = self.nhat[n]
            terms = dV * self.ind[n] * (n1 * s.u[K, J, I] + n2 * s.v[K, J, I]
                                        + n3 * w_uv[K, J, I])
            ud = _seq_sum(terms)
            if self.adm_correction:
                ud = ud / (1 + 0.25 * (1 - self.ind_funcs[n].M) * t.ct_prime)
            self.u_d[n] = ud
            # disk centre velocity (turbines.f90:538-547)
            ic, jc, kc = self.centers[n]
            self.center_vel[n] = (s.u[kc, jc, ic], s.v[kc, jc, ic], w_uv[kc, jc, ic])
            self.u_d_T[n] = (1. - eps) * self.u_d_T[n] + eps * ud
            udt = self.u_d_T[n]
            fn = -0.5 * t.ct_prime * abs(udt) * udt * 0.25 * PI * t.dia ** 2
            self.f_n[n] = fn
            s.fxa[K, J, I] = fn * n1 * self.ind[n]
            s.fya[K, J, I] = fn * n2 * self.ind[n]
            fza[K, J, I] = fn * n3 * self.ind[n]
        s.fza[1:nz + 1] = 0.5 * (fza[0:nz] + fza[1:nz + 1])


def _nint(x):
    """Fortran NINT: round half away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def _seq_sum(a):
    """Left-to-right sum (the reference's loop order)."""
    acc = 0.0
    for x in np.asarray(a).ravel():
        acc += float(x)
    return acc


def aligned_farm(p: Params, nxt=4, nyt=4, dia=0.1, hub=0.1, sx_d=7.0, sy_d=5.0,
                 ct_prime=1.33, thk=0.0):
    """A regular aligned array (read_param-style placement, turbines.f90:742-780)
    centred in the domain with streamwise spacing sx_d*D and spanwise
    spacing sy_d*D."""
    x0 = 0.5 * p.lx - 0.5 * (nxt - 1) * sx_d * dia
    y0 = 0.5 * p.ly - 0.5 * (nyt - 1) * sy_d * dia
    specs = []
    for i in range(nxt):
        for j in range(nyt):
            specs.append(TurbineSpec(x0 + i * sx_d * dia, y0 + j * sy_d * dia,
                                     hub, dia, thk, 0.0, 0.0, ct_prime))
    return specs


# --------------------------------------------------------------------------
# One time step (main.f90:130-468)
# --------------------------------------------------------------------------
def step(sp: Spectral, s: State):
    """One pass of time_loop (main.f90:146-344) for the configuration above.

    Returns a dict of the start-of-step diagnostics that depend on the
    derivative fields (rmsdiv uses them, main.f90:367)."""
    p = sp.p
    nz = p.nz
    if p.use_cfl_dt:
        cfl_clock(p, s)
    s.jt += 1
    s.jt_total += 1
    s.total_time = s.total_time + s.dt
    s.RHSx_f[:] = s.RHSx
    s.RHSy_f[:] = s.RHSy
    s.RHSz_f[:] = s.RHSz
    s.u[:], s.dudx[:], s.dudy[:] = filt_da(sp, s.u)
    s.v[:], s.dvdx[:], s.dvdy[:] = filt_da(sp, s.v)
    s.w[:], s.dwdx[:], s.dwdy[:] = filt_da(sp, s.w)
    s.dudz[:] = ddz_uv(p, s.u, s.dudz)
    s.dvdz[:] = ddz_uv(p, s.v, s.dvdz)
    s.dwdz[:] = ddz_w(p, s.w, s.dwdz)
    wallstress(sp, s)
    sgs_stag(sp, s)
    divstress_uv(sp, s)
    divstress_w(sp, s)
    convec(sp, s)
    K = np.s_[1:nz]
    s.RHSx[K] = -s.RHSx[K] - s.divtx[K]
    s.RHSy[K] = -s.RHSy[K] - s.divty[K]
    s.RHSz[K] = -s.RHSz[K] - s.divtz[K]
    s.RHSz[nz] = -s.RHSz[nz] - s.divtz[nz]
    if p.use_mean_p_force:
        s.RHSx[K] = s.RHSx[K] + p.mean_p_force_x
        s.RHSy[K] = s.RHSy[K] + p.mean_p_force_y
    if s.turbines is not None:
        s.turbines.forcing(s)
        s.RHSx[K] = s.RHSx[K] + s.fxa[K]
        s.RHSy[K] = s.RHSy[K] + s.fya[K]
        s.RHSz[K] = s.RHSz[K] + s.fza[K]
    if s.jt_total == 1 and not p.initu:
        # Euler first step of a fresh start (main.f90:273-280)
        s.RHSx_f[:] = s.RHSx
        s.RHSy_f[:] = s.RHSy
        s.RHSz_f[:] = s.RHSz
    for n in ("x", "y", "z"):
        vel = {"x": s.u, "y": s.v, "z": s.w}[n]
        R, Rf = s.f["RHS" + n], s.f["RHS" + n + "_f"]
        vel[K] = vel[K] + s.dt * (s.tadv1 * R[K] + s.tadv2 * Rf[K])
    s.w[nz] = s.w[nz] + s.dt * (s.tadv1 * s.RHSz[nz] + s.tadv2 * s.RHSz_f[nz])
    s.u[0] = BOGUS
    s.v[0] = BOGUS
    s.w[0] = BOGUS
    s.u[nz] = BOGUS
    s.v[nz] = BOGUS
    press_stag_array(sp, s)
    s.RHSx[K] = s.RHSx[K] - s.dpdx[K]
    s.RHSy[K] = s.RHSy[K] - s.dpdy[K]
    s.RHSz[K] = s.RHSz[K] - s.dpdz[K]
    s.RHSz[nz] = s.RHSz[nz] - s.dpdz[nz]
    project(p, s)


def run(p: Params, nsteps: int, seed: int = 0, turbines=None, every: int = 10,
        noise: float = 0.1):
    """Run ``nsteps`` from the seeded IC; returns (state, snapshots) where
    snapshots[jt] = (u, v, w_uv, p_real, rmsdiv) every ``every`` steps."""
    sp = Spectral(p)
    s = initial_condition(p, seed, noise)
    if turbines is not None:
        s.turbines = Turbines(p, turbines)
    snaps = {}
    for _ in range(nsteps):
        step(sp, s)
        if every and s.jt_total % every == 0:
            snaps[s.jt_total] = domain_output(p, s) + (rmsdiv(p, s),)
    return s, snaps
