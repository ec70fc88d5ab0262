"""Closed-form pins of the oracle's actuator-disk model: the indicator
function (turbine_indicator.f90:66-168), the node weights
(turbines.f90:275-462) and the disk forcing (turbines.f90:517-583).

Indicator.  R2 is the unit disk indicator h = 1/pi (|y| < 1) convolved with
the Gaussian g = 6/(pi delta^2) exp(-6 |y|^2 / delta^2), both of unit
integral, so its Hankel transform is exact:

    R2(r) = (1/pi) int_0^inf exp(-k^2 delta^2 / 24) J1(k) J0(k r) dk,
    M = pi int 2 pi r R2^2 dr = 2 int_0^inf exp(-k^2 delta^2 / 12) J1(k)^2 / k dk,

with M -> 1 as delta -> 0.  The reference evaluates the convolution as a
circular one on an N = 2048 grid of spacing d centred between cells, whose
row 1 then holds the convolution at the displacement (d, (b + 1) d) while
the table's abscissa is r_b = (b + 1/2) d; the table is normalised with the
Riemann sum of 2 pi r R2 dr.  The oracle's table is checked against the
exact profile sampled the same way (a profile without the displacement
differs by 2e-3, the disk's staircase on the grid accounts for 1.4e-4).
R1, the erf slab of thickness thk, is a box of width thk convolved with a
Gaussian of variance delta1^2 / 12: unit integral, second moment
(thk^2 + delta1^2) / 12.

Forcing.  In uniform flow the disk velocity is n.U exactly (the node
weights integrate to 1), divided by 1 + (1 - M) Ct' / 4 with the ADM
correction, and the disks' total force is -Ct'/2 |u_d| u_d (pi D^2 / 4) n,
n = (-cos th1 cos th2, -sin th1 cos th2, sin th2).  With T_avg_dim equal to
dt z_i / u_star the time filter's weight is exactly 1/2."""
import math

import numpy as np
import pytest
from scipy import integrate, special

from oracle import lesgo_oracle as o

DELTA = 0.837          # the 256 x 256 x 192 wind farm's dimensionless delta2


def r2_exact(r, delta):
    f = lambda k: math.exp(-k * k * delta * delta / 24.0) * special.j1(k) * special.j0(k * r)
    v, _ = integrate.quad(f, 0.0, math.sqrt(24.0 * 40.0) / delta, limit=2000)
    return v / math.pi


def m_exact(delta):
    f = lambda k: math.exp(-k * k * delta * delta / 12.0) * special.j1(k) ** 2 / k
    v, _ = integrate.quad(f, 0.0, math.sqrt(12.0 * 40.0) / delta, limit=4000)
    return 2.0 * v


@pytest.fixture(scope="module")
def table():
    # dia = 2: ell = 1, so delta1 = delta2 = DELTA dimensionless
    itf = o.IndicatorFunction(DELTA, DELTA, 0.5, 2.0)
    d = itf.r[1] - itf.r[0]
    prof = np.array([r2_exact(math.hypot(d, (b + 1) * d), DELTA) for b in range(len(itf.r))])
    return itf, d, prof


def test_indicator_profile_is_the_gaussian_filtered_disk(table):
    itf, d, prof = table
    pred = prof / np.sum(2.0 * math.pi * itf.r * prof * d)
    err = np.max(np.abs(itf.R2 - pred)) / pred.max()
    print(f"  R2 table vs exact profile: {err:.2e}")
    assert err < 5e-4
    # the continuous profile itself integrates to 1 and has the closed-form
    # centre value (1 - exp(-6 / delta^2)) / pi
    r = np.linspace(0.0, 6.0, 6001)
    dense = np.array([r2_exact(x, DELTA) for x in r[::50]])
    assert abs(dense[0] - (1.0 - math.exp(-6.0 / DELTA ** 2)) / math.pi) < 1e-10
    assert abs(np.trapezoid(2 * math.pi * r[::50] * dense, r[::50]) - 1.0) < 2e-3


def test_indicator_correction_factor(table):
    itf, d, prof = table
    pred = prof / np.sum(2.0 * math.pi * itf.r * prof * d)
    m_pred = math.pi * np.sum(2.0 * math.pi * itf.r * pred ** 2 * d)
    print(f"  M oracle {itf.M:.6f}  sampled exact {m_pred:.6f}  continuum {m_exact(DELTA):.6f}")
    assert abs(itf.M - m_pred) < 5e-4 * m_pred
    assert abs(itf.M - m_exact(DELTA)) < 5e-3 * m_exact(DELTA)
    # delta -> 0: the filtered disk becomes the disk, M -> 1
    sharp = o.IndicatorFunction(0.05, 0.05, 0.5, 2.0)
    assert abs(m_exact(0.05) - 1.0) < 0.03
    assert abs(sharp.M - m_exact(0.05)) < 2e-3


def test_indicator_slab_moments():
    thk, delta1, dia = 0.3, 0.4, 2.0
    itf = o.IndicatorFunction(delta1, DELTA, thk, dia)
    x = np.linspace(-4.0, 4.0, 160001)
    v = itf.val(np.zeros_like(x), x)
    r1 = v / float(itf.val(0.0, 0.0))         # R1(x) / R1(0)
    area = np.trapezoid(r1, x)
    m2 = np.trapezoid(x * x * r1, x) / area
    # R1 has unit integral: R1(0) = 1 / area
    r10 = 0.5 / thk * (math.erf(math.sqrt(6.0) / delta1 * 0.5 * thk) * 2.0)
    assert abs(1.0 / area - r10) < 1e-9
    assert abs(m2 - (thk ** 2 + delta1 ** 2) / 12.0) < 1e-9


def _farm(p, th1, th2, ct=1.33):
    return [o.TurbineSpec(0.5 * p.lx, 0.5 * p.ly, 0.4, 0.3, 0.0, th1, th2, ct)]


def test_node_weights_follow_the_exact_indicator():
    """Each node's weight is the indicator at its distance from the disk
    plane and axis, from the node's own coordinates x = (i-1) dx,
    y = (j-1) dy, z = (k - 1/2) dz (periodic images in x, y), normalised so
    that the weights integrate to 1."""
    p = o.Params(48, 32, 24)
    th1, th2 = 20.0, 10.0
    tb = o.Turbines(p, _farm(p, th1, th2))
    t = tb.specs[0]
    I, J, K = tb.nodes[0]
    dV = p.dx * p.dy * p.dz
    assert abs(np.sum(tb.ind[0]) * dV - 1.0) < 1e-12
    n = np.array(tb.nhat[0])
    np.testing.assert_allclose(n, [-math.cos(math.radians(th1)) * math.cos(math.radians(th2)),
                                   -math.sin(math.radians(th1)) * math.cos(math.radians(th2)),
                                   math.sin(math.radians(th2))], rtol=1e-14)
    rx = (I * p.dx - t.xloc + 0.5 * p.lx) % p.lx - 0.5 * p.lx
    ry = (J * p.dy - t.yloc + 0.5 * p.ly) % p.ly - 0.5 * p.ly
    rz = (K - 0.5) * p.dz - t.height
    rn = np.abs(rx * n[0] + ry * n[1] + rz * n[2])
    rd = np.sqrt(np.maximum(rx * rx + ry * ry + rz * rz - rn * rn, 0.0))
    itf = tb.ind_funcs[0]
    ell = 0.5 * t.dia
    s6 = math.sqrt(6.0) / itf.delta1
    r1 = 0.5 / itf.thk * (special.erf(s6 * (rn / ell + 0.5 * itf.thk))
                          - special.erf(s6 * (rn / ell - 0.5 * itf.thk)))
    r2 = np.array([r2_exact(x, itf.delta2) for x in rd / ell])
    want = r1 * r2
    want = want / (np.sum(want) * dV)
    err = np.max(np.abs(tb.ind[0] - want)) / want.max()
    print(f"  {len(I)} nodes, weight vs exact indicator: {err:.2e}")
    assert err < 1e-2


def _uniform_state(p, U, V, W):
    s = o.State.zeros(p)
    s.u[1:p.nz + 1] = U
    s.v[1:p.nz + 1] = V
    s.w[1:p.nz + 1] = W
    return s


@pytest.mark.parametrize("th1,th2", [(0.0, 0.0), (20.0, 10.0)])
def test_forcing_in_uniform_flow(th1, th2):
    p = o.Params(48, 32, 24)
    U, V, W = 1.7, -0.4, 0.3
    n = np.array([-math.cos(math.radians(th1)) * math.cos(math.radians(th2)),
                  -math.sin(math.radians(th1)) * math.cos(math.radians(th2)),
                  math.sin(math.radians(th2))])
    un = float(n @ [U, V, W])
    ct, dia = 1.33, 0.3
    area = 0.25 * math.pi * dia ** 2
    dV = p.dx * p.dy * p.dz
    # without the ADM correction u_d is n.U
    tb = o.Turbines(p, _farm(p, th1, th2), adm_correction=False)
    s = _uniform_state(p, U, V, W)
    tb.forcing(s)
    assert abs(tb.u_d[0] - un) < 1e-12 * abs(un)
    # with it, divided by 1 + (1 - M) Ct' / 4, M from the continuum integral
    tb = o.Turbines(p, _farm(p, th1, th2))
    s = _uniform_state(p, U, V, W)
    tb.forcing(s)
    m = m_exact(tb.ind_funcs[0].delta2)
    ud = un / (1.0 + 0.25 * (1.0 - m) * ct)
    assert abs(tb.u_d[0] - ud) < 2e-3 * abs(ud)
    # the total force is the ADM thrust along -n... with u_d < 0 along +n
    fn = -0.5 * ct * abs(tb.u_d[0]) * tb.u_d[0] * area
    assert abs(tb.f_n[0] - fn) < 1e-12 * abs(fn)
    K = np.s_[1:p.nz]
    fx = np.sum(s.fxa[K][:, :, :p.nx]) * dV
    fy = np.sum(s.fya[K][:, :, :p.nx]) * dV
    fz = np.sum(s.fza[1:p.nz + 1][:, :, :p.nx]) * dV
    np.testing.assert_allclose([fx, fy, fz], fn * n, rtol=1e-10, atol=1e-12 * abs(fn))
    assert fx < 0.0                      # drag opposes the streamwise flow


def test_time_filter_weight_half():
    """T_avg_dim = dt z_i / u_star: eps = 1/2, so from u_d_T = -1 two
    forcings give u_d_T = (u_d - 1)/2, then (3 u_d - 1)/4."""
    p = o.Params(48, 32, 24)
    tb = o.Turbines(p, _farm(p, 0.0, 0.0), T_avg_dim=p.dt, z_i=1.0, u_star=1.0)
    s = _uniform_state(p, 1.2, 0.0, 0.0)
    s.dt = p.dt
    tb.forcing(s)
    ud = tb.u_d[0]
    assert abs(tb.u_d_T[0] - 0.5 * (ud - 1.0)) < 1e-14
    tb.forcing(s)
    assert abs(tb.u_d_T[0] - 0.25 * (3.0 * ud - 1.0)) < 1e-14
