"""Pins the oracle's fresh-start branch (initu = .false., inilag = .true.):
Euler first step (main.f90:273-280), Cs_opt2 = 0.03 Smagorinsky phase until
DYN_init (sgs_stag_util.f90:187-193) and the F_* initialisation of the first
dynamic call (lagrange_Sdep.f90:266-277, 326-337)."""
import numpy as np

from oracle import lesgo_oracle as o


def uniform_state(p, U=5.0):
    """u = U on the uv levels 1..nz (u(nz) = u(nz-1)), v = w = 0,
    Cs_opt2 = 0.03: plane-uniform, so the step has a closed form away from
    the wall."""
    s = o.State.zeros(p)
    s.u[1:p.nz + 1, :, :p.nx] = U
    s.u[0] = o.BOGUS
    s.v[0] = o.BOGUS
    s.w[0] = o.BOGUS
    s.Cs_opt2[1:] = 0.03
    return s


def test_fresh_first_step_is_euler():
    """Interior levels of a plane-uniform flow see only the mean pressure
    gradient f (no shear, no vorticity, no pressure gradient in x, y):
    Euler gives u = U + dt f, AB2 from RHS_f = 0 gives U + 1.5 dt f."""
    for initu, fac in ((False, 1.0), (True, 1.5)):
        p = o.Params(16, 12, 10, initu=initu, mean_p_force_x=0.8)
        sp = o.Spectral(p)
        s = uniform_state(p)
        o.step(sp, s)
        K = np.s_[3:p.nz - 2, :, :p.nx]
        np.testing.assert_allclose(s.u[K], 5.0 + fac * p.dt * 0.8, rtol=0, atol=1e-12)
        np.testing.assert_allclose(s.v[K], 0.0, atol=1e-12)
        if not initu:
            # RHS_f holds the step's RHS before the pressure gradient came off
            J = np.s_[1:p.nz, :, :p.nx]
            np.testing.assert_allclose(s.RHSx_f[J], s.RHSx[J] + s.dpdx[J], rtol=0, atol=1e-12)


def test_fresh_smagorinsky_until_dyn_init():
    p = o.Params(16, 12, 10, initu=False, dyn_init=10)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 3)
    s.Cs_opt2[:] = 0.2              # overwritten at jt = 1
    for _ in range(9):
        o.step(sp, s)
        assert (s.Cs_opt2 == 0.03).all()
        assert not s.F_LM.any() and not s.F_NN.any() and not s.lm_init
    o.step(sp, s)
    assert s.lm_init and s.qn_init
    K = np.s_[1:p.nz + 1, :, :p.nx]
    assert (s.F_MM[K] > 0).any() and not (s.Cs_opt2[K] == 0.03).all()


def test_fresh_initialisation_gives_cs_003_as_dt_vanishes():
    """With F_MM = MM, F_LM = 0.03 MM (and F_NN, F_QN alike) and a vanishing
    Lagrangian weight, both scale coefficients stay 0.03, beta = 1 and
    Cs_opt2 = 0.03; the file start from F_* = 0 takes LM/MM instead."""
    out = {}
    for initu in (False, True):
        p = o.Params(16, 12, 10, initu=initu, dyn_init=5, dt=1e-9)
        sp = o.Spectral(p)
        s = o.initial_condition(p, 4)
        for _ in range(5):
            o.step(sp, s)
        out[initu] = s.Cs_opt2[1:p.nz, :, :p.nx].copy()
    np.testing.assert_allclose(out[False], 0.03, rtol=1e-6)
    assert np.median(np.abs(out[True] / 0.03 - 1.0)) > 0.1
