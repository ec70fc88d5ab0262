"""Whole runs from lesgo.conf (lesgo_amd.run) and the turbine controls they
use, against the oracle: the u_d_T time filter (T_avg_dim > 0), dynamic
Ct' and yaw schedules from input_turbines/*.dat, domain output, time
averages, and a restart from the run's own checkpoint files."""
import os

import numpy as np
import pytest

from lesgo_amd import samples as sm
from oracle import lesgo_oracle as o
from gpu_util import gpu_case, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfl", [False, True])
def test_time_filter_and_ct_update(cfl):
    """2x2 disks, T_avg_dim = 3 (dimensional, z_i = 1000, u_star = 0.45):
    u_d_T, thrust and the flow after 8 steps with Ct' changed on the device
    after step 4 match the oracle at 1e-10 (fixed and CFL-adaptive dt)."""
    p = o.Params(48, 32, 16, use_cfl_dt=cfl, cfl=0.05)
    specs = o.aligned_farm(p, 2, 2, dia=0.3, hub=0.3, sx_d=7, sy_d=5)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 2)
    s.turbines = o.Turbines(p, specs, T_avg_dim=3.0, z_i=1000.0, u_star=0.45)
    new_ct = [1.1, 1.5, 0.9, 1.33]
    with gpu_case(p) as c:
        for n, a in s.f.items():
            c.upload(n, a)
        c.set_counters(0, 0)
        c.turbines_init([(t.xloc, t.yloc, t.height, t.dia, t.thk, t.theta1, t.theta2, t.ct_prime)
                         for t in specs])
        c.turbines_time_filter(3.0, 1000.0, 0.45)
        for k in range(8):
            if k == 4:
                for t, ct in zip(s.turbines.specs, new_ct):
                    t.ct_prime = ct
                c.turbines_set_ct(new_ct)
            o.step(sp, s)
            c.step(1)
        udt, fn = c.turbines_state()
        np.testing.assert_allclose(udt, s.turbines.u_d_T, rtol=1e-10)
        np.testing.assert_allclose(fn, s.turbines.f_n, rtol=1e-10)
        K = np.s_[1:p.nz, :, :p.nx]
        for n in ("u", "v", "w", "p"):
            e = rel_l2(c.download(n)[K], s.f[n][K])
            assert e <= 1e-10, (n, e)


CONF = """
DOMAIN {{
  Nx = 32
  Ny = 32
  Nz = 16
  z_i = 1000.0
  Lx = 6.283185307179586
  Ly = 3.141592653589793
  Lz = 1.0
}}
MODEL {{
  cs_count = 5
  dyn_init = 5
  u_star = 0.45
}}
TIME {{
  nsteps = {nsteps}
  dt = 2.0e-4
}}
FLOW_COND {{
  lbc_mom = 2
  zo = 0.0001
  mean_p_force_x = 1.0
}}
OUTPUT {{
  wbase = 6
  nenergy = 3
  checkpoint_data = .true.
  checkpoint_nskip = 6
  domain_calc = .true.
  domain_nstart = 6
  domain_nskip = 6
  tavg_calc = {tavg}
  tavg_nstart = 2
  tavg_nskip = 1
  point_calc = .true.
  point_nstart = 12
  point_nskip = 6
  point_loc = 1.0, 0.5, 0.3 // 7.0, -0.4, 0.61
  xplane_calc = .true.
  xplane_nstart = 12
  xplane_nskip = 6
  xplane_loc = 1.3
  yplane_calc = .true.
  yplane_nstart = 12
  yplane_nskip = 6
  yplane_loc = 0.77
  zplane_calc = .true.
  zplane_nstart = 12
  zplane_nskip = 6
  zplane_loc = 0.42
}}
TURBINES {{
  num_x = 2
  num_y = 2
  orientation = 1
  dia_all = 300
  height_all = 300
  thk_all = 0
  theta1_all = 0
  theta2_all = 0
  Ct_prime = 1.33
  read_param = .false.
  dyn_theta1 = .true.
  dyn_Ct_prime = .true.
  T_avg_dim = -1
  alpha1 = 1.5
  alpha2 = 1.5
  filter_cutoff = 1e-2
  adm_correction = .true.
  tbase = 3
}}
"""


def test_run_from_conf_with_schedules_and_restart(tmp_path):
    """A 12-step fresh run from lesgo.conf with Ct' and yaw schedules
    (input_turbines/Ct_prime.dat, theta1.dat at dimensional times), then a
    6-step restart from its vel.out / total_time.dat / u_d_T.dat: the domain
    output (velocity, vorticity) at steps 6, 12, 18, the point and plane
    samples at step 12, the per-turbine histories every 3 steps, the time
    averages (continued through tavg.out by the restart: the sums go on from
    the file, the pressure sum from zero as the file has none, and the first
    summation step after the restart only starts the summation) and
    check_ke.dat match the oracle
    driven through the same schedules, at 1e-10 (the samples are the host
    writers of lesgo_amd.samples applied to the oracle's fields)."""
    from lesgo_amd.conf import place_turbines, read_conf
    from lesgo_amd.run import Runner
    os.makedirs(tmp_path / "input_turbines")
    dt_dim = 2.0e-4 * 1000.0 / 0.45
    t_ct = [0.0, 3 * dt_dim, 9 * dt_dim]
    ct = [[1.33, 1.0, 1.6], [1.33, 1.2, 1.2], [1.33, 1.5, 0.8], [1.33, 0.9, 1.4]]
    t_th = [0.0, 7.5 * dt_dim, 100.0]
    th = [[0.0, 0.0, 0.0], [0.0, 6.0, 6.0], [0.0, -4.0, -4.0], [0.0, 0.0, 0.0]]
    for name, tt, vals in (("Ct_prime", t_ct, ct), ("theta1", t_th, th)):
        with open(tmp_path / "input_turbines" / f"{name}.dat", "w") as f:
            for q, t in enumerate(tt):
                f.write(", ".join([repr(t)] + [repr(v[q]) for v in vals]) + "\n")
    with open(tmp_path / "lesgo.conf", "w") as f:
        f.write(CONF.format(nsteps=12, tavg=".true."))
    r = Runner(str(tmp_path), quiet=True)
    r.run()
    r.close()
    # the oracle through the same schedule
    conf = read_conf(str(tmp_path) + "/")
    rows = place_turbines(conf)
    p = o.Params(32, 32, 16, initu=False, dyn_init=5)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 0)
    specs = [o.TurbineSpec(*row) for row in rows]
    s.turbines = o.Turbines(p, specs)
    st = o.Stats(p, 2, 50000, 1)
    t_dim, ke = 0.0, []
    hist = [[] for _ in specs]       # turbines.f90:589-593 records, tbase = 3

    def record_history():
        tb = s.turbines
        for q, t in enumerate(tb.specs):
            hist[q].append([t_dim, *tb.center_vel[q], -tb.u_d[q], -tb.u_d_T[q], t.theta1,
                            t.theta2, t.ct_prime])

    def check_history():
        for q in range(len(specs)):
            got = np.loadtxt(tmp_path / "turbine" / f"turbine_{q + 1}.dat", ndmin=2)
            ref = np.asarray(hist[q])
            assert got.shape == ref.shape, (q, got.shape, ref.shape)
            np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-13)

    def oracle_step():
        nonlocal t_dim
        t_dim = t_dim + p.dt * 1000.0 / 0.45
        new = [o.TurbineSpec(**vars(t)) for t in s.turbines.specs]
        for k, t in enumerate(new):
            t.ct_prime = float(o._linear_interp_table(np.asarray(t_ct), np.asarray(ct[k]), t_dim))
            t.theta1 = float(o._linear_interp_table(np.asarray(t_th), np.asarray(th[k]), t_dim))
        if any(a.theta1 != b.theta1 for a, b in zip(new, s.turbines.specs)):
            udt = s.turbines.u_d_T.copy()
            s.turbines = o.Turbines(p, new)
            s.turbines.u_d_T[:] = udt
        else:
            for a, b in zip(s.turbines.specs, new):
                a.ct_prime = b.ct_prime
        o.step(sp, s)

    def check_domain(jt):
        ref = o.domain_output(p, s)
        n = 32 * 32 * 17
        raw = np.fromfile(tmp_path / "output" / f"vel.{jt}.c0.bin", dtype=np.float64)
        for q, name in enumerate(("u", "v", "w_uv")):
            g = raw[q * n:(q + 1) * n].reshape(17, 32, 32)
            e = rel_l2(g[:16], ref[q][:16])
            assert e <= 1e-10, (jt, name, e)
        # vorticity (io.f90:823-858) from the oracle's derivative fields; file
        # record level k is array level k - 1; levels 2..nz-1 compared, level
        # 1 of vortz is zero
        der = {nm: s.f[nm][:, :, :32] for nm in ("dudy", "dudz", "dvdx", "dvdz", "dwdx", "dwdy")}
        vref = sm.vorticity(der, p.nz)
        raw = np.fromfile(tmp_path / "output" / f"vort.{jt}.c0.bin", dtype=np.float64)
        for q, name in enumerate(("vortx", "vorty", "vortz")):
            g = raw[q * n:(q + 1) * n].reshape(17, 32, 32)
            e = rel_l2(g[1:16], vref[q][2:17])
            assert e <= 1e-10, (jt, name, e)
        assert not raw[2 * n:2 * n + 32 * 32].any()

    def check_samples(jt):
        g = sm.Grid(p.nx, p.ny, p.nz, conf.L_x, conf.L_y, conf.L_z, conf.dx, conf.dy, conf.dz)
        u, v, w = s.f["u"], s.f["v"], s.f["w"]
        wu = sm.w_on_uv(w, p.nz)
        out = tmp_path / "output"
        for pt, name in (((1.0, 0.5, 0.3), "vel.x-1.00000.y-0.50000.z-0.30000.dat"),
                         ((7.0 - 2 * np.pi, np.pi - 0.4, 0.61), "vel.x-0.71681.y-2.74159.z-0.61000.dat")):
            rec = np.loadtxt(out / name, ndmin=2)[0]
            assert abs(rec[0] - jt * 2.0e-4) <= 1e-12
            ref = [sm.trilinear(a, g, pt) for a in (u, v, wu)]
            np.testing.assert_allclose(rec[1:], ref, rtol=1e-10, atol=1e-12)
        for key, loc, fn, shape in (("x", 1.3, sm.x_plane, (17, 32)), ("y", 0.77, sm.y_plane, (17, 32)),
                                    ("z", 0.42, sm.z_plane, (32, 32))):
            raw = np.fromfile(out / f"vel.{key}-{sm.fmt_real(loc)}.{jt}.c0.bin", dtype=np.float64)
            m = shape[0] * shape[1]
            assert raw.size == 3 * m
            for q, a in enumerate((u, v, wu)):
                e = rel_l2(raw[q * m:(q + 1) * m].reshape(shape), fn(a, g, loc))
                assert e <= 1e-10, (key, q, e)

    for jt in range(1, 13):
        oracle_step()
        st.after_step(sp, s)
        if jt % 3 == 0:
            ke.append(o.energy(p, s))
        if jt % 6 == 0:
            check_domain(jt)
        if jt == 12:
            check_samples(jt)
        if jt % 3 == 0:
            record_history()
    check_history()
    kd = np.loadtxt(tmp_path / "output" / "check_ke.dat")
    np.testing.assert_allclose(kd[:, 1], ke, rtol=1e-10)
    n = 32 * 32 * 17
    avg = np.fromfile(tmp_path / "output" / "veluv_avg.bin", dtype=np.float64)
    e = rel_l2(avg[:n].reshape(17, 32, 32)[:16], st.mean("u")[1:17])
    assert e <= 1e-10, e
    assert os.path.exists(tmp_path / "output" / "lesgo_param.out")
    # restart for 6 more steps
    assert os.path.exists(tmp_path / "tavg.out")
    with open(tmp_path / "lesgo.conf", "w") as f:
        f.write(CONF.format(nsteps=18, tavg=".true."))
    r = Runner(str(tmp_path), quiet=True)
    assert r.restart and r.case.counters()[1] == 12
    r.run()
    r.close()
    st2 = o.Stats(p, 2, 50000, 1)       # tavg%init from tavg.out at step 13
    st2.sum = {k: (np.zeros_like(a) if k == "p" else a.copy()) for k, a in st.sum.items()}
    st2.total_time = st.total_time
    for jt in range(13, 19):
        oracle_step()
        st2.after_step(sp, s)
        if jt % 3 == 0:
            record_history()
    check_domain(18)
    check_history()
    assert st2.total_time > st.total_time
    for q, (fname, name) in enumerate((("veluv_avg.bin", "u"), ("vel2_avg.bin", "u2"),
                                       ("pres_avg.bin", "p"))):
        avg = np.fromfile(tmp_path / "output" / fname, dtype=np.float64)
        e = rel_l2(avg[:n].reshape(17, 32, 32)[:16], st2.mean(name)[1:17])
        assert e <= 1e-10, (fname, e)
    w = open(tmp_path / "total_time.dat").read().split()
    assert int(w[0]) == 18
