/*
 * lesgo_amd C ABI: MI355X (gfx950) implementation of one lesgo LES time step.
 *
 * The reference's time step is a set of Fortran subroutines acting on module
 * globals (sim_param.f90:31-93, sgs_param.f90:29-184, param.f90).  This ABI
 * keeps that shape: a case handle owns device-resident copies of those module
 * arrays, `lg_upload`/`lg_download` move them to and from host memory in the
 * reference's own layout (Fortran (ld, ny, 0:nz) column-major == C
 * [nz+1][ny][ld]), and one entry point per replaced subroutine operates on
 * the device state.  Every function returns 0 on success or a positive
 * LG_ERR_* code; lg_last_error() returns a message (the reference prints and
 * stops, messages.f90:228-241; callers of this ABI decide).
 *
 * Only plain C types cross this boundary.
 */
#ifndef LESGO_AMD_H
#define LESGO_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LG_OK = 0,
    LG_ERR_ARG = 1,
    LG_ERR_HIP = 2,
    LG_ERR_ALLOC = 3,
    LG_ERR_STATE = 4,
    LG_ERR_UNSUPPORTED = 5,
    LG_ERR_PIVOT = 6
};

/* Precision of a case. */
enum { LG_FP64 = 0, LG_MIXED = 1 };

/* Field identifiers: the sim_param / sgs_param arrays, all (ld, ny, 0:nz). */
enum {
    LG_U = 0, LG_V, LG_W,
    LG_DUDX, LG_DUDY, LG_DUDZ, LG_DVDX, LG_DVDY, LG_DVDZ,
    LG_DWDX, LG_DWDY, LG_DWDZ,
    LG_RHSX, LG_RHSY, LG_RHSZ, LG_RHSX_F, LG_RHSY_F, LG_RHSZ_F,
    LG_DPDX, LG_DPDY, LG_DPDZ,
    LG_TXX, LG_TXY, LG_TYY, LG_TXZ, LG_TYZ, LG_TZZ,
    LG_DIVTX, LG_DIVTY, LG_DIVTZ,
    LG_P, LG_FXA, LG_FYA, LG_FZA,
    LG_S11, LG_S12, LG_S13, LG_S22, LG_S23, LG_S33,
    LG_NU_T, LG_CS_OPT2, LG_F_LM, LG_F_MM, LG_F_QN, LG_F_NN, LG_BETA, LG_TN_ALL,
    LG_NFIELDS
};

/*
 * Run parameters (param.f90).  Nz is the number of uv levels in the domain;
 * the one-rank layout has nz = Nz + 1 levels above the ghost level 0
 * (input_util.f90:196-235).  Physics: sgs_model = 5, lbc_mom = 2 (equilibrium
 * wall model), ubc_mom = 0 (stress-free lid), AB2 with a static dt
 * (tadv1 = 1.5) or a CFL-adaptive one (use_cfl_dt), inflow_type = 0.
 */
typedef struct {
    int nx, ny, Nz;
    double lx, ly, lz;
    double dt;
    double zo;
    double vonk;
    int cs_count;
    int use_mean_p_force;
    double mean_p_force_x, mean_p_force_y;
    double nu;
    int precision;          /* LG_FP64 or LG_MIXED */
    /* initu (param.f90:171): 1 = the state comes from a file (initial.f90:
     * 100-102), the dynamic model runs every cs_count steps from the start;
     * 0 = fresh start (inilag): Euler first step (main.f90:273-280),
     * Cs_opt2 = 0.03 until DYN_init, F_* initialised from MM, NN at the first
     * dynamic call (sgs_stag_util.f90:187-193, lagrange_Sdep.f90:266-337).
     * Counted per member from its own first step. */
    int initu;
    int dyn_init;           /* DYN_init (param.f90:123, default 100) */
    /* use_cfl_dt (param.f90:151-152): 1 = each step takes dt = cfl / max(
     * max|u|/dx, max|v|/dy, max|w|/dz) on the device before it starts, with
     * tadv1 = 1 + dt / (2 dt_prev), tadv2 = 1 - tadv1 (cfl_util.f90:73-111,
     * main.f90:135-144); a member whose clock holds no previous step takes
     * an Euler step (initialize.f90:191-200).  dt above is then unused
     * except as lg_set_clock_member's starting value. */
    int use_cfl_dt;
    double cfl;
} lg_params;

typedef struct lg_case lg_case;

/* Lifetime (replaces sim_param_init/sgs_param_init/init_fft/test_filter_init:
 * sim_param.f90:44-99, sgs_param.f90:79-195, fft.f90:102-127,
 * test_filtermodule.f90:38-123).  Fields start at zero. */
int lg_create(const lg_params* prm, lg_case** out);
int lg_destroy(lg_case* c);
const char* lg_last_error(void);
int lg_sizes(const lg_case* c, int* nz, int* ld, int* ny);

/* Host <-> device copies of one field, host layout double [nz+1][ny][ld]. */
int lg_upload(lg_case* c, int field, const double* host);
int lg_download(lg_case* c, int field, double* host);
/* Seeded start written straight into the case (replaces the /dev/urandom
 * ic_les, initial.f90:554, with the reproducible vel.out-style state):
 * u = ln(z/z0)/vonk + U(-noise, noise) ln(z/z0)/vonk on uv levels 1..nz-1,
 * v, w get the same-amplitude noise, w = 0 at the wall and the lid,
 * u(nz) = u(nz-1), v(nz) = v(nz-1), Cs_opt2 = 0.03, everything else 0, level 0
 * BOGUS.  Uniforms: splitmix64 of (seed, field, i, j, k); counters reset. */
int lg_seed_state(lg_case* c, long seed, double noise);
/* Time counters jt, jt_total (param.f90:146-147). */
int lg_set_counters(lg_case* c, int jt, int jt_total);
int lg_get_counters(const lg_case* c, int* jt, int* jt_total);
int lg_synchronize(lg_case* c);
/* The head of the time loop (main.f90:135-151) for callers that step through
 * the subsystem entry points: with use_cfl_dt, dt = get_cfl_dt() and the AB2
 * weights of every stepped member, on the device; total_time += dt.  Call
 * once per step, after lg_set_counters has advanced the counters (lg_step
 * does this itself). */
int lg_advance_clock(lg_case* c);
/* Make the physical fields the case's state, as a restart from files does:
 * lg_step keeps the fields between steps as (kx, ky) spectra and makes the
 * physical ones when they are read; after lg_resync the next step starts
 * from the physical fields.  Writing a checkpoint calls it, so that the run
 * that wrote the files and a run restarted from them continue alike, bit for
 * bit (no counterpart in the reference, whose state is always physical). */
int lg_resync(lg_case* c);

/* --- replaced subroutines (device state in, device state out) ---------- */
/* derivatives::filt_da (derivatives.f90:166-211), levels 0..nz */
int lg_filt_da(lg_case* c, int f, int dfdx, int dfdy);
/* derivatives::ddx / ddy / ddxy (derivatives.f90:37-163), levels 0..nz */
int lg_ddx(lg_case* c, int f, int dfdx);
int lg_ddy(lg_case* c, int f, int dfdy);
int lg_ddxy(lg_case* c, int f, int dfdx, int dfdy);
/* test_filtermodule::test_filter / test_test_filter (test_filtermodule.f90:
 * 126-169) on one host plane f(ld, ny) in place: width 2 applies G_test (the
 * 2-delta cutoff), width 4 G_test_test; the padding columns come back 0. */
int lg_test_filter(lg_case* c, int width, double* plane);
/* derivatives::ddz_uv / ddz_w (derivatives.f90:214-311) */
int lg_ddz_uv(lg_case* c, int f, int dfdz);
int lg_ddz_w(lg_case* c, int f, int dfdz);
/* wallstress (wallstress.f90:21-257): lbc_mom = 2, ubc_mom = 0 */
int lg_wallstress(lg_case* c);
/* sgs_stag (sgs_stag_util.f90:33-464) incl. lagrange_Sdep when it is due */
int lg_sgs_stag(lg_case* c);
/* lagrange_Sdep (lagrange_Sdep.f90:21-432) incl. interpolag_Sdep */
int lg_lagrange_sdep(lg_case* c);
/* interpolag_Sdep (interpolag_Sdep.f90:21-270) alone */
int lg_interpolag_sdep(lg_case* c);
/* divstress_uv / divstress_w (divstress_uv.f90:21-85, divstress_w.f90:21-116) */
int lg_divstress_uv(lg_case* c);
int lg_divstress_w(lg_case* c);
/* convec (convec.f90:21-334) */
int lg_convec(lg_case* c);
/* press_stag_array + tridag_array (press_stag_array.f90:21-290,
 * tridag_array.f90:165-246) */
int lg_press_stag_array(lg_case* c);
/* forcing::project (forcing.f90:149-244) */
int lg_project(lg_case* c);
/* tridag_array(a, b, c, r, u) (tridag_array.f90:165-246) on its own: the
 * batched Thomas solve for every (jx < lh, jy) with real coefficients a, b,
 * c (lh, ny, nz+1) and interleaved complex r, u (ld = 2 lh, ny, nz+1), all
 * FP64 in the reference's column-major layout (C [nz+1][ny][lh] and
 * [nz+1][ny][ld]).  Row 1 is solved for every system; rows 2..nz+1 skip
 * jy = ny/2+1 and (jx, jy) = (1, 1), whose entries of u keep their values,
 * as does column jx = lh.  A zero pivot returns LG_ERR_PIVOT naming jx, jy
 * and the row (the reference prints and stops, 190-193, 217-222).  nzp1 is
 * nz+1.  lg_tridag_array takes host arrays; lg_tridag_array_device takes
 * device arrays and a hipStream_t (NULL: the default stream), allocates the
 * Thomas scratch gam (lh ny (nz+1) doubles) for the call and waits for the
 * stream before it returns. */
int lg_tridag_array(int lh, int ny, int nzp1, const double* a, const double* b, const double* c,
                    const double* r, double* u);
int lg_tridag_array_device(int lh, int ny, int nzp1, const double* a, const double* b,
                           const double* c, const double* r, double* u, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lg_tridag_array_device with the same arguments and
 * results, its scratch (gam and the status words) in a static device array
 * of the library instead of a per-call allocation.  Limit: lh * ny * nzp1
 * <= 2^23 (64 MiB, held on each GPU the library runs on until the process
 * ends); larger sizes return LG_ERR_ARG.  Calls that overlap on one GPU
 * share that array and race. */
int lg_tridag_array_device_noalloc(int lh, int ny, int nzp1, const double* a, const double* b,
                                   const double* c, const double* r, double* u, void* stream);
/* Inline blocks of the time loop in main.f90, exposed so a host driver can
 * mirror the loop call by call:
 *   lg_copy_field        RHSx_f = RHSx etc. (main.f90:155-157)
 *   lg_rhs_ab2           RHS = -RHS - divt + mean_p_force + f_adm; AB2 u*;
 *                        BOGUS fills (main.f90:211-214,228-232,264-308)
 *   lg_rhs_gradp         RHS -= grad p (main.f90:321-326) */
int lg_copy_field(lg_case* c, int src, int dst);
int lg_rhs_ab2(lg_case* c);
int lg_rhs_gradp(lg_case* c);
/* One or more whole time steps (main.f90:146-344). */
int lg_step(lg_case* c, int nsteps);

/* Diagnostics: rmsdiv.f90:21-60, io.f90:99-141 (energy), cfl_util.f90:35-70,
 * functions.f90:882-905 (tau_wall_bot). */
int lg_rmsdiv(lg_case* c, double* out);
int lg_energy(lg_case* c, double* out);
int lg_max_cfl(lg_case* c, double* out);
int lg_tau_wall_bot(lg_case* c, double* out);

/* --- actuator disks (turbines.f90:123-638, turbine_indicator.f90) ------- */
/* Each turbine: xloc, yloc, height, dia, thk, theta1, theta2, Ct_prime
 * (the columns of input_turbines/param.dat, nondimensional). */
int lg_turbines_init(lg_case* c, int nturb, const double* spec8,
                     double alpha1, double alpha2, double filter_cutoff,
                     int adm_correction);
int lg_turbines_forcing(lg_case* c);
/* u_d_T per turbine (turbines.f90:657-679 checkpoint content). */
int lg_turbines_get(lg_case* c, double* u_d_T, double* f_n, int nturb);
/* Time filter of the disk velocity for every disk of the case (turbines.f90:
 * 561-566, 583): u_d_T = (1 - eps) u_d_T + eps u_d with eps = r / (1 + r),
 * r = dt z_i / u_star / T_avg_dim; T_avg_dim <= 0 gives eps = 1 (the
 * default). */
int lg_turbines_time_filter(lg_case* c, double T_avg_dim, double z_i, double u_star);
/* Dynamic thrust coefficients (dyn_Ct_prime, turbines.f90:504-515): Ct' of
 * member m's nturb disks, with the ADM correction that follows from it; the
 * disks' node lists are kept. */
int lg_turbines_set_ct_member(lg_case* c, int m, const double* ct_prime, int nturb);
/* u_d_T of member m's disks (turbine_vel_init reading u_d_T.dat,
 * turbines.f90:229-250). */
int lg_turbines_set_udt_member(lg_case* c, int m, const double* u_d_T, int nturb);

/* --- per-kernel timing (no reference counterpart: the reference only has
 * whole-iteration clocks, clocks.f90:47-83).  When enabled, HIP events are
 * recorded on the case's stream around every launch of the kernel classes
 * below; lg_timing_report returns total milliseconds, launch count and the
 * algorithmic bytes (inputs read + outputs written once) of one class. */
enum {
    LG_K_ROWS_R2C = 0,          /* forward rows, nx points */
    LG_K_ROWS_C2R = 1,          /* inverse rows, nx points */
    LG_K_COLS_FWD = 2,          /* forward columns, ny points */
    LG_K_COLS_INV = 3,          /* inverse columns, ny points */
    LG_K_TRIDAG = 4,            /* pressure tridiagonal */
    LG_K_ROWS_R2C_PAD = 5,      /* forward rows on the 3/2 grid */
    LG_K_ROWS_C2R_PAD = 6,      /* inverse rows on the 3/2 grid */
    LG_K_COLS_UNPAD = 7,        /* forward columns, 3/2 grid to nx x ny */
    LG_K_COLS_PAD = 8,          /* inverse columns, nx x ny to 3/2 grid */
    LG_K_LAG_ROWS = 9,          /* Lagrangian model: inverse rows and contractions */
    LG_K_LAG_PREP = 10,         /* Lagrangian model: forward rows of its quantities */
    LG_K_ROWS_R2C_PROD = 11,    /* forward rows on the 3/2 grid forming the products */
    LG_K_NKINDS = 12
};
int lg_timing_enable(lg_case* c, int on);
int lg_timing_report(lg_case* c, int kind, double* ms, double* launches, double* bytes);

/* Plain device-to-device copy of nbytes (a multiple of 16) repeated reps
 * times: calibrates HBM byte counters and measures copy bandwidth.  No
 * counterpart in the reference. */
int lg_calibrate_copy(long nbytes, int reps, double* ms_per_copy);

/* --- batches of independent cases (no counterpart in the reference, which
 * runs one case per process).  A batch holds `members` cases of one grid,
 * each with its own fields, z0, mean pressure gradient and disks; every
 * entry point above acts on all stepped members at once (lg_step,
 * subsystems) or on member 0 (copies, seeding, diagnostics, disks).  Member
 * m's arithmetic is identical to the same case run alone. */
int lg_create_batch(const lg_params* prm, int members, lg_case** out);
/* Capacity and number of stepped members (members 0..active-1). */
int lg_batch_info(const lg_case* c, int* members, int* active);
int lg_set_active(lg_case* c, int active);
/* Per-member z0 (wallstress.f90:198) and mean_p_force (main.f90:228-232). */
int lg_member_params(lg_case* c, int m, double zo, double mean_p_force_x,
                     double mean_p_force_y);
int lg_upload_member(lg_case* c, int m, int field, const double* host);
int lg_download_member(lg_case* c, int m, int field, double* host);
/* lg_seed_state for member m with its own z0; counters untouched. */
int lg_seed_member(lg_case* c, int m, long seed, double noise);
/* All fields, z0, mean_p_force and disks of member src copied to dst. */
int lg_copy_member(lg_case* c, int src, int dst);
int lg_rmsdiv_member(lg_case* c, int m, double* out);
int lg_energy_member(lg_case* c, int m, double* out);
int lg_max_cfl_member(lg_case* c, int m, double* out);

/* Time-stepping clock of member m (param: dt, tadv1, tadv2; sgs_param:
 * lagran_dt; total_time, main.f90:149): out[0..4] = dt, tadv1, tadv2,
 * lagran_dt, total_time after the last step (with use_cfl_dt, dt is the
 * last step's; lagran_dt the displacement time pending for the next
 * dynamic call, sgs_stag_util.f90:74-82, lagrange_Sdep.f90:430). */
int lg_clock_member(lg_case* c, int m, double* out);
/* Restart a member's clock from total_time.dat (io.f90:62-95): total_time,
 * and with use_cfl_dt the previous step's dt; fresh != 0 makes the next step
 * an Euler step (jt_total = 0 or a changed cfl, initialize.f90:191-200). */
int lg_set_clock_member(lg_case* c, int m, double dt, double total_time, int fresh);
int lg_tau_wall_bot_member(lg_case* c, int m, double* out);
/* Replace member m's disks (nturb = 0 removes them). */
int lg_turbines_init_member(lg_case* c, int m, int nturb, const double* spec8,
                            double alpha1, double alpha2, double filter_cutoff,
                            int adm_correction);
int lg_turbines_get_member(lg_case* c, int m, double* u_d_T, double* f_n, int nturb);
/* The per-turbine history values of the last forcing (the record of
 * turbine_<s>.dat, turbines.f90:589-593, written every tbase steps):
 * out[5 s + 0..4] = u, v, w_uv at the disk centre cell (icp, jcp, kcp of
 * turbines.f90:196-198, sampled at the forcing, 538-547), u_d and u_d_T
 * (the file holds -u_d, -u_d_T).  A centre index past the grid is taken
 * periodically (the reference reads out of bounds there). */
int lg_turbines_history_member(lg_case* c, int m, double* out, int nturb);
/* lg_step replayed from captured HIP graphs: the same launches as lg_step,
 * captured once per (RHS storage parity, dynamic-model step) and replayed;
 * recaptured after lg_set_active or a disk change. */
int lg_step_graph(lg_case* c, int nsteps);

/* --- running time averages (time_average.f90 tavg, driven as io.f90:558-582)
 * and streamwise spectra, accumulated in FP64 on the device after each
 * step.  Between nstart and nend (jt_total) every step adds dt to the
 * pending weight; at nstart and every nskip steps after it the sums are
 * updated (the first such step only starts the summation).  Quantities
 * (tavg_t order, time_average.f90:176-319): */
enum {
    LG_TAVG_U, LG_TAVG_V, LG_TAVG_W, LG_TAVG_W_UV, LG_TAVG_U_W, LG_TAVG_V_W,
    LG_TAVG_U2, LG_TAVG_V2, LG_TAVG_W2, LG_TAVG_UV, LG_TAVG_UW, LG_TAVG_VW,
    LG_TAVG_TXX, LG_TAVG_TYY, LG_TAVG_TZZ, LG_TAVG_TXY, LG_TAVG_TXZ, LG_TAVG_TYZ,
    LG_TAVG_P, LG_TAVG_FX, LG_TAVG_FY, LG_TAVG_FZ, LG_TAVG_CS_OPT2,
    LG_TAVG_VORTX, LG_TAVG_VORTY, LG_TAVG_VORTZ, LG_TAVG_N
};
int lg_stats_enable(lg_case* c, int nstart, int nend, int nskip);
/* total_time: the averaging time so far (member 0's with use_cfl_dt, where
 * every member weighs by its own dt) */
int lg_stats_info(lg_case* c, double* total_time, int* initialized);
/* The averaging time of member m (tavg%total_time, time_average.f90:665-716;
 * with use_cfl_dt each member weighs by its own dt). */
int lg_stats_time(lg_case* c, int m, double* total_time);
/* Time average of one quantity of member m, host double [nz+1][ny][nx]. */
int lg_stats_mean(lg_case* c, int m, int quantity, double* host);
/* The raw time-weighted sum behind lg_stats_mean (the arrays tavg.out holds,
 * time_average.f90:665-716). */
int lg_stats_sums(lg_case* c, int m, int quantity, double* host);
/* Continue a previous run's sums (tavg%init reading tavg.out,
 * time_average.f90:119-166): member m's sums of all LG_TAVG_N quantities
 * (LG_TAVG_N arrays [nz+1][ny][nx] in quantity order) and the averaging
 * time, after lg_stats_enable and before the summation starts (without
 * use_cfl_dt every member shares one averaging time). */
int lg_stats_restore(lg_case* c, int m, double total_time, const double* sums);
/* Streamwise spectrum (0 u, 1 v, 2 w) of member m, host double [nz+1][nx/2]:
 * E(kx, z) = c_kx sum_ky |f_hat(kx, ky, z)|^2 of the normalised x-y
 * transform on levels 1..nz-1, time-averaged; c_0 = 1, c_kx = 2 otherwise,
 * so sum_kx E = plane mean of f^2 less its Nyquist part. */
int lg_stats_spectrum(lg_case* c, int m, int component, double* host);
/* Sums of (a - b)^2 and b^2 of one field of member ma of a and member mb of
 * b (same grid, any precisions) over x < nx and levels k0..k1. */
int lg_diff_l2(lg_case* a, int ma, lg_case* b, int mb, int field, int k0, int k1,
               double* num, double* den);

/* --- autotuner (no counterpart in the reference; FFTW_PATIENT planning,
 * fft.f90:114-121, is the nearest analogue).  Pass kinds: */
enum {
    LG_PASS_R2C = 0,        /* row pass r2c: N = row length, other = rows */
    LG_PASS_C2R = 1,        /* row pass c2r */
    LG_PASS_COLS_FWD = 2,   /* column pass forward: N = column length, other = columns */
    LG_PASS_COLS_INV = 3,   /* column pass inverse */
    LG_PASS_TRIDAG = 4,     /* tridiagonal: N = ny, other = nx/2; per_wg = block size */
    LG_PASS_COLS_UNPAD = 5, /* forward, 3/2 grid -> small */
    LG_PASS_COLS_PAD = 6    /* inverse, small -> 3/2 grid */
};
/* Times every launchable choice (four-step split n1 x n2 or the generic LDS
 * kernel n1 = 0; rows or columns per workgroup) of each pass of the case's
 * step at its real shape and records the fastest in a process-wide table
 * keyed by (kind, precision, N, other).  The table depends on the grid only,
 * so every case of a grid, alone or batched, then runs the same kernels. */
int lg_autotune(lg_case* c, int reps, int* npasses);
/* variant (row passes): 0 = rows read and written in place, 1 = the
 * workgroup's rows staged through LDS with 16-byte vector copies. */
int lg_tune_get(int kind, int precision, int N, int other, int* n1, int* n2, int* per_wg,
                int* variant);
int lg_tune_set(int kind, int precision, int N, int other, int n1, int n2, int per_wg,
                int variant);
int lg_tune_clear(void);
/* Enumerate the table (to save it and lg_tune_set it back in a later
 * process, as FFTW wisdom is kept). */
int lg_tune_count(int* n);
int lg_tune_entry(int i, int* kind, int* precision, int* N, int* other, int* n1, int* n2,
                  int* per_wg, int* variant);

/* Device bytes lg_create_batch allocates for `members` members. */
int lg_case_bytes(const lg_params* prm, int members, double* bytes);
/* Free and total device memory of the current device (hipMemGetInfo). */
int lg_mem_info(double* free_bytes, double* total_bytes);

/* --- serving loop (no counterpart in the reference).  Jobs of different
 * grids are admitted under an HBM budget into pools (batches of
 * pool_members slots, one stream each) of their grid class (all lg_params
 * fields but zo and mean_p_force_x/y), kept device-resident and stepped
 * together: each lg_server_tick admits queued jobs into free slots (a job
 * joins a pool only at a multiple of cs_count steps, so it computes exactly
 * what it computes alone), replays one captured step of every non-empty pool
 * (pools overlap on the device), waits, and retires finished jobs.  Times
 * are seconds on the server's clock (lg_server_now). */
typedef struct lg_server lg_server;
int lg_server_create(double budget_bytes, int pool_members, lg_server** out);
int lg_server_destroy(lg_server* s);
int lg_server_now(lg_server* s, double* t);
/* Queue a job: seeded start (lg_seed_state with prm->zo), nsteps steps,
 * nturb disks (spec8 as lg_turbines_init; alpha 1.5, cutoff 1e-2, ADM
 * correction on).  keep_fields keeps u, v, w, p at completion. */
int lg_server_submit(lg_server* s, const lg_params* prm, long seed, double noise, int nsteps,
                     int nturb, const double* spec8, double t_arrive, int keep_fields,
                     int* job);
int lg_server_tick(lg_server* s, double* tick_seconds, int* running, int* queued);
/* state: 0 queued, 1 running, 2 done. */
int lg_server_job(lg_server* s, int job, int* state, int* steps_done, double* t_arrive,
                  double* t_admit, double* t_done);
int lg_server_result(lg_server* s, int job, int field, double* host);
int lg_server_stats(lg_server* s, double* pts_steps, long* ticks, double* bytes_used,
                    int* npools);

#ifdef __cplusplus
}
#endif

#endif /* LESGO_AMD_H */
