!*******************************************************************************
module lesgo_amd_c
!*******************************************************************************
! iso_c_binding view of the lesgo_amd C ABI (include/lesgo_amd.h).
use iso_c_binding
implicit none

integer(c_int), parameter :: LG_FP64 = 0, LG_MIXED = 1

! field identifiers (same order as the C enum)
integer(c_int), parameter :: LG_U = 0, LG_V = 1, LG_W = 2, LG_DUDX = 3,        &
    LG_DUDY = 4, LG_DUDZ = 5, LG_DVDX = 6, LG_DVDY = 7, LG_DVDZ = 8,           &
    LG_DWDX = 9, LG_DWDY = 10, LG_DWDZ = 11, LG_RHSX = 12, LG_RHSY = 13,       &
    LG_RHSZ = 14, LG_RHSX_F = 15, LG_RHSY_F = 16, LG_RHSZ_F = 17,              &
    LG_DPDX = 18, LG_DPDY = 19, LG_DPDZ = 20, LG_TXX = 21, LG_TXY = 22,        &
    LG_TYY = 23, LG_TXZ = 24, LG_TYZ = 25, LG_TZZ = 26, LG_DIVTX = 27,         &
    LG_DIVTY = 28, LG_DIVTZ = 29, LG_P = 30, LG_FXA = 31, LG_FYA = 32,         &
    LG_FZA = 33, LG_S11 = 34, LG_S12 = 35, LG_S13 = 36, LG_S22 = 37,           &
    LG_S23 = 38, LG_S33 = 39, LG_NU_T = 40, LG_CS_OPT2 = 41, LG_F_LM = 42,     &
    LG_F_MM = 43, LG_F_QN = 44, LG_F_NN = 45, LG_BETA = 46, LG_TN_ALL = 47,    &
    LG_NFIELDS = 48

type, bind(C) :: lg_params
    integer(c_int) :: nx, ny, Nz
    real(c_double) :: lx, ly, lz
    real(c_double) :: dt
    real(c_double) :: zo
    real(c_double) :: vonk
    integer(c_int) :: cs_count
    integer(c_int) :: use_mean_p_force
    real(c_double) :: mean_p_force_x, mean_p_force_y
    real(c_double) :: nu
    integer(c_int) :: precision
    integer(c_int) :: initu
    integer(c_int) :: dyn_init
    integer(c_int) :: use_cfl_dt
    real(c_double) :: cfl
end type lg_params

interface
    integer(c_int) function lg_create(prm, out) bind(C, name='lg_create')
        import :: c_int, c_ptr, lg_params
        type(lg_params), intent(in) :: prm
        type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function lg_destroy(c) bind(C, name='lg_destroy')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    type(c_ptr) function lg_last_error() bind(C, name='lg_last_error')
        import :: c_ptr
    end function
    integer(c_int) function lg_upload(c, field, host) bind(C, name='lg_upload')
        import :: c_int, c_ptr
        type(c_ptr), value :: c, host
        integer(c_int), value :: field
    end function
    integer(c_int) function lg_download(c, field, host) bind(C, name='lg_download')
        import :: c_int, c_ptr
        type(c_ptr), value :: c, host
        integer(c_int), value :: field
    end function
    integer(c_int) function lg_seed_state(c, seed, noise) bind(C, name='lg_seed_state')
        import :: c_int, c_ptr, c_long, c_double
        type(c_ptr), value :: c
        integer(c_long), value :: seed
        real(c_double), value :: noise
    end function
    integer(c_int) function lg_set_counters(c, jt, jt_total) bind(C, name='lg_set_counters')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
        integer(c_int), value :: jt, jt_total
    end function
    integer(c_int) function lg_synchronize(c) bind(C, name='lg_synchronize')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_filt_da(c, f, dfdx, dfdy) bind(C, name='lg_filt_da')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
        integer(c_int), value :: f, dfdx, dfdy
    end function
    integer(c_int) function lg_ddx(c, f, dfdx) bind(C, name='lg_ddx')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
        integer(c_int), value :: f, dfdx
    end function
    integer(c_int) function lg_ddy(c, f, dfdy) bind(C, name='lg_ddy')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
        integer(c_int), value :: f, dfdy
    end function
    integer(c_int) function lg_ddxy(c, f, dfdx, dfdy) bind(C, name='lg_ddxy')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
        integer(c_int), value :: f, dfdx, dfdy
    end function
    integer(c_int) function lg_test_filter(c, width, plane) bind(C, name='lg_test_filter')
        import :: c_int, c_ptr
        type(c_ptr), value :: c, plane
        integer(c_int), value :: width
    end function
    integer(c_int) function lg_lagrange_sdep(c) bind(C, name='lg_lagrange_sdep')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_interpolag_sdep(c) bind(C, name='lg_interpolag_sdep')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_tridag_array(lh, ny, nzp1, a, b, c, r, u)                 &
        bind(C, name='lg_tridag_array')
        import :: c_int, c_ptr
        integer(c_int), value :: lh, ny, nzp1
        type(c_ptr), value :: a, b, c, r, u
    end function
    integer(c_int) function lg_ddz_uv(c, f, dfdz) bind(C, name='lg_ddz_uv')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
        integer(c_int), value :: f, dfdz
    end function
    integer(c_int) function lg_ddz_w(c, f, dfdz) bind(C, name='lg_ddz_w')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
        integer(c_int), value :: f, dfdz
    end function
    integer(c_int) function lg_wallstress(c) bind(C, name='lg_wallstress')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_sgs_stag(c) bind(C, name='lg_sgs_stag')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_divstress_uv(c) bind(C, name='lg_divstress_uv')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_divstress_w(c) bind(C, name='lg_divstress_w')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_convec(c) bind(C, name='lg_convec')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_press_stag_array(c) bind(C, name='lg_press_stag_array')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_project(c) bind(C, name='lg_project')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_copy_field(c, src, dst) bind(C, name='lg_copy_field')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
        integer(c_int), value :: src, dst
    end function
    integer(c_int) function lg_rhs_ab2(c) bind(C, name='lg_rhs_ab2')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_rhs_gradp(c) bind(C, name='lg_rhs_gradp')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_step(c, nsteps) bind(C, name='lg_step')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
        integer(c_int), value :: nsteps
    end function
    integer(c_int) function lg_rmsdiv(c, out) bind(C, name='lg_rmsdiv')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: c
        real(c_double), intent(out) :: out
    end function
    integer(c_int) function lg_energy(c, out) bind(C, name='lg_energy')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: c
        real(c_double), intent(out) :: out
    end function
    integer(c_int) function lg_advance_clock(c) bind(C, name='lg_advance_clock')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_resync(c) bind(C, name='lg_resync')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
    integer(c_int) function lg_clock_member(c, m, out) bind(C, name='lg_clock_member')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: c
        integer(c_int), value :: m
        real(c_double), intent(out) :: out(5)
    end function
    integer(c_int) function lg_set_clock_member(c, m, dt, total_time, fresh)             &
        bind(C, name='lg_set_clock_member')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: c
        integer(c_int), value :: m, fresh
        real(c_double), value :: dt, total_time
    end function
    integer(c_int) function lg_max_cfl(c, out) bind(C, name='lg_max_cfl')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: c
        real(c_double), intent(out) :: out
    end function
    integer(c_int) function lg_tau_wall_bot(c, out) bind(C, name='lg_tau_wall_bot')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: c
        real(c_double), intent(out) :: out
    end function
    integer(c_int) function lg_turbines_init(c, nturb, spec8, alpha1, alpha2,  &
        filter_cutoff, adm_correction) bind(C, name='lg_turbines_init')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: c
        integer(c_int), value :: nturb, adm_correction
        real(c_double), intent(in) :: spec8(*)
        real(c_double), value :: alpha1, alpha2, filter_cutoff
    end function
    integer(c_int) function lg_turbines_forcing(c) bind(C, name='lg_turbines_forcing')
        import :: c_int, c_ptr
        type(c_ptr), value :: c
    end function
end interface

end module lesgo_amd_c
