!*******************************************************************************
module types
!*******************************************************************************
! Precision kinds, as the reference's types module (types.f90:21-37).
implicit none
integer, parameter :: rprec = kind(1.d0)
end module types

!*******************************************************************************
module param
!*******************************************************************************
! The subset of the reference's param module (param.f90) the time step reads,
! with the same names; set from the main program's input, derived sizes follow
! input_util.f90:196-235 for one process (lbz = 0, nz = Nz + 1).
use types, only : rprec
implicit none
save
integer, parameter :: lbz = 0
real(rprec), parameter :: BOGUS = -1234567890._rprec
real(rprec), parameter :: pi = 3.1415926535897932384626433_rprec
integer :: Nx = 32, Ny = 32, Nz = 32       ! Nz: uv levels in the input; nz below
integer :: nz_tot, nx2, ny2, lh, ld, lh_big, ld_big
real(rprec) :: L_x = 2._rprec*pi, L_y = pi, L_z = 1._rprec
real(rprec) :: dx, dy, dz
real(rprec) :: dt = 2.0e-4_rprec, zo = 1.0e-4_rprec, vonk = 0.4_rprec
integer :: cs_count = 5
integer :: DYN_init = 100
! the seeded start stands in for a vel.out file (initial.f90:100-102)
logical :: initu = .true.
logical :: use_mean_p_force = .true.
real(rprec) :: mean_p_force_x = 1._rprec, mean_p_force_y = 0._rprec
real(rprec) :: tadv1 = 1.5_rprec, tadv2 = -0.5_rprec
! CFL-adaptive dt (param.f90:151-152): dt, tadv1, tadv2 then live on the
! device, per step (lg_advance_clock); clock_sync copies them here
logical :: use_cfl_dt = .false.
real(rprec) :: cfl = 0.05_rprec
integer :: nsteps = 100, wbase = 10, domain_nskip = 10
integer :: jt = 0, jt_total = 0
real(rprec) :: total_time = 0._rprec
integer :: precision = 0                     ! 0 FP64, 1 mixed
integer :: seed = 0
logical :: fused = .false.                   ! .true.: whole steps via lg_step
character(len=256) :: path = './'
! checkpoint (io.f90:1173-1267): vel.out, grid.out, total_time.dat
logical :: checkpoint_data = .false.
integer :: checkpoint_nskip = 10000
character(len=64) :: checkpoint_file = 'vel.out'
end module param

!*******************************************************************************
module sim_param
!*******************************************************************************
! Host mirrors of the reference's sim_param arrays (sim_param.f90:31-93),
! (ld, ny, lbz:nz).  The device copies are the working state; these are
! filled on demand by gpu_download.
use types, only : rprec
implicit none
save
real(rprec), dimension(:,:,:), allocatable, target :: u, v, w,                 &
    dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz,                      &
    RHSx, RHSy, RHSz, RHSx_f, RHSy_f, RHSz_f, dpdx, dpdy, dpdz,                &
    txx, txy, tyy, txz, tyz, tzz, divtx, divty, divtz, p, fxa, fya, fza
contains
subroutine sim_param_init()
use param, only : ld, ny, nz, lbz
allocate(u(ld,ny,lbz:nz), v(ld,ny,lbz:nz), w(ld,ny,lbz:nz))
allocate(dudx(ld,ny,lbz:nz), dudy(ld,ny,lbz:nz), dudz(ld,ny,lbz:nz))
allocate(dvdx(ld,ny,lbz:nz), dvdy(ld,ny,lbz:nz), dvdz(ld,ny,lbz:nz))
allocate(dwdx(ld,ny,lbz:nz), dwdy(ld,ny,lbz:nz), dwdz(ld,ny,lbz:nz))
allocate(RHSx(ld,ny,lbz:nz), RHSy(ld,ny,lbz:nz), RHSz(ld,ny,lbz:nz))
allocate(RHSx_f(ld,ny,lbz:nz), RHSy_f(ld,ny,lbz:nz), RHSz_f(ld,ny,lbz:nz))
allocate(dpdx(ld,ny,lbz:nz), dpdy(ld,ny,lbz:nz), dpdz(ld,ny,lbz:nz))
allocate(txx(ld,ny,lbz:nz), txy(ld,ny,lbz:nz), tyy(ld,ny,lbz:nz))
allocate(txz(ld,ny,lbz:nz), tyz(ld,ny,lbz:nz), tzz(ld,ny,lbz:nz))
allocate(divtx(ld,ny,lbz:nz), divty(ld,ny,lbz:nz), divtz(ld,ny,lbz:nz))
allocate(p(ld,ny,0:nz))
allocate(fxa(ld,ny,lbz:nz), fya(ld,ny,lbz:nz), fza(ld,ny,lbz:nz))
u = 0; v = 0; w = 0; p = 0; fxa = 0; fya = 0; fza = 0
end subroutine sim_param_init
end module sim_param

!*******************************************************************************
module sgs_param
!*******************************************************************************
! Host mirrors of the sgs_param arrays (sgs_param.f90:29-184) that the
! device case holds, all (ld, ny, lbz:nz) like the C ABI's fields (the
! reference allocates S_ij, Nu_t and Cs_opt2 from level 1; level 0 is
! unused there).
use types, only : rprec
implicit none
save
real(rprec), dimension(:,:,:), allocatable, target :: Cs_opt2, F_LM, F_MM, F_QN, F_NN, &
    S11, S12, S13, S22, S23, S33, Nu_t, Beta, Tn_all
contains
subroutine sgs_param_init()
use param, only : ld, ny, nz, lbz
allocate(Cs_opt2(ld,ny,lbz:nz), F_LM(ld,ny,lbz:nz), F_MM(ld,ny,lbz:nz))
allocate(F_QN(ld,ny,lbz:nz), F_NN(ld,ny,lbz:nz))
allocate(S11(ld,ny,lbz:nz), S12(ld,ny,lbz:nz), S13(ld,ny,lbz:nz))
allocate(S22(ld,ny,lbz:nz), S23(ld,ny,lbz:nz), S33(ld,ny,lbz:nz))
allocate(Nu_t(ld,ny,lbz:nz), Beta(ld,ny,lbz:nz), Tn_all(ld,ny,lbz:nz))
Cs_opt2 = 0; F_LM = 0; F_MM = 0; F_QN = 0; F_NN = 0
S11 = 0; S12 = 0; S13 = 0; S22 = 0; S23 = 0; S33 = 0; Nu_t = 0; Beta = 0; Tn_all = 0
end subroutine sgs_param_init
end module sgs_param
