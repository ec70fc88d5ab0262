!*******************************************************************************
module lesgo_gpu
!*******************************************************************************
! Owns the device case and maps the reference's sim_param and sgs_param arrays
! (all 48 the C ABI holds) to field ids.
! A routine that receives one of those arrays (filt_da(u, dudx, dudy, lbz))
! finds its field id by address, so the reference's call sites compile
! unchanged.  Errors stop the program like messages::error
! (messages.f90:228-241).
use iso_c_binding
use types, only : rprec
use lesgo_amd_c
implicit none
save
type(c_ptr) :: gpu_case = c_null_ptr
contains

subroutine gpu_check(rc, what)
integer(c_int), intent(in) :: rc
character(*), intent(in) :: what
character(kind=c_char), pointer :: msg(:)
integer :: n
if (rc /= 0) then
    call c_f_pointer(lg_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
        if (msg(n+1) == c_null_char) exit
        n = n + 1
    end do
    write(*,'(a,a,i0,a,1024a1)') what, ' failed, code ', rc, ': ', msg(1:n)
    stop 1
end if
end subroutine gpu_check

subroutine gpu_init()
use param
type(lg_params) :: prm
prm%nx = Nx; prm%ny = Ny; prm%Nz = nz - 1
prm%lx = L_x; prm%ly = L_y; prm%lz = L_z
prm%dt = dt; prm%zo = zo; prm%vonk = vonk
prm%cs_count = cs_count
prm%use_mean_p_force = merge(1, 0, use_mean_p_force)
prm%mean_p_force_x = mean_p_force_x; prm%mean_p_force_y = mean_p_force_y
prm%nu = 0._rprec
prm%precision = precision
prm%initu = merge(1, 0, initu)
prm%dyn_init = DYN_init
prm%use_cfl_dt = merge(1, 0, use_cfl_dt)
prm%cfl = cfl
call gpu_check(lg_create(prm, gpu_case), 'lg_create')
end subroutine gpu_init

integer(c_int) function field_id(f)
use sim_param
use sgs_param
real(rprec), dimension(:,:,0:), target, intent(in) :: f
type(c_ptr) :: a
a = c_loc(f(1,1,0))
field_id = -1
if (c_associated(a, c_loc(u))) field_id = LG_U
if (c_associated(a, c_loc(v))) field_id = LG_V
if (c_associated(a, c_loc(w))) field_id = LG_W
if (c_associated(a, c_loc(dudx))) field_id = LG_DUDX
if (c_associated(a, c_loc(dudy))) field_id = LG_DUDY
if (c_associated(a, c_loc(dudz))) field_id = LG_DUDZ
if (c_associated(a, c_loc(dvdx))) field_id = LG_DVDX
if (c_associated(a, c_loc(dvdy))) field_id = LG_DVDY
if (c_associated(a, c_loc(dvdz))) field_id = LG_DVDZ
if (c_associated(a, c_loc(dwdx))) field_id = LG_DWDX
if (c_associated(a, c_loc(dwdy))) field_id = LG_DWDY
if (c_associated(a, c_loc(dwdz))) field_id = LG_DWDZ
if (c_associated(a, c_loc(RHSx))) field_id = LG_RHSX
if (c_associated(a, c_loc(RHSy))) field_id = LG_RHSY
if (c_associated(a, c_loc(RHSz))) field_id = LG_RHSZ
if (c_associated(a, c_loc(RHSx_f))) field_id = LG_RHSX_F
if (c_associated(a, c_loc(RHSy_f))) field_id = LG_RHSY_F
if (c_associated(a, c_loc(RHSz_f))) field_id = LG_RHSZ_F
if (c_associated(a, c_loc(dpdx))) field_id = LG_DPDX
if (c_associated(a, c_loc(dpdy))) field_id = LG_DPDY
if (c_associated(a, c_loc(dpdz))) field_id = LG_DPDZ
if (c_associated(a, c_loc(txx))) field_id = LG_TXX
if (c_associated(a, c_loc(txy))) field_id = LG_TXY
if (c_associated(a, c_loc(tyy))) field_id = LG_TYY
if (c_associated(a, c_loc(txz))) field_id = LG_TXZ
if (c_associated(a, c_loc(tyz))) field_id = LG_TYZ
if (c_associated(a, c_loc(tzz))) field_id = LG_TZZ
if (c_associated(a, c_loc(divtx))) field_id = LG_DIVTX
if (c_associated(a, c_loc(divty))) field_id = LG_DIVTY
if (c_associated(a, c_loc(divtz))) field_id = LG_DIVTZ
if (c_associated(a, c_loc(p))) field_id = LG_P
if (c_associated(a, c_loc(fxa))) field_id = LG_FXA
if (c_associated(a, c_loc(fya))) field_id = LG_FYA
if (c_associated(a, c_loc(fza))) field_id = LG_FZA
if (c_associated(a, c_loc(S11))) field_id = LG_S11
if (c_associated(a, c_loc(S12))) field_id = LG_S12
if (c_associated(a, c_loc(S13))) field_id = LG_S13
if (c_associated(a, c_loc(S22))) field_id = LG_S22
if (c_associated(a, c_loc(S23))) field_id = LG_S23
if (c_associated(a, c_loc(S33))) field_id = LG_S33
if (c_associated(a, c_loc(Nu_t))) field_id = LG_NU_T
if (c_associated(a, c_loc(Cs_opt2))) field_id = LG_CS_OPT2
if (c_associated(a, c_loc(F_LM))) field_id = LG_F_LM
if (c_associated(a, c_loc(F_MM))) field_id = LG_F_MM
if (c_associated(a, c_loc(F_QN))) field_id = LG_F_QN
if (c_associated(a, c_loc(F_NN))) field_id = LG_F_NN
if (c_associated(a, c_loc(Beta))) field_id = LG_BETA
if (c_associated(a, c_loc(Tn_all))) field_id = LG_TN_ALL
if (field_id < 0) then
    write(*,*) 'lesgo_gpu: array is not a sim_param or sgs_param field'
    stop 1
end if
end function field_id

subroutine gpu_download(f)
real(rprec), dimension(:,:,0:), target, intent(inout) :: f
call gpu_check(lg_download(gpu_case, field_id(f), c_loc(f(1,1,0))), 'lg_download')
end subroutine gpu_download

subroutine gpu_upload(f)
real(rprec), dimension(:,:,0:), target, intent(in) :: f
call gpu_check(lg_upload(gpu_case, field_id(f), c_loc(f(1,1,0))), 'lg_upload')
end subroutine gpu_upload

subroutine gpu_counters()
use param, only : jt, jt_total
call gpu_check(lg_set_counters(gpu_case, int(jt, c_int), int(jt_total, c_int)), 'lg_set_counters')
end subroutine gpu_counters

end module lesgo_gpu

!*******************************************************************************
module derivatives
!*******************************************************************************
! Drop-in for derivatives.f90 (ddx 37-76, ddy 79-118, ddxy 121-163, filt_da
! 166-211, ddz_uv 214-264, ddz_w 267-311) on the device state.
use types, only : rprec
use lesgo_gpu
implicit none
contains
subroutine ddx(f, dfdx, lbz)
integer, intent(in) :: lbz
real(rprec), dimension(:,:,lbz:), target, intent(in) :: f
real(rprec), dimension(:,:,lbz:), target, intent(inout) :: dfdx
call gpu_check(lg_ddx(gpu_case, field_id(f), field_id(dfdx)), 'ddx')
end subroutine ddx
subroutine ddy(f, dfdy, lbz)
integer, intent(in) :: lbz
real(rprec), dimension(:,:,lbz:), target, intent(in) :: f
real(rprec), dimension(:,:,lbz:), target, intent(inout) :: dfdy
call gpu_check(lg_ddy(gpu_case, field_id(f), field_id(dfdy)), 'ddy')
end subroutine ddy
subroutine ddxy(f, dfdx, dfdy, lbz)
integer, intent(in) :: lbz
real(rprec), dimension(:,:,lbz:), target, intent(in) :: f
real(rprec), dimension(:,:,lbz:), target, intent(inout) :: dfdx, dfdy
call gpu_check(lg_ddxy(gpu_case, field_id(f), field_id(dfdx), field_id(dfdy)), 'ddxy')
end subroutine ddxy
subroutine filt_da(f, dfdx, dfdy, lbz)
integer, intent(in) :: lbz
real(rprec), dimension(:,:,lbz:), target, intent(inout) :: f, dfdx, dfdy
call gpu_check(lg_filt_da(gpu_case, field_id(f), field_id(dfdx), field_id(dfdy)), 'filt_da')
end subroutine filt_da
subroutine ddz_uv(f, dfdz, lbz)
integer, intent(in) :: lbz
real(rprec), dimension(:,:,lbz:), target, intent(inout) :: f, dfdz
call gpu_check(lg_ddz_uv(gpu_case, field_id(f), field_id(dfdz)), 'ddz_uv')
end subroutine ddz_uv
subroutine ddz_w(f, dfdz, lbz)
integer, intent(in) :: lbz
real(rprec), dimension(:,:,lbz:), target, intent(inout) :: f, dfdz
call gpu_check(lg_ddz_w(gpu_case, field_id(f), field_id(dfdz)), 'ddz_w')
end subroutine ddz_w
end module derivatives

!*******************************************************************************
module sgs_stag_util
!*******************************************************************************
! Drop-in for sgs_stag_util::sgs_stag (sgs_stag_util.f90:33-464), which calls
! lagrange_Sdep every cs_count steps.
use lesgo_gpu
implicit none
contains
subroutine sgs_stag()
call gpu_counters()
call gpu_check(lg_sgs_stag(gpu_case), 'sgs_stag')
end subroutine sgs_stag
end module sgs_stag_util

!*******************************************************************************
module forcing
!*******************************************************************************
! Drop-in for forcing::project and forcing::forcing_applied
! (forcing.f90:79-117, 149-244).
use lesgo_gpu
implicit none
logical :: turbines_on = .false.
contains
subroutine forcing_applied()
if (turbines_on) call gpu_check(lg_turbines_forcing(gpu_case), 'turbines_forcing')
end subroutine forcing_applied
subroutine project()
call gpu_check(lg_project(gpu_case), 'project')
end subroutine project
end module forcing

!*******************************************************************************
module test_filtermodule
!*******************************************************************************
! Drop-in for test_filtermodule::test_filter and test_test_filter
! (test_filtermodule.f90:126-169): one (ld, ny) plane filtered in place with
! the case's G_test (2 delta) or G_test_test (4 delta).  The filters are set
! up with the case (test_filter_init, 38-123, is part of lg_create).
use types, only : rprec
use lesgo_gpu
implicit none
contains
subroutine test_filter_init()
end subroutine test_filter_init
subroutine test_filter(f)
real(rprec), dimension(:,:), contiguous, target, intent(inout) :: f
call gpu_check(lg_test_filter(gpu_case, 2_c_int, c_loc(f(1,1))), 'test_filter')
end subroutine test_filter
subroutine test_test_filter(f)
real(rprec), dimension(:,:), contiguous, target, intent(inout) :: f
call gpu_check(lg_test_filter(gpu_case, 4_c_int, c_loc(f(1,1))), 'test_test_filter')
end subroutine test_test_filter
end module test_filtermodule

!*******************************************************************************
module turbines
!*******************************************************************************
! Drop-in for turbines::turbines_forcing (turbines.f90:465-638): the disks
! set up with lg_turbines_init force the device state; fxa, fya, fza are
! the device's (gpu_download fills the host mirrors).
use lesgo_gpu
implicit none
private
public :: turbines_forcing
contains
subroutine turbines_forcing()
call gpu_check(lg_turbines_forcing(gpu_case), 'turbines_forcing')
end subroutine turbines_forcing
end module turbines

! External subroutines with the reference's names and argument lists.
subroutine lagrange_Sdep()
! lagrange_Sdep.f90:21-432, including its interpolag_Sdep call
use lesgo_gpu
call gpu_counters()
call gpu_check(lg_lagrange_sdep(gpu_case), 'lagrange_Sdep')
end subroutine lagrange_Sdep

subroutine interpolag_Sdep()
! interpolag_Sdep.f90:21-270
use lesgo_gpu
call gpu_check(lg_interpolag_sdep(gpu_case), 'interpolag_Sdep')
end subroutine interpolag_Sdep

subroutine tridag_array(a, b, c, r, u)
! tridag_array.f90:166-246 on host arrays (the serial variant's shapes)
use types, only : rprec
use param, only : lh, ld, ny, nz
use lesgo_gpu
real(rprec), dimension(lh,ny,nz+1), target, intent(in) :: a, b, c
real(rprec), dimension(ld,ny,nz+1), target, intent(in) :: r
real(rprec), dimension(ld,ny,nz+1), target, intent(inout) :: u
call gpu_check(lg_tridag_array(int(lh, c_int), int(ny, c_int), int(nz + 1, c_int),          &
    c_loc(a), c_loc(b), c_loc(c), c_loc(r), c_loc(u)), 'tridag_array')
end subroutine tridag_array

subroutine wallstress()
use lesgo_gpu
call gpu_check(lg_wallstress(gpu_case), 'wallstress')
end subroutine wallstress

subroutine divstress_uv(divtx, divty, txx, txy, txz, tyy, tyz)
use types, only : rprec
use param, only : ld, ny, nz, lbz
use lesgo_gpu
real(rprec), dimension(ld,ny,lbz:nz), target :: divtx, divty, txx, txy, txz, tyy, tyz
if (field_id(divtx) /= LG_DIVTX .or. field_id(txx) /= LG_TXX) stop 'divstress_uv: arrays'
call gpu_check(lg_divstress_uv(gpu_case), 'divstress_uv')
end subroutine divstress_uv

subroutine divstress_w(divt, tx, ty, tz)
use types, only : rprec
use param, only : ld, ny, nz, lbz
use lesgo_gpu
real(rprec), dimension(ld,ny,lbz:nz), target :: divt, tx, ty, tz
if (field_id(divt) /= LG_DIVTZ .or. field_id(tz) /= LG_TZZ) stop 'divstress_w: arrays'
call gpu_check(lg_divstress_w(gpu_case), 'divstress_w')
end subroutine divstress_w

subroutine convec()
use lesgo_gpu
call gpu_check(lg_convec(gpu_case), 'convec')
end subroutine convec

subroutine press_stag_array()
use lesgo_gpu
call gpu_check(lg_press_stag_array(gpu_case), 'press_stag_array')
end subroutine press_stag_array

subroutine rmsdiv(rms)
use types, only : rprec
use lesgo_gpu
real(rprec) :: rms
call gpu_check(lg_rmsdiv(gpu_case, rms), 'rmsdiv')
end subroutine rmsdiv
