!*******************************************************************************
program lesgo_amd_dropins
!*******************************************************************************
! Calls the drop-ins of the reference procedures that the time loop itself
! does not call through their own names: derivatives::ddx, ddy, ddxy,
! test_filtermodule::test_filter, test_test_filter, the lagrange_Sdep and
! interpolag_Sdep externals, tridag_array and turbines::turbines_forcing.
! A 32 x 24 x 12 case with a 2 x 2 disk array is stepped three times, its
! state is written to state.bin (all 48 fields, (ld, ny, 0:nz) each), then
! each drop-in runs on that state and its outputs are written to out.bin in
! the order listed in the body (records of doubles, stream access), for a
! caller to compare with a CPU restatement of the same procedures.
use iso_c_binding
use types, only : rprec
use param
use sim_param
use sgs_param
use lesgo_amd_c
use lesgo_gpu
use derivatives, only : ddx, ddy, ddxy
use test_filtermodule, only : test_filter, test_test_filter
use turbines, only : turbines_forcing
implicit none
interface
    subroutine lagrange_Sdep()
    end subroutine lagrange_Sdep
    subroutine interpolag_Sdep()
    end subroutine interpolag_Sdep
    subroutine tridag_array(a, b, c, r, u)
    use types, only : rprec
    use param, only : lh, ld, ny, nz
    real(rprec), dimension(lh,ny,nz+1), intent(in) :: a, b, c
    real(rprec), dimension(ld,ny,nz+1), intent(in) :: r
    real(rprec), dimension(ld,ny,nz+1), intent(inout) :: u
    end subroutine tridag_array
end interface
integer :: un, i, j, k
real(rprec), allocatable :: plane(:,:), ta(:,:,:), tb(:,:,:), tc(:,:,:), tr(:,:,:), tu(:,:,:)
real(rprec), allocatable :: F0(:,:,:,:)
real(rprec) :: farm(16)

Nx = 32; Ny = 24; Nz = 12
nz = Nz + 1
nz_tot = nz
nx2 = 3*Nx/2; ny2 = 3*Ny/2
lh = Nx/2 + 1; ld = 2*lh
lh_big = nx2/2 + 1; ld_big = 2*lh_big
dx = L_x/Nx; dy = L_y/Ny; dz = L_z/(nz_tot - 1)
initu = .true.
call sim_param_init()
call sgs_param_init()
call gpu_init()
call gpu_check(lg_seed_state(gpu_case, 6_c_long, 0.1_rprec), 'lg_seed_state')
farm(1:8) = [2.0_rprec, 1.0_rprec, 0.3_rprec, 0.3_rprec, 0._rprec, 0._rprec, 0._rprec, 1.33_rprec]
farm(9:16) = [4.5_rprec, 2.2_rprec, 0.3_rprec, 0.3_rprec, 0._rprec, 0._rprec, 0._rprec, 1.33_rprec]
call gpu_check(lg_turbines_init(gpu_case, 2_c_int, farm, 1.5_rprec, 1.5_rprec, 1.e-2_rprec,    &
    1_c_int), 'lg_turbines_init')
call gpu_check(lg_step(gpu_case, 3_c_int), 'lg_step')
jt = 3; jt_total = 3

open(newunit=un, file='state.bin', access='stream', form='unformatted', status='replace')
call put_all(un)
close(un)
allocate(F0(ld,ny,0:nz,4))
F0(:,:,:,1) = F_LM; F0(:,:,:,2) = F_MM; F0(:,:,:,3) = F_QN; F0(:,:,:,4) = F_NN

open(newunit=un, file='out.bin', access='stream', form='unformatted', status='replace')
! 1-4: ddx(u), ddy(v), ddxy(w) -> dudx, dvdy, dwdx, dwdy
call ddx(u, dudx, lbz)
call ddy(v, dvdy, lbz)
call ddxy(w, dwdx, dwdy, lbz)
call gpu_download(dudx); call gpu_download(dvdy)
call gpu_download(dwdx); call gpu_download(dwdy)
write(un) dudx, dvdy, dwdx, dwdy
! 5-6: test_filter of u at level 2, test_test_filter of v at level 3
allocate(plane(ld, ny))
plane = u(:, :, 2)
call test_filter(plane)
write(un) plane
plane = v(:, :, 3)
call test_test_filter(plane)
write(un) plane
! 7: tridag_array on the systems below (also written, 8-11: a, b, c, r)
allocate(ta(lh,ny,nz+1), tb(lh,ny,nz+1), tc(lh,ny,nz+1), tr(ld,ny,nz+1), tu(ld,ny,nz+1))
do k = 1, nz + 1
    do j = 1, ny
        do i = 1, lh
            ta(i,j,k) = 1._rprec + 0.1_rprec*sin(real(i + 2*j + 3*k, rprec))
            tc(i,j,k) = 1._rprec + 0.1_rprec*cos(real(2*i + j + k, rprec))
            tb(i,j,k) = -3._rprec - real(i*i + j, rprec)/50._rprec
        end do
        do i = 1, ld
            tr(i,j,k) = sin(0.37_rprec*i + 0.11_rprec*j*k)
        end do
    end do
end do
tu = 0._rprec
call tridag_array(ta, tb, tc, tr, tu)
write(un) tu, ta, tb, tc, tr
! 12-15: interpolag_Sdep on the written state -> F_LM, F_MM, F_QN, F_NN
call interpolag_Sdep()
call gpu_download(F_LM); call gpu_download(F_MM)
call gpu_download(F_QN); call gpu_download(F_NN)
write(un) F_LM, F_MM, F_QN, F_NN
! 16-21: lagrange_Sdep on the written state (the F_* restored first) at
! jt = jt_total = cs_count -> F_LM, F_MM, F_QN, F_NN, Cs_opt2, Beta
F_LM = F0(:,:,:,1); F_MM = F0(:,:,:,2); F_QN = F0(:,:,:,3); F_NN = F0(:,:,:,4)
call gpu_upload(F_LM); call gpu_upload(F_MM); call gpu_upload(F_QN); call gpu_upload(F_NN)
jt = cs_count; jt_total = cs_count
call lagrange_Sdep()
call gpu_download(F_LM); call gpu_download(F_MM)
call gpu_download(F_QN); call gpu_download(F_NN)
call gpu_download(Cs_opt2); call gpu_download(Beta)
write(un) F_LM, F_MM, F_QN, F_NN, Cs_opt2, Beta
! 22-24: turbines_forcing on the written velocity -> fxa, fya, fza
call turbines_forcing()
call gpu_download(fxa); call gpu_download(fya); call gpu_download(fza)
write(un) fxa, fya, fza
close(un)
call gpu_check(lg_destroy(gpu_case), 'lg_destroy')
write(*,'(a)') 'dropins complete'

contains

subroutine put_all(unit)
! every field of the device state, in the C ABI's field order
integer, intent(in) :: unit
call gpu_download(u); call gpu_download(v); call gpu_download(w)
call gpu_download(dudx); call gpu_download(dudy); call gpu_download(dudz)
call gpu_download(dvdx); call gpu_download(dvdy); call gpu_download(dvdz)
call gpu_download(dwdx); call gpu_download(dwdy); call gpu_download(dwdz)
call gpu_download(RHSx); call gpu_download(RHSy); call gpu_download(RHSz)
call gpu_download(RHSx_f); call gpu_download(RHSy_f); call gpu_download(RHSz_f)
call gpu_download(dpdx); call gpu_download(dpdy); call gpu_download(dpdz)
call gpu_download(txx); call gpu_download(txy); call gpu_download(tyy)
call gpu_download(txz); call gpu_download(tyz); call gpu_download(tzz)
call gpu_download(divtx); call gpu_download(divty); call gpu_download(divtz)
call gpu_download(p); call gpu_download(fxa); call gpu_download(fya); call gpu_download(fza)
call gpu_download(S11); call gpu_download(S12); call gpu_download(S13)
call gpu_download(S22); call gpu_download(S23); call gpu_download(S33)
call gpu_download(Nu_t); call gpu_download(Cs_opt2)
call gpu_download(F_LM); call gpu_download(F_MM); call gpu_download(F_QN); call gpu_download(F_NN)
call gpu_download(Beta); call gpu_download(Tn_all)
write(unit) u, v, w, dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz,      &
    RHSx, RHSy, RHSz, RHSx_f, RHSy_f, RHSz_f, dpdx, dpdy, dpdz,                &
    txx, txy, tyy, txz, tyz, tzz, divtx, divty, divtz, p, fxa, fya, fza,       &
    S11, S12, S13, S22, S23, S33, Nu_t, Cs_opt2, F_LM, F_MM, F_QN, F_NN, Beta, Tn_all
end subroutine put_all

end program lesgo_amd_dropins
