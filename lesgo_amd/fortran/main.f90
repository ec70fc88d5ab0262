!*******************************************************************************
program lesgo_amd_main
!*******************************************************************************
! Fortran driver of the time loop (main.f90:130-468) with every replaced
! subsystem running on the GPU through the lesgo_amd C ABI.  Input: a namelist
! file (first argument, default lesgo_amd.nml) with group &lesgo; output:
! output/vel.<jt>.c0.bin (u, v, w_uv) and output/pres.<jt>.c0.bin (p - |u|^2/2)
! every domain_nskip steps as direct-access records (io.f90:786-884), and a
! stdout block with rmsdiv and KE every wbase steps (main.f90:402-436).
use iso_c_binding
use types, only : rprec
use param
use sim_param
use sgs_param
use lesgo_amd_c
use lesgo_gpu
use derivatives, only : filt_da, ddz_uv, ddz_w
use sgs_stag_util, only : sgs_stag
use forcing, only : forcing_applied, project, turbines_on
implicit none
namelist /lesgo/ Nx, Ny, Nz, L_x, L_y, L_z, dt, zo, cs_count, mean_p_force_x, &
    nsteps, wbase, domain_nskip, precision, seed, fused, path, turbines_on, initu, DYN_init, &
    checkpoint_data, checkpoint_nskip, checkpoint_file, use_cfl_dt, cfl
character(len=512) :: nml
integer :: jt_step, fid, ios, count0, count1, rate
logical :: restart
real(rprec) :: rms, ke, tw
real(rprec), allocatable :: farm(:)

nml = 'lesgo_amd.nml'
if (command_argument_count() >= 1) call get_command_argument(1, nml)
open(newunit=fid, file=trim(nml), status='old', iostat=ios)
if (ios /= 0) then
    write(*,*) 'cannot open ', trim(nml)
    stop 1
end if
read(fid, nml=lesgo)
close(fid)

! derived sizes (input_util.f90:196-235), one process
nz = Nz + 1
nz_tot = nz
nx2 = 3*Nx/2; ny2 = 3*Ny/2
lh = Nx/2 + 1; ld = 2*lh
lh_big = nx2/2 + 1; ld_big = 2*lh_big
dx = L_x/Nx; dy = L_y/Ny; dz = L_z/(nz_tot - 1)

call sim_param_init()
call sgs_param_init()
! a checkpoint in path is read as the reference reads it (initial.f90:96-110,
! ic_file 226-239): initu = .true.; otherwise the seeded start
inquire(file=trim(path)//trim(checkpoint_file), exist=restart)
if (restart) initu = .true.
call gpu_init()
if (restart) then
    call ic_file()
else
    call gpu_check(lg_seed_state(gpu_case, int(seed, c_long), 0.1_rprec), 'lg_seed_state')
end if
if (turbines_on) then
    call aligned_farm(farm)
    call gpu_check(lg_turbines_init(gpu_case, int(size(farm)/8, c_int), farm,      &
        1.5_rprec, 1.5_rprec, 1.e-2_rprec, 1_c_int), 'lg_turbines_init')
end if
call system('mkdir -p ' // trim(path) // 'output')

call system_clock(count0, rate)
time_loop: do jt_step = 1, nsteps
    jt_total = jt_total + 1
    jt = jt + 1
    if (.not. use_cfl_dt) total_time = total_time + dt
    if (fused) then
        call gpu_check(lg_step(gpu_case, 1_c_int), 'lg_step')
    else
        call gpu_counters()
        ! dt = get_cfl_dt(), tadv1, tadv2, total_time (main.f90:135-151)
        call gpu_check(lg_advance_clock(gpu_case), 'lg_advance_clock')
        ! RHS_f <- RHS (main.f90:155-157)
        call gpu_check(lg_copy_field(gpu_case, LG_RHSX, LG_RHSX_F), 'copy')
        call gpu_check(lg_copy_field(gpu_case, LG_RHSY, LG_RHSY_F), 'copy')
        call gpu_check(lg_copy_field(gpu_case, LG_RHSZ, LG_RHSZ_F), 'copy')
        call filt_da(u, dudx, dudy, lbz)
        call filt_da(v, dvdx, dvdy, lbz)
        call filt_da(w, dwdx, dwdy, lbz)
        call ddz_uv(u, dudz, lbz)
        call ddz_uv(v, dvdz, lbz)
        call ddz_w(w, dwdz, lbz)
        call wallstress()
        call sgs_stag()
        call divstress_uv(divtx, divty, txx, txy, txz, tyy, tyz)
        call divstress_w(divtz, txz, tyz, tzz)
        call convec()
        call forcing_applied()
        call gpu_check(lg_rhs_ab2(gpu_case), 'rhs_ab2')
        call press_stag_array()
        call gpu_check(lg_rhs_gradp(gpu_case), 'rhs_gradp')
        call project()
    end if
    if (modulo(jt_total, domain_nskip) == 0) call write_domain()
    if (checkpoint_data .and. modulo(jt_total, checkpoint_nskip) == 0) call checkpoint()
    if (modulo(jt_total, wbase) == 0) then
        call clock_sync()
        call rmsdiv(rms)
        call gpu_check(lg_energy(gpu_case, ke), 'energy')
        call gpu_check(lg_tau_wall_bot(gpu_case, tw), 'tau_wall')
        write(*,'(a)') '==================================================='
        write(*,'(a,i9)') '  Iteration: ', jt_total
        write(*,'(a,E15.7)') '  Velocity divergence metric: ', rms
        write(*,'(a,E15.7)') '  Kinetic energy: ', ke
        write(*,'(a,E15.7)') '  Bot wall stress: ', tw
        write(*,'(a,E15.7)') '  Time step: ', dt
        write(*,'(a,2E15.7)') '  AB2 TADV1, TADV2: ', tadv1, tadv2
    end if
end do time_loop
call gpu_check(lg_synchronize(gpu_case), 'synchronize')
if (checkpoint_data) call checkpoint()
call system_clock(count1)
write(*,"(a,e15.7)") 'Simulation wall time (s) : ', real(count1 - count0, rprec)/rate
call gpu_check(lg_destroy(gpu_case), 'lg_destroy')
write(*,'(a)') 'Simulation complete'

contains

subroutine ic_file()
! vel.out (initial.f90:226-239) and total_time.dat (jt_total, total_time)
integer :: un, jt0
real(rprec) :: t0, tdim, dt0, cfl0
open(newunit=un, file=trim(path)//trim(checkpoint_file), form='unformatted', status='old')
read(un) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz),                          &
    RHSx(:, :, 1:nz), RHSy(:, :, 1:nz), RHSz(:, :, 1:nz),                      &
    Cs_opt2(:,:,1:nz), F_LM(:,:,1:nz), F_MM(:,:,1:nz),                         &
    F_QN(:,:,1:nz), F_NN(:,:,1:nz)
close(un)
u(:,:,0) = BOGUS; v(:,:,0) = BOGUS; w(:,:,0) = BOGUS
call gpu_upload(u); call gpu_upload(v); call gpu_upload(w)
call gpu_upload(RHSx); call gpu_upload(RHSy); call gpu_upload(RHSz)
call gpu_upload(Cs_opt2); call gpu_upload(F_LM); call gpu_upload(F_MM)
call gpu_upload(F_QN); call gpu_upload(F_NN)
open(newunit=un, file=trim(path)//'total_time.dat', status='old', iostat=ios)
if (ios == 0) then
    read(un, *) jt0, t0, tdim, dt0, cfl0
    close(un)
    jt_total = jt0
    total_time = t0
    ! with use_cfl_dt the file's dt is the previous step's; a new run or a
    ! changed cfl restarts with an Euler step (io.f90:91-94,
    ! initialize.f90:191-200)
    if (use_cfl_dt) dt = dt0
    call gpu_check(lg_set_clock_member(gpu_case, 0_c_int, dt, total_time,                 &
        merge(1_c_int, 0_c_int, use_cfl_dt .and. (jt0 == 0 .or.                          &
        abs((cfl0 - cfl)/cfl) > 1.e-2_rprec))), 'lg_set_clock_member')
end if
call gpu_counters()
end subroutine ic_file

subroutine clock_sync()
! dt, tadv1, tadv2, total_time of the last step from the device clock
real(c_double) :: ck(5)
if (.not. use_cfl_dt) return
call gpu_check(lg_clock_member(gpu_case, 0_c_int, ck), 'lg_clock_member')
dt = ck(1); tadv1 = ck(2); tadv2 = ck(3); total_time = ck(5)
end subroutine clock_sync

subroutine checkpoint()
! io.f90:1173-1267: vel.out, grid.out, total_time.dat (one process)
integer :: un
real(rprec) :: cfl_w
! the written fields are the state from here on, as in a restart from them
call gpu_check(lg_resync(gpu_case), 'lg_resync')
call gpu_download(u); call gpu_download(v); call gpu_download(w)
call gpu_download(RHSx); call gpu_download(RHSy); call gpu_download(RHSz)
call gpu_download(Cs_opt2); call gpu_download(F_LM); call gpu_download(F_MM)
call gpu_download(F_QN); call gpu_download(F_NN)
open(newunit=un, file=trim(path)//trim(checkpoint_file), form='unformatted',      &
    status='replace')
write(un) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz),                         &
    RHSx(:, :, 1:nz), RHSy(:, :, 1:nz), RHSz(:, :, 1:nz),                      &
    Cs_opt2(:,:,1:nz), F_LM(:,:,1:nz), F_MM(:,:,1:nz),                         &
    F_QN(:,:,1:nz), F_NN(:,:,1:nz)
close(un)
open(newunit=un, file=trim(path)//'grid.out', form='unformatted', status='replace')
write(un) 1, Nx, Ny, Nz, L_x, L_y, L_z
close(un)
! the cfl written is the target with use_cfl_dt, the measured one otherwise
! (io.f90:1232-1238)
call clock_sync()
if (use_cfl_dt) then
    cfl_w = cfl
else
    call gpu_check(lg_max_cfl(gpu_case, cfl_w), 'max_cfl')
end if
open(newunit=un, file=trim(path)//'total_time.dat', status='replace')
write(un, *) jt_total, total_time, total_time, dt, cfl_w
close(un)
end subroutine checkpoint

subroutine write_domain()
! io.f90:786-884 (binary branch): u, v, w on the uv grid; p - |u|^2/2
real(rprec), allocatable :: w_uv(:,:,:), pr(:,:,:)
character(len=512) :: fname
integer :: un
call gpu_download(u)
call gpu_download(v)
call gpu_download(w)
call gpu_download(p)
allocate(w_uv(Nx,Ny,1:nz), pr(Nx,Ny,1:nz))
w_uv(:,:,1:nz-1) = 0.5_rprec*(w(1:Nx,:,2:nz) + w(1:Nx,:,1:nz-1))
w_uv(:,:,nz) = w_uv(:,:,nz-1)
pr = p(1:Nx,:,1:nz) - 0.5*(u(1:Nx,:,1:nz)**2 + w_uv**2 + v(1:Nx,:,1:nz)**2)
write(fname,'(a,i0,a)') trim(path)//'output/vel.', jt_total, '.c0.bin'
open(newunit=un, file=trim(fname), form='unformatted', access='direct',      &
    recl=Nx*Ny*nz*rprec, status='replace')
write(un, rec=1) u(1:Nx,:,1:nz)
write(un, rec=2) v(1:Nx,:,1:nz)
write(un, rec=3) w_uv
close(un)
write(fname,'(a,i0,a)') trim(path)//'output/pres.', jt_total, '.c0.bin'
open(newunit=un, file=trim(fname), form='unformatted', access='direct',      &
    recl=Nx*Ny*nz*rprec, status='replace')
write(un, rec=1) pr
close(un)
end subroutine write_domain

subroutine aligned_farm(spec)
! 4x4 aligned array centred in the domain, 7D x 5D spacing, D = hub = 0.1
real(rprec), allocatable, intent(out) :: spec(:)
integer :: i, j, k
real(rprec) :: x0, y0
allocate(spec(8*16))
x0 = 0.5_rprec*L_x - 0.5_rprec*3*7._rprec*0.1_rprec
y0 = 0.5_rprec*L_y - 0.5_rprec*3*5._rprec*0.1_rprec
k = 0
do i = 0, 3
    do j = 0, 3
        spec(k+1:k+8) = [x0 + i*7._rprec*0.1_rprec, y0 + j*5._rprec*0.1_rprec,    &
            0.1_rprec, 0.1_rprec, 0._rprec, 0._rprec, 0._rprec, 1.33_rprec]
        k = k + 8
    end do
end do
end subroutine aligned_farm

end program lesgo_amd_main
