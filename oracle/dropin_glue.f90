! dropin_glue.f90 -- a driver for the Fortran drop-in of the GPU step.
!
! TEST INFRASTRUCTURE ONLY.  Linked with the reference's main_module.f90,
! timing_module.f90 and parallel_mpi.f90 compiled unchanged from for_src/,
! the drop-in integrate / pressure_solver (ps3d_amd/fortran/integrate_gpu.f90)
! and the C-ABI interface module (ps3d_amd/fortran/ps3d_capi.f90) into
! oracle/_ref/libps3d_dropin.so.  The tests fill main_module's arrays through
! dropin_ptr and step the model exactly as the main loop of main.f90:138-162
! does (call integrate, rotate the tau slots, advance itt and time), so the
! drop-in runs against the reference's own module layout and state.
!
! What this file supplies:
!  * dropin_setup: the one-rank decomposition that p3dfft_initialization
!    (my_fft.f90:5-82) would set, the parameters a config's set_parameter
!    sets, and the reference's allocate_main_module.
!  * dropin_run: the main loop, followed by a refresh of the host arrays;
!    dropin_pressure: one call of pressure_solver on the host arrays.
!  * set_forcing: the config hook; empty as in for_config/eddy1.f90:141-142
!    unless dropin_set_forcing switched on a forcing on the host mirrors
!    (db(:,:,:,tau) += a fixed array, whole arrays as lee_wave1.f90:321-324
!    updates them), which takes the drop-in's host phase path.
module dropin_glue
 use iso_c_binding
 use main_module
 implicit none
 real(real_type), allocatable :: dropin_force(:,:,:)

contains

 subroutine addr_of(a, ptr)
  real(real_type), target :: a(*)
  type(c_ptr), intent(out) :: ptr
  ptr = c_loc(a(1))
 end subroutine addr_of

 ! iprm = (nx, ny, nz, enable_nonlinear, enable_vertical_boundaries,
 !         enable_AB_3_order, mixed)
 ! rprm = (Lx, Ly, Lz, dx, dy, dz, dt, f0, N0, Ro, dsqr, eps_ab,
 !         Ah, Av, Kh, Kv, Ahbi, Avbi, Khbi, Kvbi)
 subroutine dropin_setup(iprm, rprm) bind(C, name="dropin_setup")
  use ps3d_gpu_state
  integer(c_int), intent(in) :: iprm(7)
  real(c_double), intent(in) :: rprm(20)
  nx = iprm(1); ny = iprm(2); nz = iprm(3)
  enable_nonlinear = iprm(4) /= 0
  enable_vertical_boundaries = iprm(5) /= 0
  enable_AB_3_order = iprm(6) /= 0
  ps3d_use_mixed = iprm(7) /= 0
  ps3d_forcing_on_host = .false.
  ps3d_autotune = .false.
  if (allocated(dropin_force)) deallocate(dropin_force)
  Lx = rprm(1); Ly = rprm(2); Lz = rprm(3)
  dx = rprm(4); dy = rprm(5); dz = rprm(6); dt = rprm(7)
  f0 = rprm(8); N0 = rprm(9); Ro = rprm(10); dsqr = rprm(11); eps_ab = rprm(12)
  Ah = rprm(13); Av = rprm(14); Kh = rprm(15); Kv = rprm(16)
  Ahbi = rprm(17); Avbi = rprm(18); Khbi = rprm(19); Kvbi = rprm(20)
  ! monitor and snapshot cadences beyond the run: the drop-in refreshes the
  ! host arrays after the first step only, and dropin_run once at the end
  tsmonint = 1d30; snapint = 1d30
  n_pes = 1; my_pe = 0; n_pes_x = 1; n_pes_y = 1
  n_pes_i = 1; n_pes_j = 1; n_pes_k = 1
  my_blk_i = 1; my_blk_j = 1; my_blk_k = 1
  istart = (/1, 1, 1/); iend = (/nx, ny, nz/); isize = iend
  fstart = (/1, 1, 1/); fend = (/nx/2+1, ny, nz/); fsize = fend
  is_pe = 1; ie_pe = nx; js_pe = 1; je_pe = ny; ks_pe = 1; ke_pe = nz
  if (allocated(u)) then
   deallocate(u, v, w, b, p, du, dv, dw, db, flux_east, flux_north, flux_top, ksqr, FS, kx, ky, kz)
   if (allocated(psi)) deallocate(psi, phi, psib)
  endif
  call allocate_main_module()
  ! main.f90:112-113
  time = 0.; itt = 0
  taum2 = 1; taum1 = 2; tau = 3
 end subroutine dropin_setup

 ! the drop-in's launch-shape autotuner at set-up (no cache file)
 subroutine dropin_set_autotune() bind(C, name="dropin_set_autotune")
  use ps3d_gpu_state
  ps3d_autotune = .true.
  ps3d_tune_cache = ''
 end subroutine dropin_set_autotune

 ! a forcing of db on the host mirrors (set its array through dropin_ptr(6))
 subroutine dropin_set_forcing() bind(C, name="dropin_set_forcing")
  use ps3d_gpu_state
  allocate(dropin_force(is_pe-onx:ie_pe+onx, js_pe-onx:je_pe+onx, ks_pe-onx:ke_pe+onx))
  dropin_force = 0d0
  ps3d_forcing_on_host = .true.
 end subroutine dropin_set_forcing

 ! 0..4 u v w b p, 5 ksqr, 6 the forcing array
 function dropin_ptr(id) bind(C, name="dropin_ptr") result(ptr)
  integer(c_int), value :: id
  type(c_ptr) :: ptr
  select case (id)
  case (0); call addr_of(u, ptr)
  case (1); call addr_of(v, ptr)
  case (2); call addr_of(w, ptr)
  case (3); call addr_of(b, ptr)
  case (4); call addr_of(p, ptr)
  case (5); call addr_of(ksqr, ptr)
  case (6)
   if (allocated(dropin_force)) then
    call addr_of(dropin_force, ptr)
   else
    ptr = c_null_ptr
   endif
  case default; ptr = c_null_ptr
  end select
 end function dropin_ptr

 ! nsteps passes of the main loop (main.f90:138-162 without the diagnostics);
 ! st returns (itt, tau, taum1, taum2)
 subroutine dropin_run(nsteps, st) bind(C, name="dropin_run")
  integer(c_int), value :: nsteps
  integer(c_int), intent(out) :: st(4)
  integer :: n
  do n = 1, nsteps
   call integrate
   tau   = mod(tau,3)+1
   taum1 = mod(taum1,3)+1
   taum2 = mod(taum2,3)+1
   time = time + dt
   itt = itt + 1
  enddo
  call ps3d_gpu_sync(.false.)
  st = (/itt, tau, taum1, taum2/)
 end subroutine dropin_run

 ! pressure_solver called outside integrate (as a config's initial
 ! conditions may call it), then the host arrays as it leaves them
 subroutine dropin_pressure() bind(C, name="dropin_pressure")
  call pressure_solver
 end subroutine dropin_pressure

 ! the device model of the drop-in is released (a new setup starts another)
 subroutine dropin_finish() bind(C, name="dropin_finish")
  use ps3d_capi
  use ps3d_gpu_state
  integer(c_int) :: ierr
  if (c_associated(ps3d_model_handle)) then
   ierr = ps3d_destroy(ps3d_model_handle)
   ps3d_model_handle = c_null_ptr
  endif
 end subroutine dropin_finish

end module dropin_glue


! the config hook (integrate.f90:82)
subroutine set_forcing
 use main_module
 use dropin_glue
 if (allocated(dropin_force)) db(:,:,:,tau) = db(:,:,:,tau) + dropin_force
end subroutine set_forcing
