! config_glue.f90 -- a reference config, linked unchanged, driven through the
! Fortran drop-in of the GPU step.
!
! TEST INFRASTRUCTURE ONLY.  Linked with the reference's main_module.f90,
! timing_module.f90, parallel_mpi.f90 and one config (for_config/
! jets_unscaled.f90: set_parameter, set_initial_conditions, set_forcing),
! all compiled unchanged, the drop-in integrate / pressure_solver
! (ps3d_amd/fortran/integrate_gpu.f90) and the C-ABI interface module, into
! oracle/_ref/libps3d_jets.so.  The calls follow main.f90: set_parameter
! (:43), the one-rank decomposition that p3dfft_initialization
! (my_fft.f90:5-82) would set, allocate_main_module (:74), the clock and tau
! slots (:112-113), set_initial_conditions and the lid zeroing of w
! (:117-122), then the main loop (:138-162).  wavenumber_initialisation
! (my_fft.f90:87-156) needs P3DFFT and is not built: the caller fills ksqr.
module config_glue
 use iso_c_binding
 use main_module
 implicit none

contains

 subroutine addr_of(a, ptr)
  real(real_type), target :: a(*)
  type(c_ptr), intent(out) :: ptr
  ptr = c_loc(a(1))
 end subroutine addr_of

 subroutine config_setup() bind(C, name="config_setup")
  call set_parameter()
  n_pes = 1; my_pe = 0; n_pes_x = 1; n_pes_y = 1
  n_pes_i = 1; n_pes_j = 1; n_pes_k = 1
  my_blk_i = 1; my_blk_j = 1; my_blk_k = 1
  istart = (/1, 1, 1/); iend = (/nx, ny, nz/); isize = iend
  fstart = (/1, 1, 1/); fend = (/nx/2+1, ny, nz/); fsize = fend
  is_pe = 1; ie_pe = nx; js_pe = 1; je_pe = ny; ks_pe = 1; ke_pe = nz
  call allocate_main_module()
  time = 0.; itt = 0
  taum2 = 1; taum1 = 2; tau = 3
 end subroutine config_setup

 ! the parameters set_parameter left in main_module
 ! iprm = (nx, ny, nz, enable_nonlinear, enable_vertical_boundaries, enable_AB_3_order)
 ! rprm = (Lx, Ly, Lz, dx, dy, dz, dt, f0, N0, Ro, dsqr, eps_ab,
 !         Ah, Av, Kh, Kv, Ahbi, Avbi, Khbi, Kvbi)
 subroutine config_params(iprm, rprm) bind(C, name="config_params")
  integer(c_int), intent(out) :: iprm(6)
  real(c_double), intent(out) :: rprm(20)
  iprm = (/nx, ny, nz, merge(1, 0, enable_nonlinear), merge(1, 0, enable_vertical_boundaries), &
           merge(1, 0, enable_AB_3_order)/)
  rprm = (/Lx, Ly, Lz, dx, dy, dz, dt, f0, N0, Ro, dsqr, eps_ab, Ah, Av, Kh, Kv, Ahbi, Avbi, Khbi, Kvbi/)
 end subroutine config_params

 subroutine config_initial_conditions() bind(C, name="config_initial_conditions")
  call set_initial_conditions()
  if (enable_vertical_boundaries) then
   if (my_blk_k == 1      ) w(:,:,1-onx:0) = 0d0
   if (my_blk_k == n_pes_k) w(:,:,nz:nz+onx) = 0d0
  endif
 end subroutine config_initial_conditions

 ! 0..4 u v w b p, 5 ksqr
 function config_ptr(id) bind(C, name="config_ptr") result(ptr)
  integer(c_int), value :: id
  type(c_ptr) :: ptr
  select case (id)
  case (0); call addr_of(u, ptr)
  case (1); call addr_of(v, ptr)
  case (2); call addr_of(w, ptr)
  case (3); call addr_of(b, ptr)
  case (4); call addr_of(p, ptr)
  case (5); call addr_of(ksqr, ptr)
  case default; ptr = c_null_ptr
  end select
 end function config_ptr

 ! nsteps passes of the main loop (without the diagnostics), then the host
 ! arrays refreshed; st returns (itt, tau, taum1, taum2)
 subroutine config_run(nsteps, st) bind(C, name="config_run")
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
 end subroutine config_run

 subroutine config_finish() bind(C, name="config_finish")
  use ps3d_capi
  use ps3d_gpu_state
  integer(c_int) :: ierr
  if (c_associated(ps3d_model_handle)) then
   ierr = ps3d_destroy(ps3d_model_handle)
   ps3d_model_handle = c_null_ptr
  endif
 end subroutine config_finish

end module config_glue
