! integrate_gpu.f90 -- drop-in replacements of integrate (for_src/integrate.f90)
! and pressure_solver (for_src/pressure_solver.f90) that run the step on the
! GPU through libps3d.so.  Link this file instead of those two; main.f90, the
! rest of for_src/ and any for_config/*.f90 link unchanged.
!
! The driver keeps owning main_module's host arrays.  ps3d_gpu_init (called
! once, after set_initial_conditions and read_restart, e.g. from the first
! integrate) copies them to the device; afterwards the device state is
! authoritative and the host mirrors are refreshed only where the driver
! reads them:
!  * before diagnose / snapshots, on the cadences of main.f90:144-145;
!  * around set_forcing when the config's forcing works on the fields
!    (ps3d_forcing_on_host = .true.): tendencies and fields go to the host,
!    set_forcing runs there (integrate.f90:82), and the results come back.
module ps3d_gpu_state
 use iso_c_binding
 implicit none
 type(c_ptr) :: ps3d_model_handle = c_null_ptr
 logical :: ps3d_forcing_on_host = .false.
 logical :: ps3d_use_mixed = .false.
 ! time the launch shapes once at set-up (results do not depend on them),
 ! cached in ps3d_tune_cache ('' : not cached)
 logical :: ps3d_autotune = .false.
 character(len=256) :: ps3d_tune_cache = ''
end module ps3d_gpu_state


subroutine ps3d_gpu_check(ierr, where)
 use ps3d_capi
 implicit none
 integer(c_int), intent(in) :: ierr
 character(len=*), intent(in) :: where
 if (ierr /= PS3D_OK) call halt_stop(where//': '//trim(ps3d_error_message()))
end subroutine ps3d_gpu_check


subroutine ps3d_gpu_sync(to_device)
 use main_module
 use ps3d_capi
 use ps3d_gpu_state
 implicit none
 logical, intent(in) :: to_device
 if (to_device) then
  call ps3d_gpu_check(ps3d_set_field(ps3d_model_handle, PS3D_U, u), 'ps3d_set_field u')
  call ps3d_gpu_check(ps3d_set_field(ps3d_model_handle, PS3D_V, v), 'ps3d_set_field v')
  call ps3d_gpu_check(ps3d_set_field(ps3d_model_handle, PS3D_W, w), 'ps3d_set_field w')
  call ps3d_gpu_check(ps3d_set_field(ps3d_model_handle, PS3D_B, b), 'ps3d_set_field b')
  call ps3d_gpu_check(ps3d_set_field(ps3d_model_handle, PS3D_P, p), 'ps3d_set_field p')
  ! a mixed model's tendency slots hold its own (pressure-referenced FP32)
  ! quantities and are not exchanged with the host
  if (ps3d_use_mixed) return
  call ps3d_gpu_check(ps3d_set_field(ps3d_model_handle, PS3D_DU, du), 'ps3d_set_field du')
  call ps3d_gpu_check(ps3d_set_field(ps3d_model_handle, PS3D_DV, dv), 'ps3d_set_field dv')
  call ps3d_gpu_check(ps3d_set_field(ps3d_model_handle, PS3D_DW, dw), 'ps3d_set_field dw')
  call ps3d_gpu_check(ps3d_set_field(ps3d_model_handle, PS3D_DB, db), 'ps3d_set_field db')
 else
  call ps3d_gpu_check(ps3d_get_field(ps3d_model_handle, PS3D_U, u), 'ps3d_get_field u')
  call ps3d_gpu_check(ps3d_get_field(ps3d_model_handle, PS3D_V, v), 'ps3d_get_field v')
  call ps3d_gpu_check(ps3d_get_field(ps3d_model_handle, PS3D_W, w), 'ps3d_get_field w')
  call ps3d_gpu_check(ps3d_get_field(ps3d_model_handle, PS3D_B, b), 'ps3d_get_field b')
  call ps3d_gpu_check(ps3d_get_field(ps3d_model_handle, PS3D_P, p), 'ps3d_get_field p')
  if (ps3d_use_mixed) return
  call ps3d_gpu_check(ps3d_get_field(ps3d_model_handle, PS3D_DU, du), 'ps3d_get_field du')
  call ps3d_gpu_check(ps3d_get_field(ps3d_model_handle, PS3D_DV, dv), 'ps3d_get_field dv')
  call ps3d_gpu_check(ps3d_get_field(ps3d_model_handle, PS3D_DW, dw), 'ps3d_get_field dw')
  call ps3d_gpu_check(ps3d_get_field(ps3d_model_handle, PS3D_DB, db), 'ps3d_get_field db')
 endif
end subroutine ps3d_gpu_sync


subroutine ps3d_gpu_init
 use main_module
 use ps3d_capi
 use ps3d_gpu_state
 implicit none
 type(ps3d_params) :: prm
 prm%nx = nx; prm%ny = ny; prm%nz = nz
 prm%enable_nonlinear = merge(1, 0, enable_nonlinear)
 prm%enable_vertical_boundaries = merge(1, 0, enable_vertical_boundaries)
 prm%enable_AB_3_order = merge(1, 0, enable_AB_3_order)
 prm%Lx = Lx; prm%Ly = Ly; prm%Lz = Lz
 prm%dx = dx; prm%dy = dy; prm%dz = dz; prm%dt = dt
 prm%f0 = f0; prm%N0 = N0; prm%Ro = Ro; prm%dsqr = dsqr; prm%eps_ab = eps_ab
 prm%Ah = Ah; prm%Av = Av; prm%Kh = Kh; prm%Kv = Kv
 prm%Ahbi = Ahbi; prm%Avbi = Avbi; prm%Khbi = Khbi; prm%Kvbi = Kvbi
 prm%precision = merge(PS3D_MIXED, PS3D_FP64, ps3d_use_mixed)
 prm%reserved = 0
 if (n_pes /= 1) call halt_stop('ps3d: the GPU step runs the whole domain on one rank')
 if (ps3d_use_mixed .and. ps3d_forcing_on_host) &
  call halt_stop('ps3d: set_forcing on host mirrors needs the FP64 form (ps3d_use_mixed = .false.)')
 call ps3d_gpu_check(ps3d_create(prm, ksqr, c_null_ptr, ps3d_model_handle), 'ps3d_create')
 call ps3d_gpu_sync(.true.)
 if (ps3d_autotune) call ps3d_gpu_check(ps3d_tune(ps3d_model_handle, int(tau, c_int), &
                                        trim(ps3d_tune_cache)//c_null_char, 5_c_int), 'ps3d_tune')
end subroutine ps3d_gpu_init


! integrate.f90:2-115 on the GPU
subroutine integrate
 use main_module
 use ps3d_capi
 use ps3d_gpu_state
 use timing_module
 implicit none
 logical :: need_host

 if (.not. c_associated(ps3d_model_handle)) call ps3d_gpu_init

 call tic('integrate')
 if (.not. ps3d_forcing_on_host) then
  call ps3d_gpu_check(ps3d_integrate(ps3d_model_handle, itt, tau, taum1, taum2), 'ps3d_integrate')
 else
  call ps3d_gpu_check(ps3d_tendency(ps3d_model_handle, tau), 'ps3d_tendency')
  call ps3d_gpu_check(ps3d_friction(ps3d_model_handle), 'ps3d_friction')
  call ps3d_gpu_sync(.false.)
  call set_forcing()
  call ps3d_gpu_sync(.true.)
  call ps3d_gpu_check(ps3d_ab_update(ps3d_model_handle, itt, tau, taum1, taum2), 'ps3d_ab_update')
  call ps3d_gpu_check(ps3d_pressure_solver(ps3d_model_handle), 'ps3d_pressure_solver')
 endif
 call toc('integrate')

 ! refresh the host mirrors where the driver reads them next (main.f90:144-145)
 need_host = itt == 0 .or. mod(itt, max(1, int(tsmonint/dt))) == 0 &
             .or. mod(itt, max(1, int(snapint/dt))) == 0
 if (need_host) call ps3d_gpu_sync(.false.)
end subroutine integrate


! pressure_solver.f90:2-120 for callers outside integrate (the state on the
! host is pushed, projected on the GPU and pulled back)
subroutine pressure_solver
 use main_module
 use ps3d_capi
 use ps3d_gpu_state
 implicit none
 if (.not. c_associated(ps3d_model_handle)) call ps3d_gpu_init
 call ps3d_gpu_sync(.true.)
 call ps3d_gpu_check(ps3d_pressure_solver(ps3d_model_handle), 'ps3d_pressure_solver')
 call ps3d_gpu_sync(.false.)
end subroutine pressure_solver
