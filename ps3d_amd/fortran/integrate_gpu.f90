! integrate_gpu.f90 -- drop-in replacements of integrate (for_src/integrate.f90)
! and pressure_solver (for_src/pressure_solver.f90) that run the step on the
! GPU through libps3d.so.  Link this file instead of those two; main.f90, the
! rest of for_src/ and any for_config/*.f90 link unchanged.
!
! The driver keeps owning main_module's host arrays.  ps3d_gpu_init (called
! from the first integrate) copies them to the device.  Calls reach
! integrate from two kinds of loops:
!  * the main loop (main.f90:139-161): each call continues the last one
!    (itt advanced by one, the tau slots rotated, dt, Ro and friction
!    unchanged).  The device state is authoritative and the host mirrors
!    are refreshed only where the driver reads them: before diagnose and
!    snapshots, on the cadences of main.f90:144-145.
!  * loops that set the state on the host and integrate it themselves:
!    diag_balance's model_run (diag_balance.f90:465-497: u = u_in, itt = 0,
!    ten calls, then the saved state put back) and diag_opt_balance
!    (diag_opt_balance.f90:190-292: dt = -|dt|, Ro ramped, friction off,
!    itt = 0 for every step).  Any call that does not continue the last one
!    is one of these: the host arrays (u, v, w, b, p, du..db) and the
!    parameters are pushed, the step runs, and the results come back to the
!    host, where such a caller reads them.  The main loop's next call then
!    starts again from the host arrays, which the caller has restored --
!    also when its counters happen to continue the outside caller's last
!    call (model_run keeps itt = 0 and rotates the tau slots, so the main
!    loop's first call after read_restart at itt = 1 can look like its
!    continuation): after a call from outside the main loop the next call
!    always starts from the host arrays.
! Around set_forcing, when the config's forcing works on the fields
! (ps3d_forcing_on_host = .true.), tendencies and fields go to the host,
! set_forcing runs there (integrate.f90:82), and the results come back.
! The mixed form keeps its tendency history on the device only, so it
! cannot resume the main loop after such a caller; it stops with a message
! pointing to ps3d_diag_balance / ps3d_opt_balance instead.
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
 ! the last call of integrate, (itt, tau, taum1, taum2), and the parameters
 ! the device model steps with
 integer :: ps3d_last(4) = -1
 real(8) :: ps3d_dyn(13) = 0d0
 logical :: ps3d_switches(2) = .false.
 ! the last call came from outside the main loop: its caller may have
 ! changed the host arrays since, so the next call starts from them
 logical :: ps3d_host_ahead = .false.
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


! main_module's parameters as the C-ABI takes them
subroutine ps3d_gpu_params(prm)
 use main_module
 use ps3d_capi
 use ps3d_gpu_state
 implicit none
 type(ps3d_params), intent(out) :: prm
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
end subroutine ps3d_gpu_params


! the parameters that callers change between steps (diag_opt_balance.f90)
subroutine ps3d_gpu_dynamic(dyn, sw)
 use main_module
 implicit none
 real(8), intent(out) :: dyn(13)
 logical, intent(out) :: sw(2)
 dyn = (/dt, Ro, eps_ab, f0, N0, Ah, Av, Kh, Kv, Ahbi, Avbi, Khbi, Kvbi/)
 sw = (/enable_nonlinear, enable_AB_3_order/)
end subroutine ps3d_gpu_dynamic


subroutine ps3d_gpu_init
 use main_module
 use ps3d_capi
 use ps3d_gpu_state
 implicit none
 type(ps3d_params) :: prm
 call ps3d_gpu_params(prm)
 if (n_pes /= 1) call halt_stop('ps3d: the GPU step runs the whole domain on one rank')
 if (ps3d_use_mixed .and. ps3d_forcing_on_host) &
  call halt_stop('ps3d: set_forcing on host mirrors needs the FP64 form (ps3d_use_mixed = .false.)')
 call ps3d_gpu_check(ps3d_create(prm, ksqr, c_null_ptr, ps3d_model_handle), 'ps3d_create')
 call ps3d_gpu_dynamic(ps3d_dyn, ps3d_switches)
 ps3d_last = -1
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
 logical :: need_host, resumes, first, sw(2)
 real(8) :: dyn(13)
 type(ps3d_params) :: prm

 first = .not. c_associated(ps3d_model_handle)
 if (first) call ps3d_gpu_init

 ! a call of the main loop continues the last one; any other call starts
 ! from the host arrays and parameters (see the head of this file)
 call ps3d_gpu_dynamic(dyn, sw)
 resumes = .not. first .and. itt == ps3d_last(1) + 1 .and. tau == mod(ps3d_last(2), 3) + 1 &
           .and. taum1 == mod(ps3d_last(3), 3) + 1 .and. taum2 == mod(ps3d_last(4), 3) + 1 &
           .and. all(dyn == ps3d_dyn) .and. all(sw .eqv. ps3d_switches)
 if (.not. resumes) then
  if (ps3d_use_mixed .and. .not. first) &
   call halt_stop('ps3d: integrate called outside the main loop (diag_balance model_run or diag_opt_balance);'// &
                  ' the mixed form cannot resume its device-side history after that: use the FP64 form,'// &
                  ' or ps3d_diag_balance / ps3d_opt_balance')
  if (any(dyn /= ps3d_dyn) .or. any(sw .neqv. ps3d_switches)) then
   call ps3d_gpu_params(prm)
   call ps3d_gpu_check(ps3d_set_params(ps3d_model_handle, prm), 'ps3d_set_params')
   ps3d_dyn = dyn
   ps3d_switches = sw
  endif
  if (.not. first) call ps3d_gpu_sync(.true.)
 else if (ps3d_host_ahead) then
  call ps3d_gpu_sync(.true.)
 endif
 ps3d_host_ahead = .not. resumes .and. .not. first
 ps3d_last = (/itt, tau, taum1, taum2/)

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
 ! (and always after a call from outside the main loop, whose caller reads
 ! the results from the host arrays)
 need_host = .not. resumes .or. itt == 0 .or. mod(itt, max(1, int(tsmonint/dt))) == 0 &
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
