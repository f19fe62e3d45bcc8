! ps3d_capi.f90 -- ISO_C_BINDING interfaces to libps3d.so (include/ps3d.h).
!
! Self-contained: needs no other module.  Every function returns 0 on success;
! ps3d_error_message() fetches the message of the last failure.
module ps3d_capi
 use iso_c_binding
 implicit none

 integer(c_int), parameter :: PS3D_OK = 0
 integer(c_int), parameter :: PS3D_FP64 = 0, PS3D_MIXED = 1
 integer(c_int), parameter :: PS3D_U = 0, PS3D_V = 1, PS3D_W = 2, PS3D_B = 3, PS3D_P = 4
 integer(c_int), parameter :: PS3D_DU = 5, PS3D_DV = 6, PS3D_DW = 7, PS3D_DB = 8
 integer(c_int), parameter :: PS3D_FS = 9, PS3D_KSQR = 10
 integer(c_int), parameter :: PS3D_HARMONIC = 0, PS3D_BIHARMONIC_H = 1, PS3D_BIHARMONIC_V = 2

 ! struct ps3d_params
 type, bind(C) :: ps3d_params
  integer(c_int32_t) :: nx, ny, nz
  integer(c_int32_t) :: enable_nonlinear, enable_vertical_boundaries, enable_AB_3_order
  real(c_double) :: Lx, Ly, Lz, dx, dy, dz, dt
  real(c_double) :: f0, N0, Ro, dsqr, eps_ab
  real(c_double) :: Ah, Av, Kh, Kv, Ahbi, Avbi, Khbi, Kvbi
  integer(c_int32_t) :: precision
  integer(c_int32_t) :: reserved(7)
 end type ps3d_params

 ! ps3d_opt_balance_params (include/ps3d.h)
 type, bind(C) :: ps3d_opt_balance_params
  real(c_double) :: period, average, tol
  integer(c_int32_t) :: max_iterations, average_times, time_average, reserved
 end type

 interface
  integer(c_int) function ps3d_last_error(buf, len) bind(C, name="ps3d_last_error")
   import :: c_int, c_char
   character(kind=c_char), intent(out) :: buf(*)
   integer(c_int), value :: len
  end function
  integer(c_int) function ps3d_create(prm, ksqr, stream, model) bind(C, name="ps3d_create")
   import :: c_int, c_double, c_ptr, ps3d_params
   type(ps3d_params), intent(in) :: prm
   real(c_double), intent(in) :: ksqr(*)
   type(c_ptr), value :: stream
   type(c_ptr), intent(out) :: model
  end function
  integer(c_int) function ps3d_destroy(model) bind(C, name="ps3d_destroy")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
  end function
  ! launch-shape autotuner; cache_path NUL-terminated ('' + c_null_char: no file)
  integer(c_int) function ps3d_tune(model, next_tau, cache_path, reps) bind(C, name="ps3d_tune")
   import :: c_int, c_ptr, c_char
   type(c_ptr), value :: model
   integer(c_int), value :: next_tau, reps
   character(kind=c_char), intent(in) :: cache_path(*)
  end function
  integer(c_int) function ps3d_set_field(model, field, host) bind(C, name="ps3d_set_field")
   import :: c_int, c_double, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: field
   real(c_double), intent(in) :: host(*)
  end function
  integer(c_int) function ps3d_get_field(model, field, host) bind(C, name="ps3d_get_field")
   import :: c_int, c_double, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: field
   real(c_double), intent(out) :: host(*)
  end function
  integer(c_int) function ps3d_integrate(model, itt, tau, taum1, taum2) bind(C, name="ps3d_integrate")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: itt, tau, taum1, taum2
  end function
  integer(c_int) function ps3d_run(model, nsteps, state, use_graph) bind(C, name="ps3d_run")
   import :: c_int, c_int32_t, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: nsteps, use_graph
   integer(c_int32_t), intent(inout) :: state(4)
  end function
  integer(c_int) function ps3d_synchronize(model) bind(C, name="ps3d_synchronize")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
  end function
  integer(c_int) function ps3d_tendency(model, tau) bind(C, name="ps3d_tendency")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: tau
  end function
  integer(c_int) function ps3d_friction(model) bind(C, name="ps3d_friction")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
  end function
  integer(c_int) function ps3d_mixing(model, op, field, hmix, vmix) bind(C, name="ps3d_mixing")
   import :: c_int, c_double, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: op, field
   real(c_double), value :: hmix, vmix
  end function
  integer(c_int) function ps3d_ab_update(model, itt, tau, taum1, taum2) bind(C, name="ps3d_ab_update")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: itt, tau, taum1, taum2
  end function
  integer(c_int) function ps3d_border_exchg(model, field) bind(C, name="ps3d_border_exchg")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: field
  end function
  integer(c_int) function ps3d_pressure_solver(model) bind(C, name="ps3d_pressure_solver")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
  end function
  integer(c_int) function ps3d_my_fft(model, three_d, qp, qs) bind(C, name="ps3d_my_fft")
   import :: c_int, c_double, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: three_d
   real(c_double), intent(in) :: qp(*)
   real(c_double), intent(out) :: qs(*)
  end function
  integer(c_int) function ps3d_my_ifft(model, three_d, qs, qp) bind(C, name="ps3d_my_ifft")
   import :: c_int, c_double, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: three_d
   real(c_double), intent(in) :: qs(*)
   real(c_double), intent(out) :: qp(*)
  end function
  ! diag_opt_balance.f90:55-179 (ps3d_opt_balance_params as a bind(C) type)
  integer(c_int) function ps3d_opt_balance(model, ob, kx, ky, kz, out, norms, iterations) &
    bind(C, name="ps3d_opt_balance")
   import :: c_int, c_int32_t, c_ptr, c_double, ps3d_opt_balance_params
   type(c_ptr), value :: model
   type(ps3d_opt_balance_params), intent(in) :: ob
   real(c_double), intent(in) :: kx(*), ky(*), kz(*)
   type(c_ptr), intent(in) :: out(4)
   real(c_double), intent(out) :: norms(*)
   integer(c_int32_t), intent(out) :: iterations
  end function
  ! set_forcing hook (integrate.f90:82): fn is a bind(C) function
  ! integer(c_int) fn(model, user, itt, tau, stream), passed with c_funloc
  integer(c_int) function ps3d_set_forcing(model, fn, user) bind(C, name="ps3d_set_forcing")
   import :: c_int, c_ptr, c_funptr
   type(c_ptr), value :: model
   type(c_funptr), value :: fn
   type(c_ptr), value :: user
  end function
  integer(c_int) function ps3d_device_ptr(model, field, slot, ptr) bind(C, name="ps3d_device_ptr")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
   integer(c_int), value :: field, slot
   type(c_ptr), intent(out) :: ptr
  end function
  ! wavemaker forcing of for_src/config.f90:224-237; re(f), im(f) = c_loc of
  ! u_r.. b_r and u_i.. b_i (c_null_ptr leaves a field unforced)
  integer(c_int) function ps3d_set_wavemaker(model, omega, re, im) bind(C, name="ps3d_set_wavemaker")
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: model
   real(c_double), value :: omega
   type(c_ptr), intent(in) :: re(4), im(4)
  end function
  integer(c_int) function ps3d_clear_wavemaker(model) bind(C, name="ps3d_clear_wavemaker")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
  end function
  ! lee_wave1.f90:239-314 background advection; u_back, b_back = c_loc of
  ! field-layout arrays (c_null_ptr: off / zero)
  integer(c_int) function ps3d_set_background_flow(model, u_back, b_back) bind(C, name="ps3d_set_background_flow")
   import :: c_int, c_ptr
   type(c_ptr), value :: model, u_back, b_back
  end function
  ! steady body force on du, dv, dw, db (c_loc of field-layout arrays or c_null_ptr)
  integer(c_int) function ps3d_set_body_force(model, f) bind(C, name="ps3d_set_body_force")
   import :: c_int, c_ptr
   type(c_ptr), value :: model
   type(c_ptr), intent(in) :: f(4)
  end function
  ! dt, Ro, friction, eps_ab, f0, N0 and the two switches of the next steps
  integer(c_int) function ps3d_set_params(model, prm) bind(C, name="ps3d_set_params")
   import :: c_int, c_ptr, ps3d_params
   type(c_ptr), value :: model
   type(ps3d_params), intent(in) :: prm
  end function
  ! the FFT library of my_fft.f90 on the device (op strings NUL-terminated)
  integer(c_int) function ps3d_fft_create(nx, ny, nz, stream, plan) bind(C, name="ps3d_fft_create")
   import :: c_int, c_ptr
   integer(c_int), value :: nx, ny, nz
   type(c_ptr), value :: stream
   type(c_ptr), intent(out) :: plan
  end function
  integer(c_int) function ps3d_fft_destroy(plan) bind(C, name="ps3d_fft_destroy")
   import :: c_int, c_ptr
   type(c_ptr), value :: plan
  end function
  integer(c_int) function ps3d_fft_forward(plan, op, qp, qs) bind(C, name="ps3d_fft_forward")
   import :: c_int, c_ptr, c_char, c_double
   type(c_ptr), value :: plan
   character(kind=c_char), intent(in) :: op(*)
   real(c_double), intent(in) :: qp(*)
   real(c_double), intent(out) :: qs(*)
  end function
  integer(c_int) function ps3d_fft_backward(plan, op, qs, qp) bind(C, name="ps3d_fft_backward")
   import :: c_int, c_ptr, c_char, c_double
   type(c_ptr), value :: plan
   character(kind=c_char), intent(in) :: op(*)
   real(c_double), intent(in) :: qs(*)
   real(c_double), intent(out) :: qp(*)
  end function
 end interface

contains

 ! the last error as a Fortran string
 function ps3d_error_message() result(msg)
  character(len=512) :: msg
  character(kind=c_char) :: buf(512)
  integer(c_int) :: n, i
  n = ps3d_last_error(buf, 512_c_int)
  msg = ' '
  do i = 1, min(int(n), 511)
   msg(i:i) = buf(i)
  enddo
 end function ps3d_error_message

end module ps3d_capi
