! ref_glue.f90 -- test harness around the reference's own Fortran.
!
! TEST INFRASTRUCTURE ONLY.  Linked with the reference sources compiled
! unchanged from for_src/ (main_module.f90, timing_module.f90, advection.f90,
! mixing.f90, integrate.f90, parallel_mpi.f90, eigenvectors.f90) into
! oracle/_ref/libps3d_ref.so, which the tests use to pin the CPU oracle.
!
! What this file supplies, and why:
!  * ref_setup: the one-rank decomposition that p3dfft_initialization
!    (my_fft.f90:5-82) would set, then the reference's allocate_main_module.
!  * set_forcing: the config hook, empty as in for_config/eddy1.f90:141-142.
!  * pressure_solver: a capture point.  pressure_solver.f90 calls the P3DFFT
!    wrappers of my_fft.f90, and P3DFFT is not in this image, so the
!    reference's pressure solve is not built; with this empty hook the
!    reference's integrate (integrate.f90:2-115) stops right before the
!    projection, which is the state the oracle's orc_integrate_pre returns.
module ref_glue
 use iso_c_binding
 use main_module
 implicit none

contains

 subroutine addr_of(a, ptr)
  real(real_type), target :: a(*)
  type(c_ptr), intent(out) :: ptr
  ptr = c_loc(a(1))
 end subroutine addr_of

 ! iprm = (nx, ny, nz, enable_nonlinear, enable_vertical_boundaries, enable_AB_3_order)
 ! rprm = (Lx, Ly, Lz, dx, dy, dz, dt, f0, N0, Ro, dsqr, eps_ab,
 !         Ah, Av, Kh, Kv, Ahbi, Avbi, Khbi, Kvbi)
 subroutine ref_setup(iprm, rprm) bind(C, name="ref_setup")
  integer(c_int), intent(in) :: iprm(6)
  real(c_double), intent(in) :: rprm(20)
  nx = iprm(1); ny = iprm(2); nz = iprm(3)
  enable_nonlinear = iprm(4) /= 0
  enable_vertical_boundaries = iprm(5) /= 0
  enable_AB_3_order = iprm(6) /= 0
  Lx = rprm(1); Ly = rprm(2); Lz = rprm(3)
  dx = rprm(4); dy = rprm(5); dz = rprm(6); dt = rprm(7)
  f0 = rprm(8); N0 = rprm(9); Ro = rprm(10); dsqr = rprm(11); eps_ab = rprm(12)
  Ah = rprm(13); Av = rprm(14); Kh = rprm(15); Kv = rprm(16)
  Ahbi = rprm(17); Avbi = rprm(18); Khbi = rprm(19); Kvbi = rprm(20)
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
 end subroutine ref_setup

 ! 0..4 u v w b p, 5..8 du dv dw db (all 3 slots), 9..11 flux_east/north/top
 function ref_ptr(id) bind(C, name="ref_ptr") result(ptr)
  integer(c_int), value :: id
  type(c_ptr) :: ptr
  select case (id)
  case (0); call addr_of(u, ptr)
  case (1); call addr_of(v, ptr)
  case (2); call addr_of(w, ptr)
  case (3); call addr_of(b, ptr)
  case (4); call addr_of(p, ptr)
  case (5); call addr_of(du, ptr)
  case (6); call addr_of(dv, ptr)
  case (7); call addr_of(dw, ptr)
  case (8); call addr_of(db, ptr)
  case (9); call addr_of(flux_east, ptr)
  case (10); call addr_of(flux_north, ptr)
  case (11); call addr_of(flux_top, ptr)
  case default; ptr = c_null_ptr
  end select
 end function ref_ptr

 subroutine ref_set_time(itt_, tau_, taum1_, taum2_) bind(C, name="ref_set_time")
  integer(c_int), value :: itt_, tau_, taum1_, taum2_
  itt = itt_; tau = tau_; taum1 = taum1_; taum2 = taum2_
 end subroutine ref_set_time

 ! main_module's model time (restart.f90 records it)
 subroutine ref_set_clock(t) bind(C, name="ref_set_clock")
  real(c_double), value :: t
  time = t
 end subroutine ref_set_clock

 ! advection.f90 into the current tau slot
 subroutine ref_buoyancy_advection() bind(C, name="ref_buoyancy_advection")
  call buoyancy_advection_2nd(db(:,:,:,tau))
 end subroutine ref_buoyancy_advection

 subroutine ref_momentum_advection() bind(C, name="ref_momentum_advection")
  call momentum_advection_2nd(du(:,:,:,tau), dv(:,:,:,tau), dw(:,:,:,tau))
 end subroutine ref_momentum_advection

 ! mixing.f90 on field id (0..3 = u v w b); op 0 harmonic, 1 biharmonic_h, 2 biharmonic_v
 subroutine ref_mixing(op, id, hmix, vmix) bind(C, name="ref_mixing")
  integer(c_int), value :: op, id
  real(c_double), value :: hmix, vmix
  select case (id)
  case (0); call mix_one(u)
  case (1); call mix_one(v)
  case (2); call mix_one(w)
  case (3); call mix_one(b)
  end select
 contains
  subroutine mix_one(a)
   real(real_type) :: a(*)
   select case (op)
   case (0); call harmonic(a, hmix, vmix)
   case (1); call biharmonic_h(a, hmix)
   case (2); call biharmonic_v(a, vmix)
   end select
  end subroutine mix_one
 end subroutine ref_mixing

 subroutine ref_border_exchg(id) bind(C, name="ref_border_exchg")
  integer(c_int), value :: id
  select case (id)
  case (0); call border_exchg_3D(u)
  case (1); call border_exchg_3D(v)
  case (2); call border_exchg_3D(w)
  case (3); call border_exchg_3D(b)
  case (4); call border_exchg_3D(p)
  end select
 end subroutine ref_border_exchg

 ! integrate.f90:2-115 up to the (captured) call of pressure_solver
 subroutine ref_integrate() bind(C, name="ref_integrate")
  call integrate
 end subroutine ref_integrate

 function ref_hat_ksqr(k, d) bind(C, name="ref_hat_ksqr") result(r)
  real(c_double), value :: k, d
  real(c_double) :: r
  real*8, external :: hat_ksqr
  r = hat_ksqr(k, d)
 end function ref_hat_ksqr

 function ref_ab3(i) bind(C, name="ref_ab3") result(r)
  integer(c_int), value :: i
  real(c_double) :: r
  select case (i)
  case (0); r = AB3_a
  case (1); r = AB3_b
  case default; r = AB3_c
  end select
 end function ref_ab3

end module ref_glue

subroutine set_forcing
end subroutine set_forcing

subroutine pressure_solver
end subroutine pressure_solver
