! p3dfft_gpu.f90 -- module p3dfft for my_fft.f90, on one GPU.
!
! my_fft.f90 (for_src/my_fft.f90:5-82, :159-252) reaches its transforms
! through the distributed FFT library P3DFFT (2.7.9, make.inc:10):
! p3dfft_setup, p3dfft_get_dims, p3dfft_ftran_r2c and p3dfft_btran_c2r.
! This module provides those procedures with P3DFFT's interface and
! semantics for the case this build runs: one GPU holding the whole domain,
! one rank.  The transforms run on the device through libps3d.so
! (ps3d_fft_* in include/ps3d.h); the pencil transposes P3DFFT does over MPI
! become the LDS transposes inside the FFT kernels.  Link it in place of
! P3DFFT; my_fft.f90, main.f90 and the diagnostics compile unchanged
! against it.
!
! Semantics kept from P3DFFT: forward r2c unnormalised with e^{-ikx}; the
! spectral array is (nx/2+1, ny, nz) with x fastest (P3DFFT built without
! STRIDE1, the layout my_fft.f90 and diag_balance.f90 index); op strings
! 'fft' and 'ffn' forward, 'fft' and 'nff' backward.  Errors stop the run
! through the driver's halt_stop (parallel_mpi.f90:5-18), as P3DFFT's own
! errors abort MPI.
module p3dfft
 use iso_c_binding
 implicit none
 private
 integer, parameter, public :: mytype = 8
 integer, parameter, public :: p3dfft_type = 8
 public :: p3dfft_setup, p3dfft_get_dims, p3dfft_ftran_r2c, p3dfft_btran_c2r, p3dfft_clean

 type(c_ptr), save :: plan = c_null_ptr
 integer, save :: gnx = 0, gny = 0, gnz = 0

 interface
  integer(c_int) function c_fft_create(nx, ny, nz, stream, out) bind(C, name="ps3d_fft_create")
   import :: c_int, c_ptr
   integer(c_int), value :: nx, ny, nz
   type(c_ptr), value :: stream
   type(c_ptr), intent(out) :: out
  end function
  integer(c_int) function c_fft_destroy(f) bind(C, name="ps3d_fft_destroy")
   import :: c_int, c_ptr
   type(c_ptr), value :: f
  end function
  integer(c_int) function c_fft_forward(f, op, qp, qs) bind(C, name="ps3d_fft_forward")
   import :: c_int, c_ptr, c_char, c_double
   type(c_ptr), value :: f
   character(kind=c_char), intent(in) :: op(*)
   real(c_double), intent(in) :: qp(*)
   complex(c_double), intent(out) :: qs(*)
  end function
  integer(c_int) function c_fft_backward(f, op, qs, qp) bind(C, name="ps3d_fft_backward")
   import :: c_int, c_ptr, c_char, c_double
   type(c_ptr), value :: f
   character(kind=c_char), intent(in) :: op(*)
   complex(c_double), intent(in) :: qs(*)
   real(c_double), intent(out) :: qp(*)
  end function
  integer(c_int) function c_last_error(buf, len) bind(C, name="ps3d_last_error")
   import :: c_int, c_char
   character(kind=c_char), intent(out) :: buf(*)
   integer(c_int), value :: len
  end function
 end interface

contains

 subroutine check(ierr, where)
  integer(c_int), intent(in) :: ierr
  character(len=*), intent(in) :: where
  character(kind=c_char) :: buf(400)
  character(len=400) :: msg
  integer :: n, i
  if (ierr == 0) return
  n = c_last_error(buf, 400_c_int)
  msg = ' '
  do i = 1, min(n, 399)
   msg(i:i) = buf(i)
  enddo
  call halt_stop('p3dfft ('//where//'): '//trim(msg))
 end subroutine check

 ! p3dfft_setup(dims, nx, ny, nz, comm, nx_cut, ny_cut, nz_cut, overwrite,
 ! memsize) as called at my_fft.f90:10.  One rank: dims must be (1, 1); the
 ! cut sizes, overwrite and memsize have nothing to select on one device.
 subroutine p3dfft_setup(dims, nx, ny, nz, mpi_comm_in, nx_cut, ny_cut, nz_cut, overwrite, memsize)
  integer, intent(in) :: dims(2), nx, ny, nz, mpi_comm_in
  integer, intent(in), optional :: nx_cut, ny_cut, nz_cut
  logical, intent(in), optional :: overwrite
  integer, intent(out), optional :: memsize(3)
  if (dims(1) * dims(2) /= 1) &
   call halt_stop('p3dfft: the GPU transforms hold the whole domain on one rank (n_pes_x = n_pes_y = 1)')
  if (c_associated(plan)) call p3dfft_clean
  call check(c_fft_create(int(nx, c_int), int(ny, c_int), int(nz, c_int), c_null_ptr, plan), 'setup')
  gnx = nx; gny = ny; gnz = nz
  if (present(memsize)) memsize = 0
 end subroutine p3dfft_setup

 ! conf = 1: the real array (nx, ny, nz); conf = 2: the complex array
 ! (nx/2+1, ny, nz), both whole on one rank
 subroutine p3dfft_get_dims(dstart, dend, dsize, conf)
  integer, intent(out) :: dstart(3), dend(3), dsize(3)
  integer, intent(in) :: conf
  if (.not. c_associated(plan)) call halt_stop('p3dfft: p3dfft_get_dims before p3dfft_setup')
  dstart = (/1, 1, 1/)
  if (conf == 1) then
   dend = (/gnx, gny, gnz/)
  else if (conf == 2) then
   dend = (/gnx/2 + 1, gny, gnz/)
  else
   call halt_stop('p3dfft: p3dfft_get_dims supports conf = 1 and 2')
  endif
  dsize = dend
 end subroutine p3dfft_get_dims

 subroutine p3dfft_ftran_r2c(XgYZ, XYZg, op)
  real(mytype), intent(in) :: XgYZ(*)
  complex(mytype), intent(out) :: XYZg(*)
  character(len=*), intent(in) :: op
  if (.not. c_associated(plan)) call halt_stop('p3dfft: p3dfft_ftran_r2c before p3dfft_setup')
  call check(c_fft_forward(plan, op(1:min(3, len(op)))//c_null_char, XgYZ, XYZg), 'ftran_r2c '//op)
 end subroutine p3dfft_ftran_r2c

 subroutine p3dfft_btran_c2r(XYZg, XgYZ, op)
  complex(mytype), intent(in) :: XYZg(*)
  real(mytype), intent(out) :: XgYZ(*)
  character(len=*), intent(in) :: op
  if (.not. c_associated(plan)) call halt_stop('p3dfft: p3dfft_btran_c2r before p3dfft_setup')
  call check(c_fft_backward(plan, op(1:min(3, len(op)))//c_null_char, XYZg, XgYZ), 'btran_c2r '//op)
 end subroutine p3dfft_btran_c2r

 subroutine p3dfft_clean
  if (c_associated(plan)) call check(c_fft_destroy(plan), 'clean')
  plan = c_null_ptr
 end subroutine p3dfft_clean

end module p3dfft
