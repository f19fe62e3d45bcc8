! nf_stub.f90 -- TEST INFRASTRUCTURE ONLY.  No-op netCDF routines for the
! link of the reference driver (main.f90 and its diagnostics, compiled
! unchanged) with the GPU drop-in, in an image without netCDF-Fortran.
! They write no netCDF file.  For the tests they record, under the folder
! named by PS3D_DRIVER_CAPTURE (nothing when unset):
!  * params.txt, at the first file creation (init_snap_cdf / init_diag_*,
!    after set_parameter and wavenumber_initialisation): main_module's
!    parameters, and kx, ky, kz, ksqr as raw doubles (kx.bin ...);
!  * every nf_put_vara_double of a 3-D variable as raw doubles, one file per
!    (netCDF file, variable, time index), e.g. snap.cdf.u.0001.
! With PS3D_DRIVER_SNAPSHOTS = K > 0, closing snap.cdf after its K-th
! snapshot (diag_snap, main.f90:147) sets runlen = time, so main.f90's own
! loop ends after that step's diagnostics and writes its restart file.
module nf_stub_state
 implicit none
 integer, parameter :: maxf = 64, maxv = 512
 character(len=128) :: fname(maxf) = ' '
 logical :: opened_for_write(maxf) = .false.
 integer :: writes(maxf) = 0        ! completed write sessions per file
 integer :: nfiles = 0
 character(len=64) :: vname(maxv) = ' '
 integer :: vndims(maxv) = 0
 integer :: nvars = 0
 character(len=512) :: capture = ' '
 integer :: stop_after = 0, snapshots = 0
 logical :: ready = .false., params_written = .false.
contains
 subroutine setup
  character(len=32) :: s
  integer :: l, st
  if (ready) return
  ready = .true.
  call get_environment_variable('PS3D_DRIVER_CAPTURE', capture, l, st)
  if (st /= 0) capture = ' '
  call get_environment_variable('PS3D_DRIVER_SNAPSHOTS', s, l, st)
  if (st == 0 .and. l > 0) read(s, *) stop_after
 end subroutine setup

 integer function file_id(name)
  character(len=*), intent(in) :: name
  integer :: i
  do i = 1, nfiles
   if (trim(fname(i)) == trim(name)) then
    file_id = i
    return
   endif
  enddo
  nfiles = nfiles + 1
  fname(nfiles) = name
  file_id = nfiles
 end function file_id

 integer function var_id(name, ndims)
  character(len=*), intent(in) :: name
  integer, intent(in) :: ndims
  integer :: i
  do i = 1, nvars
   if (trim(vname(i)) == trim(name)) then
    var_id = i
    if (ndims >= 0) vndims(i) = ndims
    return
   endif
  enddo
  nvars = nvars + 1
  vname(nvars) = name
  vndims(nvars) = ndims
  var_id = nvars
 end function var_id

 subroutine write_raw(path, a, n)
  character(len=*), intent(in) :: path
  integer, intent(in) :: n
  real(8), intent(in) :: a(n)
  integer :: u
  open(newunit=u, file=path, access='stream', form='unformatted', status='replace')
  write(u) a
  close(u)
 end subroutine write_raw

 subroutine write_params
  use main_module
  integer :: io
  if (params_written .or. len_trim(capture) == 0) return
  params_written = .true.
  open(newunit=io, file=trim(capture)//'/params.txt', status='replace')
  write(io, '(a,3i8)') 'grid', nx, ny, nz
  write(io, '(a,3l2)') 'switches', enable_nonlinear, enable_vertical_boundaries, enable_AB_3_order
  write(io, '(a,20es26.17e3)') 'reals', Lx, Ly, Lz, dx, dy, dz, dt, f0, N0, Ro, dsqr, eps_ab, &
                              Ah, Av, Kh, Kv, Ahbi, Avbi, Khbi, Kvbi
  write(io, '(a,3es26.17e3)') 'cadence', tsmonint, snapint, runlen
  close(io)
  call write_raw(trim(capture)//'/kx.bin', kx, size(kx))
  call write_raw(trim(capture)//'/ky.bin', ky, size(ky))
  call write_raw(trim(capture)//'/kz.bin', kz, size(kz))
  call write_raw(trim(capture)//'/ksqr.bin', ksqr, size(ksqr))
 end subroutine write_params
end module nf_stub_state


integer function nf_create(path, cmode, ncid)
 use nf_stub_state
 character(len=*) :: path
 integer :: cmode, ncid
 call setup
 call write_params
 ncid = file_id(path)
 opened_for_write(ncid) = .false.
 nf_create = 0
end function

integer function nf_create_par(path, cmode, comm, info, ncid)
 use nf_stub_state
 character(len=*) :: path
 integer :: cmode, comm, info, ncid
 call setup
 call write_params
 ncid = file_id(path)
 nf_create_par = 0
end function

integer function nf_open(path, mode, ncid)
 use nf_stub_state
 character(len=*) :: path
 integer :: mode, ncid
 call setup
 ncid = file_id(path)
 opened_for_write(ncid) = .true.
 nf_open = 0
end function

integer function nf_open_par(path, mode, comm, info, ncid)
 use nf_stub_state
 character(len=*) :: path
 integer :: mode, comm, info, ncid
 call setup
 ncid = file_id(path)
 opened_for_write(ncid) = .true.
 nf_open_par = 0
end function

integer function nf_close(ncid)
 use nf_stub_state
 use main_module, only: runlen, time
 integer :: ncid
 if (opened_for_write(ncid)) then
  writes(ncid) = writes(ncid) + 1
  if (trim(fname(ncid)) == 'snap.cdf') then
   snapshots = writes(ncid)
   if (stop_after > 0 .and. snapshots >= stop_after) runlen = time
  endif
 endif
 opened_for_write(ncid) = .false.
 nf_close = 0
end function

integer function nf_def_dim(ncid, name, len, dimid)
 character(len=*) :: name
 integer :: ncid, len, dimid
 dimid = 1
 nf_def_dim = 0
end function

integer function ncddef(ncid, name, len, iret)
 character(len=*) :: name
 integer :: ncid, len, iret
 iret = 0
 ncddef = 1
end function

integer function nf_def_var(ncid, name, xtype, ndims, dimids, varid)
 use nf_stub_state
 character(len=*) :: name
 integer :: ncid, xtype, ndims, dimids(*), varid
 varid = var_id(name, ndims)
 nf_def_var = 0
end function

integer function ncvdef(ncid, name, xtype, ndims, dimids, iret)
 use nf_stub_state
 character(len=*) :: name
 integer :: ncid, xtype, ndims, dimids(*), iret
 iret = 0
 ncvdef = var_id(name, ndims)
end function

integer function nf_enddef(ncid)
 integer :: ncid
 nf_enddef = 0
end function

integer function nf_set_fill(ncid, mode, old)
 integer :: ncid, mode, old
 old = 0
 nf_set_fill = 0
end function

integer function nf_var_par_access(ncid, varid, mode)
 integer :: ncid, varid, mode
 nf_var_par_access = 0
end function

integer function nf_inq_dimid(ncid, name, dimid)
 character(len=*) :: name
 integer :: ncid, dimid
 dimid = 1
 nf_inq_dimid = 0
end function

! the time dimension's length: records written to this file so far
integer function nf_inq_dimlen(ncid, dimid, len)
 use nf_stub_state
 integer :: ncid, dimid, len
 len = writes(ncid)
 nf_inq_dimlen = 0
end function

integer function nf_inq_varid(ncid, name, varid)
 use nf_stub_state
 character(len=*) :: name
 integer :: ncid, varid
 varid = var_id(name, -1)
 nf_inq_varid = 0
end function

integer function nf_put_att_text(ncid, varid, name, len, text)
 character(len=*) :: name, text
 integer :: ncid, varid, len
 nf_put_att_text = 0
end function

integer function nf_put_att_double(ncid, varid, name, xtype, len, vals)
 character(len=*) :: name
 integer :: ncid, varid, xtype, len
 real(8) :: vals(*)
 nf_put_att_double = 0
end function

integer function nf_put_att_int(ncid, varid, name, xtype, len, vals)
 character(len=*) :: name
 integer :: ncid, varid, xtype, len, vals(*)
 nf_put_att_int = 0
end function

! records 3-D (or 3-D + time) variables: file.var.NNNN, NNNN the time index
integer function nf_put_vara_double(ncid, varid, start, count, vals)
 use nf_stub_state
 integer :: ncid, varid, start(*), count(*)
 real(8) :: vals(*)
 integer :: nd, n, i
 character(len=600) :: path
 nf_put_vara_double = 0
 if (len_trim(capture) == 0) return
 nd = vndims(varid)
 if (nd < 3) return
 n = 1
 do i = 1, 3
  n = n * count(i)
 enddo
 write(path, '(a,"/",a,".",a,".",i4.4)') trim(capture), trim(fname(ncid)), trim(vname(varid)), &
      merge(start(nd), 0, nd > 3)
 call write_raw(trim(path), vals, n)
end function

character(len=80) function nf_strerror(iret)
 integer :: iret
 nf_strerror = 'netCDF stand-in: no error'
end function
