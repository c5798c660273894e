! parphom_hip.f90 - ISO_C_BINDING interface of libparphom_hip.so and
! libparphom_h5.so for a Fortran host (the module a parphom.x maintainer adds;
! the C declarations are include/parphom_hip.h and include/parphom_h5.h).
!
! The derived types are interoperable with the C structs field by field, so a
! Fortran host passes its own arrays (lat(3,3), real_pos(N,3), q points(nq,3),
! the fc(3,3,N,N) dataset) by address, unchanged.
module parphom_hip
  use iso_c_binding
  implicit none
  private
  public :: pph_system, pph_opts, pph_out
  public :: pph_create, pph_run, pph_destroy, pph_last_error, pph_version
  public :: pph_h5_read_fc, pph_h5_open_output, pph_h5_write_group, pph_h5_close, &
            pph_h5_last_error
  public :: c_to_string, PPH_OK

  integer(c_int), parameter :: PPH_OK = 0

  ! the fields of parphom's `moire` state (mod_variables.F90:83-89)
  type, bind(C) :: pph_system
    integer(c_int64_t) :: natom, ntypes
    type(c_ptr)        :: lat      ! real(c_double) lat(3,3): lat(i,k) = component k of vector i
    type(c_ptr)        :: pos      ! real(c_double) real_pos(N,3), Cartesian (A)
    type(c_ptr)        :: mass     ! real(c_double) mass(ntypes), amu
    type(c_ptr)        :: type     ! integer(c_int64_t) type(N), 1-based
  end type

  ! pzheevx_vars and the velocity settings (read_input.F90:368-516)
  type, bind(C) :: pph_opts
    character(kind=c_char) :: jobz, range, vel_method, precision
    integer(c_int32_t)     :: vel, device
    real(c_double)         :: vl, vu          ! cm^-1 as in the input file
    integer(c_int64_t)     :: il, iu
    real(c_double)         :: abstol, orfac
    integer(c_int64_t)     :: batch           ! 0 = automatic
    integer(c_int32_t)     :: streams         ! batches in flight, 0 = automatic
  end type

  ! outputs of pph_run (host pointers)
  type, bind(C) :: pph_out
    type(c_ptr)        :: eval_cm   ! real(c_double) eval(ld_eval, nq)
    integer(c_int64_t) :: ld_eval
    type(c_ptr)        :: m         ! integer(c_int64_t) m(nq)
    type(c_ptr)        :: nz        ! integer(c_int64_t) nz(nq)
    type(c_ptr)        :: evec      ! complex(c_double_complex) evec(3N, evec_cols, nq) or null
    integer(c_int64_t) :: evec_cols
    type(c_ptr)        :: vel       ! real(c_double) vel(3, evec_cols, nq) or null
    type(c_ptr)        :: info      ! integer(c_int32_t) info(nq)
  end type

  interface
    ! read_force_constants + pdgeadd and the lattice set-up (force_read.F90:44-213)
    integer(c_int) function pph_create(sys, fc, opts, h) bind(C, name="pph_create")
      import :: c_int, c_ptr, c_double, pph_system, pph_opts
      type(pph_system), intent(in) :: sys
      real(c_double), intent(in)   :: fc(*)
      type(pph_opts), intent(in)   :: opts
      type(c_ptr), intent(out)     :: h
    end function
    ! the q loop of calculations.F90:49-53
    integer(c_int) function pph_run(h, nq, q, ldq, out) bind(C, name="pph_run")
      import :: c_int, c_ptr, c_int64_t, c_double, pph_out
      type(c_ptr), value           :: h
      integer(c_int64_t), value    :: nq
      real(c_double), intent(in)   :: q(*)
      integer(c_int64_t), value    :: ldq
      type(pph_out), intent(inout) :: out
    end function
    integer(c_int) function pph_destroy(h) bind(C, name="pph_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function pph_last_error(h) bind(C, name="pph_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function pph_version() bind(C, name="pph_version")
      import :: c_ptr
    end function

    ! HDF5 in parphom.x's layout (libparphom_h5.so)
    integer(c_int) function pph_h5_read_fc(path, dset, natom, out) bind(C, name="pph_h5_read_fc")
      import :: c_int, c_char, c_int64_t, c_double
      character(kind=c_char), intent(in) :: path(*), dset(*)
      integer(c_int64_t), value          :: natom
      real(c_double), intent(out)        :: out(*)
    end function
    integer(c_int) function pph_h5_open_output(path, create, fid) &
        bind(C, name="pph_h5_open_output")
      import :: c_int, c_char, c_int32_t, c_int64_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value          :: create
      integer(c_int64_t), intent(out)    :: fid
    end function
    ! the per-q group of write_output.F90:95-464: q_vec, eigenvalues,
    ! evec {r,i} (nz, n), vel (nz, 3); evec and vel may be c_null_ptr
    integer(c_int) function pph_h5_write_group(fid, qindex, qvec, eval_cm, m, evec, n, nz, vel) &
        bind(C, name="pph_h5_write_group")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int64_t), value  :: fid, qindex
      real(c_double), intent(in) :: qvec(3), eval_cm(*)
      integer(c_int64_t), value  :: m
      type(c_ptr), value         :: evec
      integer(c_int64_t), value  :: n, nz
      type(c_ptr), value         :: vel
    end function
    integer(c_int) function pph_h5_close(fid) bind(C, name="pph_h5_close")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: fid
    end function
    type(c_ptr) function pph_h5_last_error() bind(C, name="pph_h5_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! a NUL-terminated C string as a Fortran string
  function c_to_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: a(:)
    integer :: k
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    call c_f_pointer(p, a, [4096])
    k = 0
    do while (k < 4096)
      if (a(k + 1) == c_null_char) exit
      k = k + 1
    end do
    allocate(character(len=k) :: s)
    s = transfer(a(1:k), s)
  end function

end module parphom_hip
