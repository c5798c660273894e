! C entry points around the reference's bz_integration module (test
! infrastructure: oracle/Makefile links this with the reference's own
! Utilities/pyphutil/fortran_routines/bz_integration.f90 into
! oracle/_ref/libbzi_ref.so).  Arrays are explicit-shape so a C caller can
! pass plain pointers; rank 1 keeps the reference's progress bar silent.
module bzi_ref_wrap
  use iso_c_binding
  use bz_integration
  implicit none
contains
  ! tri: ntri x 3 column-major, 1-based; freq, func: nk x nbands column-major
  subroutine ref_linear_triangulation(nk, nbands, ntri, tri, ne, e, freq, func, dos, nd) &
      bind(C, name="ref_linear_triangulation")
    integer(c_int), value :: nk, nbands, ntri, ne
    integer(c_int), intent(in) :: tri(ntri, 3)
    real(c_double), intent(in) :: e(ne), freq(nk, nbands), func(nk, nbands)
    real(c_double), intent(out) :: dos(ne), nd(ne)
    integer, allocatable :: t(:, :)
    real(c_double), allocatable :: fr(:, :), fu(:, :)
    allocate(t(ntri, 3), fr(nk, nbands), fu(nk, nbands))
    t = tri
    fr = freq
    fu = func
    dos = 0.0d0
    nd = 0.0d0
    call linear_triangulation(nbands, t, ne, e, fr, fu, 1, dos, nd)
  end subroutine

  subroutine ref_standard_integration(method, weight, nbands, freq, width, ne, e, natom, dos) &
      bind(C, name="ref_standard_integration")
    integer(c_int), value :: method, nbands, ne, natom
    real(c_double), value :: weight, width
    real(c_double), intent(in) :: freq(nbands), e(ne)
    real(c_double), intent(out) :: dos(ne)
    call standard_integration(method, weight, freq, width, e, ne, natom, nbands, dos)
  end subroutine
end module
