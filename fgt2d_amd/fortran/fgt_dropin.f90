! Fortran drop-in for the plane-wave FGT drivers of flatironinstitute/fgt2d.
!
! Same names, arguments and meaning as the reference subroutines, so a Fortran
! program linked against this library instead of libfgt2d runs the transform
! on the GPU:
!   fgt2d                 2d/planewave/src/fgt2d.f:19-22
!   fgt2dpart_direct_vec  2d/planewave/src/fgt2d.f:1354-1357
!   fgt3d                 3d/planewave/src/fgt3d.f:19-22
!   fgt3dpart_direct_vec  3d/planewave/src/fgt3d.f:1533-1536
!   prini, prin2, prinf, prinf_long   common/prini_new.f (the reference test
!                                      driver calls them for its printing)
! Each call goes through the C-ABI of include/fgtpw.h (iso_c_binding, all
! scalars by value).  The reference reports no errors; the signatures have no
! status argument, so a failure prints the library's message and stops.

module fgtpw_c
  use iso_c_binding
  implicit none
  interface
     integer(c_int) function fgtpw_fgt2d(nd, delta, eps, ns, sources, ifcharge, charge, &
          ifdipole, rnormal, dipstr, iper, ifpgh, pot, grad, hess, nt, targ, ifpghtarg, &
          pottarg, gradtarg, hesstarg) bind(C, name="fgtpw_fgt2d")
       import :: c_int, c_double
       integer(c_int), value :: nd, ns, ifcharge, ifdipole, iper, ifpgh, nt, ifpghtarg
       real(c_double), value :: delta, eps
       real(c_double) :: sources(*), charge(*), rnormal(*), dipstr(*)
       real(c_double) :: pot(*), grad(*), hess(*), targ(*), pottarg(*), gradtarg(*), hesstarg(*)
     end function fgtpw_fgt2d

     integer(c_int) function fgtpw_fgt2dpart_direct(nd, delta, eps, istart, iend, jstart, jend, &
          source, ifcharge, charge, ifdipole, rnormal, dipstr, targ, ifpgh, pot, grad, hess) &
          bind(C, name="fgtpw_fgt2dpart_direct")
       import :: c_int, c_double
       integer(c_int), value :: nd, istart, iend, jstart, jend, ifcharge, ifdipole, ifpgh
       real(c_double), value :: delta, eps
       real(c_double) :: source(*), charge(*), rnormal(*), dipstr(*), targ(*)
       real(c_double) :: pot(*), grad(*), hess(*)
     end function fgtpw_fgt2dpart_direct

     integer(c_int) function fgtpw_fgt3d(nd, delta, eps, ns, sources, ifcharge, charge, &
          ifdipole, rnormal, dipstr, iper, ifpgh, pot, grad, hess, nt, targ, ifpghtarg, &
          pottarg, gradtarg, hesstarg) bind(C, name="fgtpw_fgt3d")
       import :: c_int, c_double
       integer(c_int), value :: nd, ns, ifcharge, ifdipole, iper, ifpgh, nt, ifpghtarg
       real(c_double), value :: delta, eps
       real(c_double) :: sources(*), charge(*), rnormal(*), dipstr(*)
       real(c_double) :: pot(*), grad(*), hess(*), targ(*), pottarg(*), gradtarg(*), hesstarg(*)
     end function fgtpw_fgt3d

     integer(c_int) function fgtpw_fgt3dpart_direct(nd, delta, dmax, istart, iend, jstart, jend, &
          source, ifcharge, charge, ifdipole, rnormal, dipstr, targ, ifpgh, pot, grad, hess) &
          bind(C, name="fgtpw_fgt3dpart_direct")
       import :: c_int, c_double
       integer(c_int), value :: nd, istart, iend, jstart, jend, ifcharge, ifdipole, ifpgh
       real(c_double), value :: delta, dmax
       real(c_double) :: source(*), charge(*), rnormal(*), dipstr(*), targ(*)
       real(c_double) :: pot(*), grad(*), hess(*)
     end function fgtpw_fgt3dpart_direct

     type(c_ptr) function fgtpw_last_error() bind(C, name="fgtpw_last_error")
       import :: c_ptr
     end function fgtpw_last_error
  end interface

contains

  subroutine fgtpw_stop_on_error(ier, who)
    integer(c_int), intent(in) :: ier
    character(len=*), intent(in) :: who
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (ier == 0) return
    call c_f_pointer(fgtpw_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(*, '(a, a, i0, a, a)') who, ': error ', ier, ': ', trim(text)
    error stop 1
  end subroutine fgtpw_stop_on_error

end module fgtpw_c


subroutine fgt2d(nd, delta, eps, ns, sources, ifcharge, charge, ifdipole, rnormal, dipstr, &
     iper, ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg)
  use fgtpw_c
  implicit none
  integer :: nd, ns, ifcharge, ifdipole, iper, ifpgh, nt, ifpghtarg
  real(8) :: delta, eps
  real(8) :: sources(2, *), charge(nd, *), rnormal(2, *), dipstr(nd, *)
  real(8) :: pot(nd, *), grad(nd, 2, *), hess(nd, 3, *)
  real(8) :: targ(2, *), pottarg(nd, *), gradtarg(nd, 2, *), hesstarg(nd, 3, *)
  integer(c_int) :: ier
  ier = fgtpw_fgt2d(nd, delta, eps, ns, sources, ifcharge, charge, ifdipole, rnormal, dipstr, &
       iper, ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg)
  call fgtpw_stop_on_error(ier, 'fgt2d')
end subroutine fgt2d


subroutine fgt2dpart_direct_vec(nd, delta, eps, istart, iend, jstart, jend, source, ifcharge, &
     charge, ifdipole, rnormal, dipstr, targ, ifpgh, pot, grad, hess)
  use fgtpw_c
  implicit none
  integer :: nd, istart, iend, jstart, jend, ifcharge, ifdipole, ifpgh
  real(8) :: delta, eps
  real(8) :: source(2, *), charge(nd, *), rnormal(2, *), dipstr(nd, *), targ(2, *)
  real(8) :: pot(nd, *), grad(nd, 2, *), hess(nd, 3, *)
  integer(c_int) :: ier
  ier = fgtpw_fgt2dpart_direct(nd, delta, eps, istart, iend, jstart, jend, source, ifcharge, &
       charge, ifdipole, rnormal, dipstr, targ, ifpgh, pot, grad, hess)
  call fgtpw_stop_on_error(ier, 'fgt2dpart_direct_vec')
end subroutine fgt2dpart_direct_vec


subroutine fgt3d(nd, delta, eps, ns, sources, ifcharge, charge, ifdipole, rnormal, dipstr, &
     iper, ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg)
  use fgtpw_c
  implicit none
  integer :: nd, ns, ifcharge, ifdipole, iper, ifpgh, nt, ifpghtarg
  real(8) :: delta, eps
  real(8) :: sources(3, *), charge(nd, *), rnormal(3, *), dipstr(nd, *)
  real(8) :: pot(nd, *), grad(nd, 3, *), hess(nd, 6, *)
  real(8) :: targ(3, *), pottarg(nd, *), gradtarg(nd, 3, *), hesstarg(nd, 6, *)
  integer(c_int) :: ier
  ier = fgtpw_fgt3d(nd, delta, eps, ns, sources, ifcharge, charge, ifdipole, rnormal, dipstr, &
       iper, ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg)
  call fgtpw_stop_on_error(ier, 'fgt3d')
end subroutine fgt3d


subroutine fgt3dpart_direct_vec(nd, delta, dmax, istart, iend, jstart, jend, source, ifcharge, &
     charge, ifdipole, rnormal, dipstr, targ, ifpgh, pot, grad, hess)
  use fgtpw_c
  implicit none
  integer :: nd, istart, iend, jstart, jend, ifcharge, ifdipole, ifpgh
  real(8) :: delta, dmax
  real(8) :: source(3, *), charge(nd, *), rnormal(3, *), dipstr(nd, *), targ(3, *)
  real(8) :: pot(nd, *), grad(nd, 3, *), hess(nd, 6, *)
  integer(c_int) :: ier
  ier = fgtpw_fgt3dpart_direct(nd, delta, dmax, istart, iend, jstart, jend, source, ifcharge, &
       charge, ifdipole, rnormal, dipstr, targ, ifpgh, pot, grad, hess)
  call fgtpw_stop_on_error(ier, 'fgt3dpart_direct_vec')
end subroutine fgt3dpart_direct_vec


! Printing helpers with the calling sequences of the reference's prini family
! (common/prini_new.f): prini selects the output unit, prin2 prints reals,
! prinf integers; a message ends at its first '*'.
module fgtpw_print
  implicit none
  integer, save :: unit_out = 6
  logical, save :: printing = .true.
contains
  function upto_star(mes) result(s)
    character(len=*), intent(in) :: mes
    character(len=len(mes)) :: s
    integer :: k
    k = index(mes, '*')
    if (k == 0) then
       s = mes
    else
       s = mes(1:k - 1)
    end if
  end function upto_star
end module fgtpw_print

subroutine prini(ip1, iq1)
  use fgtpw_print
  implicit none
  integer :: ip1, iq1
  unit_out = ip1
  printing = ip1 > 0
end subroutine prini

subroutine prin2(mes, a, n)
  use fgtpw_print
  implicit none
  character(len=*) :: mes
  integer :: n
  real(8) :: a(*)
  if (.not. printing) return
  write(unit_out, '(1x, a)') trim(upto_star(mes))
  if (n > 0) write(unit_out, '(6(2x, es13.6))') a(1:n)
end subroutine prin2

subroutine prinf(mes, ia, n)
  use fgtpw_print
  implicit none
  character(len=*) :: mes
  integer :: n
  integer :: ia(*)
  if (.not. printing) return
  write(unit_out, '(1x, a)') trim(upto_star(mes))
  if (n > 0) write(unit_out, '(10(1x, i7))') ia(1:n)
end subroutine prinf

subroutine prinf_long(mes, ia, n)
  use fgtpw_print
  implicit none
  character(len=*) :: mes
  integer :: n
  integer :: ia(*)
  if (.not. printing) return
  write(unit_out, '(1x, a)') trim(upto_star(mes))
  if (n > 0) write(unit_out, '(6(1x, i12))') ia(1:n)
end subroutine prinf_long
