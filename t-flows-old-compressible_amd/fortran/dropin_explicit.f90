!==============================================================================!
!  explicit_cgs -- the CGS call of dropin_driver through the typed
!  ISO_C_BINDING interface of module tflows_krylov (bound to cgs_).
!==============================================================================!
subroutine explicit_cgs(n, nb, nonz, a, acol, arow, adia, ab, x, r1, prec, niter, tol, inires, finres)
  use tflows_krylov, only: tflows_cgs
  implicit none
  integer :: n, nb, nonz, prec, niter
  real    :: a(nonz), ab(-nb:-1), x(-nb:n), r1(n), tol, inires, finres
  integer :: acol(n + 1), arow(nonz), adia(n)
  call tflows_cgs(n, nb, nonz, a, acol, arow, adia, ab, x, r1, prec, niter, tol, inires, finres)
end subroutine explicit_cgs
