!  ref_glue.f90 -- set-up glue for the reference Fortran solvers.
!
!  TEST INFRASTRUCTURE ONLY: linked into oracle/_ref/libtflows_ref.so next
!  to the reference's own Prec1/Resid/CG/BiCG/CGS sources (compiled where
!  they lie) so that tests can call the reference as a checker.
!
!  The reference allocates its solver workspace in GeoAloc
!  (Process/GeoAloc.f90:40-49) and sets the single-domain partition state in
!  StaPar (Sequential/StaPar.f90:16-17).  This glue does the same two jobs
!  for a stand-alone caller; it holds no solver arithmetic.
subroutine ref_setup(n, nb)
  use sol_mod
  use par_mod
  implicit none
  integer :: n, nb
  this = 0
  npro = 0
  if (allocated(d)) deallocate(d, p1, p2, q1, q2, r2, u1, u2, v1, v2)
  allocate(d(n)); d = 0
  allocate(p1(-nb:n)); p1 = 0
  allocate(p2(-nb:n)); p2 = 0
  allocate(q1(n)); q1 = 0
  allocate(q2(n)); q2 = 0
  allocate(r2(n)); r2 = 0
  allocate(u1(n)); u1 = 0
  allocate(u2(n)); u2 = 0
  allocate(v1(n)); v1 = 0
  allocate(v2(n)); v2 = 0
end subroutine ref_setup

!  Value of sol_mod's persistent `error` (Library/sol_mod.f90:11), which the
!  solvers hand back as FinRes on their early exit.
subroutine ref_get_error(e)
  use sol_mod
  implicit none
  real :: e
  e = error
end subroutine ref_get_error

subroutine ref_set_error(e)
  use sol_mod
  implicit none
  real :: e
  error = e
end subroutine ref_set_error
