!==============================================================================!
!  dropin_driver -- a Fortran host that calls the GPU solvers exactly the way
!  the reference's equation assemblers call its own:
!
!    call cg  (NC, NbC, NONZERO, Aval, Acol, Arow, Adia, Abou, PP%n, b,
!              PREC, niter, PP%STol, res(4), error)        ! CalcPS.f90:217
!    call bicg(NC, Nbc, NONZERO, Aval, Acol, Arow, Adia, Abou, rhoUi, b,
!              PREC, niter, Ui%STol, res(var), error)      ! NewUVW.f90:506
!
!  with implicit interfaces, x dimensioned (-NbC:NC), REAL = REAL(8).  The
!  executable links libtflows_krylov.so in place of the reference's CG.o,
!  BiCG.o, CGS.o; nothing in the calling code changes.  CGS is called
!  through the explicit ISO_C_BINDING interface of module tflows_krylov
!  (dropin_explicit.f90).
!
!  The matrix is the pressure-correction Laplacian of an nx*ny*nz box of
!  unit cells (off-diagonal -1 per interior face, diagonal = number of
!  interior faces, cell 1 pinned by +1 on its diagonal), built here in the
!  T-FlowS layout that TopolM produces (sorted columns, NONZ = nnz + 1).
!  Prints one line per solve: method, niter, IniRes, FinRes, true ||b-Ax||.
!==============================================================================!
program dropin_driver
  implicit none
  integer, parameter :: nx = 16, ny = 12, nz = 10
  integer :: nc, nbc, nonzero, prec, niter, c, j, n, cnt
  integer, allocatable :: acol(:), arow(:), adia(:)
  real, allocatable :: aval(:), abou(:), x(:), b(:), bsave(:)
  real :: tol, res, error, bnorm, tr
  integer :: nb_list(6), cols(7), m
  logical :: inserted

  nc = nx * ny * nz
  nbc = 3                           ! buffer cells of an MPI build; unused here
  allocate(acol(nc + 1), adia(nc))
  ! count entries: the diagonal plus one per interior face
  acol(1) = 1
  do c = 1, nc
    call neighbours(c, nb_list, cnt)
    acol(c + 1) = acol(c) + cnt + 1
  end do
  nonzero = acol(nc + 1)
  allocate(arow(nonzero), aval(nonzero), abou(-nbc:-1))
  arow = 0; aval = 0.0; abou = 0.0
  do c = 1, nc
    call neighbours(c, nb_list, cnt)
    ! columns of row c, ascending: the neighbours with c itself in its place
    m = 0
    inserted = .false.
    do j = 1, cnt
      if (.not. inserted .and. nb_list(j) > c) then
        m = m + 1; cols(m) = c; inserted = .true.
      end if
      m = m + 1; cols(m) = nb_list(j)
    end do
    if (.not. inserted) then
      m = m + 1; cols(m) = c
    end if
    n = acol(c)
    do j = 1, m
      arow(n) = cols(j)
      if (cols(j) == c) then
        aval(n) = real(cnt); adia(c) = n
      else
        aval(n) = -1.0
      end if
      n = n + 1
    end do
  end do
  aval(adia(1)) = aval(adia(1)) + 1.0   ! pin cell 1

  allocate(x(-nbc:nc), b(nc), bsave(nc))
  do c = 1, nc
    bsave(c) = sin(0.37 * c) + 0.25 * cos(1.3 * c)
  end do
  bsave = bsave - sum(bsave) / nc
  bnorm = sqrt(sum(bsave**2))
  tol = 1.0e-8 * bnorm
  prec = 1

  ! --- CG through the implicit interface, as CalcPS.f90:217 ---------------
  x = 0.0; b = bsave; niter = 800
  call cg(nc, nbc, nonzero, aval, acol, arow, adia, abou, x, b, prec, niter, tol, res, error)
  call report('cg', niter, res, error)

  ! --- BiCG through the implicit interface, as NewUVW.f90:506 -------------
  x = 0.0; b = bsave; niter = 200
  call bicg(nc, nbc, nonzero, aval, acol, arow, adia, abou, x, b, prec, niter, tol, res, error)
  call report('bicg', niter, res, error)

  ! --- CGS through the explicit ISO_C_BINDING interface (dropin_explicit.f90;
  !     flang keeps a BIND(C) "cgs_" and an implicit cgs apart per unit) ----
  x = 0.0; b = bsave; niter = 400
  call explicit_cgs(nc, nbc, nonzero, aval, acol, arow, adia, abou, x, b, prec, niter, tol, res, error)
  call report('cgs', niter, res, error)

contains

  subroutine neighbours(c, list, cnt)
    integer, intent(in) :: c
    integer, intent(out) :: list(6), cnt
    integer :: i0, j0, k0
    i0 = mod(c - 1, nx); j0 = mod((c - 1) / nx, ny); k0 = (c - 1) / (nx * ny)
    cnt = 0
    if (k0 > 0)      then; cnt = cnt + 1; list(cnt) = c - nx * ny; end if
    if (j0 > 0)      then; cnt = cnt + 1; list(cnt) = c - nx;      end if
    if (i0 > 0)      then; cnt = cnt + 1; list(cnt) = c - 1;       end if
    if (i0 < nx - 1) then; cnt = cnt + 1; list(cnt) = c + 1;       end if
    if (j0 < ny - 1) then; cnt = cnt + 1; list(cnt) = c + nx;      end if
    if (k0 < nz - 1) then; cnt = cnt + 1; list(cnt) = c + nx * ny; end if
  end subroutine neighbours

  subroutine report(name, it, ini, fin)
    character(*), intent(in) :: name
    integer, intent(in) :: it
    real, intent(in) :: ini, fin
    real :: r(nc)
    integer :: ii, jj
    r = bsave
    do ii = 1, nc
      do jj = acol(ii), acol(ii + 1) - 1
        r(ii) = r(ii) - aval(jj) * x(arow(jj))
      end do
    end do
    tr = sqrt(sum(r**2))
    write(*, '(a,1x,a,1x,i6,3(1x,es14.6),1x,es14.6)') 'DROPIN', name, it, ini, fin, tr, tol
  end subroutine report

end program dropin_driver
