!==============================================================================!
!  tflows_krylov -- ISO_C_BINDING interfaces of libtflows_krylov.so, the
!  MI355X (gfx950) Krylov solvers.
!
!  The drop-in entry points keep the reference's external subroutines
!  (Process/CG.f90:2-3, BiCG.f90:2-3, CGS.f90:2-3, Prec1.f90:2,
!  Resid.f90:2): every argument by reference, the same order and meaning,
!  linked under the names cg_, bicg_, cgs_, prec1_, resid_.  A caller that
!  keeps its implicit interface (as CalcPS.f90:217 and CalcZmix.f90:394 do)
!  links to them unchanged; a caller that USEs this module gets explicit,
!  typed interfaces to the same symbols.  Compile the calling code with
!  -fdefault-real-8 as the reference does, so that REAL is C double and
!  INTEGER is a 32-bit C int.
!
!  The device-resident API (tflows_system_*, tflows_solve*) is declared
!  here as well for callers that keep their data in GPU memory.
!==============================================================================!
module tflows_krylov
  use iso_c_binding, only: c_int32_t, c_double, c_ptr, c_char
  implicit none

  integer(c_int32_t), parameter :: TFLOWS_METHOD_CG = 0, TFLOWS_METHOD_BICG = 1, TFLOWS_METHOD_CGS = 2
  integer(c_int32_t), parameter :: TFLOWS_MEM_HOST = 0, TFLOWS_MEM_DEVICE = 1

  ! flags and constants of one momentum / mixture-fraction assembly
  ! (tflows_uvw_params); codes as Process/IniPar.f90:22-32, per-material
  ! arrays for materials 1..8
  type, bind(C) :: tflows_uvw_params
    integer(c_int32_t) :: var, convec, diffus, cross, inert
    real(c_double)     :: visc, dt, urf
    integer(c_int32_t) :: blend(8), blend_tem(8)
    real(c_double)     :: urfc(8)
    integer(c_int32_t) :: visc_rho = 0   ! 1: VISeff = visc * RhoS (Moin_Problem() == 4); 0: visc
  end type tflows_uvw_params

  interface
    !--- drop-ins (host arrays, reference layout) -----------------------------
    subroutine tflows_cg(n, nb, nonz, a, acol, arow, adia, ab, x, r1, prec, niter, tol, inires, finres) &
        bind(C, name="cg_")
      import :: c_int32_t, c_double
      integer(c_int32_t), intent(in)    :: n, nb, nonz, prec
      real(c_double),     intent(in)    :: a(nonz), ab(*), tol
      integer(c_int32_t), intent(in)    :: acol(n+1), arow(nonz), adia(n)
      real(c_double),     intent(inout) :: x(-nb:n), r1(n)
      integer(c_int32_t), intent(inout) :: niter
      real(c_double),     intent(inout) :: inires, finres
    end subroutine tflows_cg

    subroutine tflows_bicg(n, nb, nonz, a, acol, arow, adia, ab, x, r1, prec, niter, tol, inires, finres) &
        bind(C, name="bicg_")
      import :: c_int32_t, c_double
      integer(c_int32_t), intent(in)    :: n, nb, nonz, prec
      real(c_double),     intent(in)    :: a(nonz), ab(*), tol
      integer(c_int32_t), intent(in)    :: acol(n+1), arow(nonz), adia(n)
      real(c_double),     intent(inout) :: x(-nb:n), r1(n)
      integer(c_int32_t), intent(inout) :: niter
      real(c_double),     intent(inout) :: inires, finres
    end subroutine tflows_bicg

    subroutine tflows_cgs(n, nb, nonz, a, acol, arow, adia, ab, x, r1, prec, niter, tol, inires, finres) &
        bind(C, name="cgs_")
      import :: c_int32_t, c_double
      integer(c_int32_t), intent(in)    :: n, nb, nonz, prec
      real(c_double),     intent(in)    :: a(nonz), ab(*), tol
      integer(c_int32_t), intent(in)    :: acol(n+1), arow(nonz), adia(n)
      real(c_double),     intent(inout) :: x(-nb:n), r1(n)
      integer(c_int32_t), intent(inout) :: niter
      real(c_double),     intent(inout) :: inires, finres
    end subroutine tflows_cgs

    subroutine tflows_prec1(n, nonz, a, acol, arow, adia, d, prec) bind(C, name="prec1_")
      import :: c_int32_t, c_double
      integer(c_int32_t), intent(in)  :: n, nonz, prec
      real(c_double),     intent(in)  :: a(nonz)
      integer(c_int32_t), intent(in)  :: acol(n+1), arow(nonz), adia(n)
      real(c_double),     intent(out) :: d(n)
    end subroutine tflows_prec1

    subroutine tflows_resid(n, nb, nonz, a, acol, arow, ab, x, r1) bind(C, name="resid_")
      import :: c_int32_t, c_double
      integer(c_int32_t), intent(in)    :: n, nb, nonz
      real(c_double),     intent(in)    :: a(nonz), ab(*), x(-nb:n)
      integer(c_int32_t), intent(in)    :: acol(n+1), arow(nonz)
      real(c_double),     intent(inout) :: r1(n)
    end subroutine tflows_resid

    !--- device-resident API -------------------------------------------------
    function tflows_system_create(n, nonz, acol, arow, adia, mem, stream, sys) result(rc) &
        bind(C, name="tflows_system_create")
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: n, nonz, mem
      type(c_ptr), value        :: acol, arow, adia, stream
      type(c_ptr)               :: sys
      integer(c_int32_t)        :: rc
    end function tflows_system_create

    function tflows_system_destroy(sys) result(rc) bind(C, name="tflows_system_destroy")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: sys
      integer(c_int32_t) :: rc
    end function tflows_system_destroy

    function tflows_system_set_values(sys, aval, mem, stream) result(rc) bind(C, name="tflows_system_set_values")
      import :: c_int32_t, c_ptr
      type(c_ptr), value        :: sys, aval, stream
      integer(c_int32_t), value :: mem
      integer(c_int32_t)        :: rc
    end function tflows_system_set_values

    function tflows_solve(sys, method, prec, x, r1, niter, tol, inires, finres, stream) result(rc) &
        bind(C, name="tflows_solve")
      import :: c_int32_t, c_double, c_ptr
      type(c_ptr), value        :: sys, x, r1, stream
      integer(c_int32_t), value :: method, prec
      integer(c_int32_t)        :: niter
      real(c_double), value     :: tol
      real(c_double)            :: inires, finres
      integer(c_int32_t)        :: rc
    end function tflows_solve

    !--- pressure-correction assembly (CalcPS.f90:51-202) ------------------
    ! Array arguments by address (c_loc of the module arrays, which keep
    ! their reference bases: SideC(0:2,NS), TypeBC(-NbC:NC), volume(-NbC:NC))
    function tflows_calcps_create(nc, nbc, ns, nonz, sidec, sidaij, adia, typebc, sx, sy, sz, dx, dy, dz, f, &
                                  volume, mem, stream, h) result(rc) bind(C, name="tflows_calcps_create")
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: nc, nbc, ns, nonz, mem
      type(c_ptr), value        :: sidec, sidaij, adia, typebc, sx, sy, sz, dx, dy, dz, f, volume, stream
      type(c_ptr)               :: h
      integer(c_int32_t)        :: rc
    end function tflows_calcps_create

    function tflows_calcps_destroy(h) result(rc) bind(C, name="tflows_calcps_destroy")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t) :: rc
    end function tflows_calcps_destroy

    function tflows_calcps_assemble(h, nmat, dt, asave, rho_n, rho_o, rho_oo, u, v, w, p, px, py, pz, aval, b, &
                                    flux, abou, errmax, mean, mem, stream) result(rc) &
        bind(C, name="tflows_calcps_assemble")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value        :: h, asave, rho_n, rho_o, rho_oo, u, v, w, p, px, py, pz, aval, b, flux, abou, stream
      integer(c_int32_t), value :: nmat, mem
      real(c_double), value     :: dt
      real(c_double)            :: errmax, mean
      integer(c_int32_t)        :: rc
    end function tflows_calcps_assemble

    !--- the correction after the pressure solve (CorUVW.f90:36-121) -------
    function tflows_calcps_correct(h, dt, pp, aval, abou, asave, rho_n, rho_o, rho_oo, px, py, pz, u, v, w, &
                                   flux, b, errmax, mem, stream) result(rc) bind(C, name="tflows_calcps_correct")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value        :: h, pp, aval, abou, asave, rho_n, rho_o, rho_oo, px, py, pz, u, v, w, flux, b, &
                                   stream
      real(c_double), value     :: dt
      real(c_double)            :: errmax
      integer(c_int32_t), value :: mem
      integer(c_int32_t)        :: rc
    end function tflows_calcps_correct

    ! CFLmax(1:Nmat), PeMax(1:Nmat) of CorUVW.f90:124-147 (after uvw_setup)
    function tflows_calcps_cfl_peclet(h, dt, visc, rho_n, flux, cflmax, pemax, mem, stream) result(rc) &
        bind(C, name="tflows_calcps_cfl_peclet")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value        :: h, rho_n, flux, stream
      real(c_double), value     :: dt, visc
      real(c_double)            :: cflmax(*), pemax(*)
      integer(c_int32_t), value :: mem
      integer(c_int32_t)        :: rc
    end function tflows_calcps_cfl_peclet

    !--- momentum assembly (NewUVW.f90:137-499) -----------------------------
    function tflows_calcps_uvw_setup(h, scoef, ff, material, nmaterials, mem, stream) result(rc) &
        bind(C, name="tflows_calcps_uvw_setup")
      import :: c_int32_t, c_ptr
      type(c_ptr), value        :: h, scoef, ff, material, stream
      integer(c_int32_t), value :: nmaterials, mem
      integer(c_int32_t)        :: rc
    end function tflows_calcps_uvw_setup

    function tflows_calcps_uvw_assemble(h, p, rho_n, rho_o, rho_oo, ui_n, ui_o, ui_oo, source, pi, dUidi, &
                                        dUidj, dUidk, dUjdi, dUkdi, dUjdj, dUkdk, flux_u, flux, aval, b, abou, &
                                        asave, rhoui, mem, stream) result(rc) &
        bind(C, name="tflows_calcps_uvw_assemble")
      import :: c_int32_t, c_ptr, tflows_uvw_params
      type(c_ptr), value          :: h, rho_n, rho_o, rho_oo, ui_n, ui_o, ui_oo, source, pi, dUidi, dUidj, dUidk, &
                                     dUjdi, dUkdi, dUjdj, dUkdk, flux_u, flux, aval, b, abou, asave, rhoui, &
                                     stream
      type(tflows_uvw_params)     :: p
      integer(c_int32_t), value   :: mem
      integer(c_int32_t)          :: rc
    end function tflows_calcps_uvw_assemble

    function tflows_calcps_uvw_update(h, rhoui, rho_n, ui_n, mem, stream) result(rc) &
        bind(C, name="tflows_calcps_uvw_update")
      import :: c_int32_t, c_ptr
      type(c_ptr), value        :: h, rhoui, rho_n, ui_n, stream
      integer(c_int32_t), value :: mem
      integer(c_int32_t)        :: rc
    end function tflows_calcps_uvw_update

    !--- mixture-fraction assembly (CalcZmix.f90:100-405) -------------------
    function tflows_calcps_zmix_assemble(h, p, rho_n, rho_o, rho_oo, z_n, z_o, z_oo, source, dZdx, dZdy, dZdz, &
                                         flux_u, flux, aval, b, abou, rhoz, mem, stream) result(rc) &
        bind(C, name="tflows_calcps_zmix_assemble")
      import :: c_int32_t, c_ptr, tflows_uvw_params
      type(c_ptr), value          :: h, rho_n, rho_o, rho_oo, z_n, z_o, z_oo, source, dZdx, dZdy, dZdz, flux_u, &
                                     flux, aval, b, abou, rhoz, stream
      type(tflows_uvw_params)     :: p
      integer(c_int32_t), value   :: mem
      integer(c_int32_t)          :: rc
    end function tflows_calcps_zmix_assemble

    function tflows_calcps_zmix_update(h, rhoz, rho_n, z_n, mem, stream) result(rc) &
        bind(C, name="tflows_calcps_zmix_update")
      import :: c_int32_t, c_ptr
      type(c_ptr), value        :: h, rhoz, rho_n, z_n, stream
      integer(c_int32_t), value :: mem
      integer(c_int32_t)        :: rc
    end function tflows_calcps_zmix_update
  end interface
end module tflows_krylov
