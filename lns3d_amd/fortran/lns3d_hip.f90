!-----------------------------------------------------------------------------
! Fortran host layer of lns3d_amd (built into liblns3d_fortran.so, linked
! against liblns3d_amd.so).
!
! 1. Drop-in externals with the reference's names and argument lists, so a
!    Fortran driver links them in place of src/cpenta1bc.f, cpenta2bc.f,
!    cpenta1p.f, cpenta2p.f, lrhs3d.f90, lhs1f3d.f90, lhs2f3d.f90, cgrad.f90
!    and cgrad2.f90:
!       call lrhs3D(vl, rl, vml, xl, yl)                   ! lrhs3d.f90:2
!       call lhs1f3D(mat, Ah, Dh, Dhi, Vh11, ABhi, dtl, vl) ! lhs1f3d.f90:2
!       call lhs2f3D(mat, Bh, Dh, Dhi, Vh22, ABhi, dtl, vl) ! lhs2f3d.f90:2
!       call cgrad(ndof, nx, ny, v, g1v, g2v, dx, dy, optx, opty, xper, yper,
!                  lsym, rsym, bsym, tsym, carp)           ! cgrad.f90:2-3
!       call cgrad2(ndof, nx, ny, v, g1v, g11v, g12v, g22v, dx, dy, ...)
!    The first three read the module state the host forwards once, after
!    genmtrx and set_time_step, with c_set_module_state(st) (type
!    lns3d_module_state below, one bind(C) struct).
!       call cpenta1bc(nsys, n, nblk, mat, rhs, code)      ! cpenta1bc.f:2
!       call cpenta2bc(n, nsys, nblk, mat, rhs, code)      ! cpenta2bc.f:2
!       call cpenta1p (nsys, np, nblk, mat, rhs, per, per2, code)
!       call cpenta2p (np, nsys, nblk, mat, rhs, per, per2, code)
!    Arguments keep the reference's meaning: mat and rhs in its column-major
!    layouts, code 0 factor / 1 solve / 2 both.  The factors stay on the GPU
!    (keyed by the address of mat and the shape), so mat is not overwritten;
!    per and per2 are not written.  The reference reports no errors; the
!    status of the last call is kept for lns3d_hip_status().
! 2. Module lns3d_hip: iso_c_binding interfaces to the C ABI
!    (include/lns3d_amd.h) and a driver of the implicit step with host
!    arrays (the time loop of src/impdrv3d.f90:126-406):
!       call lns3d_hip_init(prm, metric, vm)   ! coefficients, dtl, factors
!       call lns3d_hip_scheme(impl, niter)     ! optional: the deck's impl, niter
!       call lns3d_hip_sponge / _top / _inflow / _smoother, then
!       call lns3d_hip_factor()                ! the deck's sponge, boundary values, eps_e
!       call lns3d_hip_put(v)                  ! state to the device
!       call lns3d_hip_advance(nsteps, dv)     ! steps on the device
!       call lns3d_hip_get(v)                  ! state back
!       call lns3d_hip_free()
!-----------------------------------------------------------------------------
module lns3d_hip
  use iso_c_binding
  implicit none
  private

  ! mirror of lns3d_params (include/lns3d_amd.h)
  type, bind(C), public :: lns3d_params
    integer(c_int) :: nx, ny, optx, opty, xper, yper, wall, wallt, top, left, right
    integer(c_int) :: linear, mattyp, loctime
    real(c_double) :: Ma, Re, Pr, gamma, kz, omega, cflmax, dtmax, alfa
    integer(c_int) :: periodic_exact, precision
    integer(c_int) :: wall_is = 0, wall_ie = 0, idamp = 0, use_calcd = 0, update_lhs = 0
  end type lns3d_params

  ! mirror of lns3d_module_state (include/lns3d_amd.h): the module state of
  ! src/global.f90, local2d.f90 and local3d.f90 that lrhs3D, lhs1f3D and
  ! lhs2f3D read, forwarded once by the host through lns3d_set_module_state
  type, bind(C), public :: lns3d_module_state
    integer(c_int) :: nx = 0, ny = 0, ndof = 5
    real(c_double) :: alfa = 1.0_c_double, dxi = 0, deta = 0
    integer(c_int) :: optx = 0, opty = 0, carp = 0
    integer(c_int) :: xper = 0, yper = 0, lsym = 0, rsym = 0, bsym = 0, tsym = 0
    integer(c_int) :: wall = 0, wallt = 0, top = 0, left = 0, right = 0
    integer(c_int) :: is = 0, ie = 0, navier = 1
    integer(c_int) :: linear = 2, mattyp = 0, loctime = 0, idamp = 0
    real(c_double) :: Ma = 0, Re = 0, Pr = 0, gamma = 0, gamma1 = 0, kz = 0, omega = 0, x0 = 0, S0 = 0
    real(c_double) :: cflmax = 0, dtmax = 0
    integer(c_int) :: ispg = 0
    type(c_ptr) :: spg = c_null_ptr, spg2 = c_null_ptr
    real(c_double) :: eps_e = 0
    type(c_ptr) :: buff = c_null_ptr
    integer(c_int) :: use_calcd = 0, update_lhs = 0
    type(c_ptr) :: istep = c_null_ptr, iter = c_null_ptr   ! the caller's counters, read per call
    type(c_ptr) :: wamp = c_null_ptr, vm = c_null_ptr, x = c_null_ptr, bnb = c_null_ptr, eig = c_null_ptr
    real(c_double) :: lalpha(2) = 0
    type(c_ptr) :: bump = c_null_ptr
    type(c_ptr) :: Ah = c_null_ptr, Bh = c_null_ptr, Dh = c_null_ptr, Dhi = c_null_ptr
    type(c_ptr) :: Vh11 = c_null_ptr, Vh12 = c_null_ptr, Vh22 = c_null_ptr, ABhi = c_null_ptr
    type(c_ptr) :: dtl = c_null_ptr
  end type lns3d_module_state

  integer(c_int), public :: lns3d_hip_last = 0
  type(c_ptr), save :: ctx = c_null_ptr
  real(c_double), save :: s0_given = -1.0_c_double   ! Sutherland's constant for mattyp = 1
  ! the deck's boundary values (top = 1, left 1/4/5, right 4/5, wall = 3) are
  ! still to come: lns3d_hip_init did not factor, and lns3d_hip_scheme does not
  logical, save :: factor_deferred = .false.

  public :: lns3d_hip_status, lns3d_hip_message, lns3d_hip_record
  public :: lns3d_hip_init, lns3d_hip_put, lns3d_hip_get, lns3d_hip_advance, lns3d_hip_free
  public :: lns3d_hip_factor, lns3d_hip_scheme, lns3d_hip_sponge, lns3d_hip_top, lns3d_hip_inflow
  public :: lns3d_hip_smoother, lns3d_hip_wall, lns3d_hip_sutherland
  public :: c_cpenta1bc, c_cpenta2bc, c_cpenta1p, c_cpenta2p, c_set_periodic_exact
  public :: c_set_module_state, c_drop_lrhs3d, c_drop_lhs1f3d, c_drop_lhs2f3d, c_drop_cgrad, c_drop_cgrad2
  public :: lns3d_hip_flag

  interface
    function c_last_error() bind(C, name="lns3d_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    integer(c_int) function c_cpenta1bc(nsys, n, nblk, mat, rhs, code) bind(C, name="lns3d_cpenta1bc")
      import :: c_int, c_ptr
      integer(c_int), value :: nsys, n, nblk, code
      type(c_ptr), value :: mat, rhs
    end function
    integer(c_int) function c_cpenta2bc(n, nsys, nblk, mat, rhs, code) bind(C, name="lns3d_cpenta2bc")
      import :: c_int, c_ptr
      integer(c_int), value :: n, nsys, nblk, code
      type(c_ptr), value :: mat, rhs
    end function
    integer(c_int) function c_cpenta1p(nsys, np, nblk, mat, rhs, per, per2, code) &
        bind(C, name="lns3d_cpenta1p")
      import :: c_int, c_ptr
      integer(c_int), value :: nsys, np, nblk, code
      type(c_ptr), value :: mat, rhs, per, per2
    end function
    integer(c_int) function c_cpenta2p(np, nsys, nblk, mat, rhs, per, per2, code) &
        bind(C, name="lns3d_cpenta2p")
      import :: c_int, c_ptr
      integer(c_int), value :: np, nsys, nblk, code
      type(c_ptr), value :: mat, rhs, per, per2
    end function
    integer(c_int) function c_set_periodic_exact(exact) bind(C, name="lns3d_set_periodic_exact")
      import :: c_int
      integer(c_int), value :: exact
    end function
    integer(c_int) function c_ctx_create(p, out) bind(C, name="lns3d_ctx_create")
      import :: c_int, c_ptr, lns3d_params
      type(lns3d_params), intent(in) :: p
      type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function c_ctx_destroy(c) bind(C, name="lns3d_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
    end function
    integer(c_int) function c_setup(c, metric, vm) bind(C, name="lns3d_setup")
      import :: c_int, c_ptr
      type(c_ptr), value :: c, metric, vm
    end function
    integer(c_int) function c_factor(c) bind(C, name="lns3d_factor")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
    end function
    integer(c_int) function c_state_upload(c, v) bind(C, name="lns3d_state_upload")
      import :: c_int, c_ptr
      type(c_ptr), value :: c, v
    end function
    integer(c_int) function c_state_download(c, v) bind(C, name="lns3d_state_download")
      import :: c_int, c_ptr
      type(c_ptr), value :: c, v
    end function
    integer(c_int) function c_advance(c, nsteps, dv) bind(C, name="lns3d_advance")
      import :: c_int, c_ptr
      type(c_ptr), value :: c, dv
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function c_set_time_scheme(c, impl) bind(C, name="lns3d_set_time_scheme")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: impl
    end function
    integer(c_int) function c_set_corrector_iterations(c, niter) &
        bind(C, name="lns3d_set_corrector_iterations")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: niter
    end function
    integer(c_int) function c_set_sponge(c, ispg, spg, spg2, x, wamp, x0) bind(C, name="lns3d_set_sponge")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c, spg, spg2, x, wamp
      integer(c_int), value :: ispg
      real(c_double), value :: x0
    end function
    integer(c_int) function c_set_top_forcing(c, x, wamp, x0) bind(C, name="lns3d_set_top_forcing")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c, x, wamp
      real(c_double), value :: x0
    end function
    integer(c_int) function c_set_inflow(c, x, x0, eig, lalpha) bind(C, name="lns3d_set_inflow")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c, x, eig, lalpha
      real(c_double), value :: x0
    end function
    integer(c_int) function c_set_sutherland(c, s0) bind(C, name="lns3d_set_sutherland")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), value :: s0
    end function
    integer(c_int) function c_set_wall(c, uw) bind(C, name="lns3d_set_wall")
      import :: c_int, c_ptr
      type(c_ptr), value :: c, uw
    end function
    integer(c_int) function c_set_smoother(c, eps_e, buff) bind(C, name="lns3d_set_smoother")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c, buff
      real(c_double), value :: eps_e
    end function
    integer(c_int) function c_set_module_state(st) bind(C, name="lns3d_set_module_state")
      import :: c_int, lns3d_module_state
      type(lns3d_module_state), intent(in) :: st
    end function
    integer(c_int) function c_drop_lrhs3d(vl, rl, vml, xl, yl) bind(C, name="lns3d_drop_lrhs3d")
      import :: c_int, c_ptr
      type(c_ptr), value :: vl, rl, vml, xl, yl
    end function
    integer(c_int) function c_drop_lhs1f3d(mat, ah, dh, dhi, vh11, abhi, dtl, vl) bind(C, name="lns3d_drop_lhs1f3d")
      import :: c_int, c_ptr
      type(c_ptr), value :: mat, ah, dh, dhi, vh11, abhi, dtl, vl
    end function
    integer(c_int) function c_drop_lhs2f3d(mat, bh, dh, dhi, vh22, abhi, dtl, vl) bind(C, name="lns3d_drop_lhs2f3d")
      import :: c_int, c_ptr
      type(c_ptr), value :: mat, bh, dh, dhi, vh22, abhi, dtl, vl
    end function
    integer(c_int) function c_drop_cgrad(ndof, nx, ny, v, g1v, g2v, dx, dy, optx, opty, xper, yper, &
        lsym, rsym, bsym, tsym, carp) bind(C, name="lns3d_drop_cgrad")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: ndof, nx, ny, optx, opty, xper, yper, lsym, rsym, bsym, tsym, carp
      type(c_ptr), value :: v, g1v, g2v
      real(c_double), value :: dx, dy
    end function
    integer(c_int) function c_drop_cgrad2(ndof, nx, ny, v, g1v, g11v, g12v, g22v, dx, dy, optx, opty, &
        xper, yper, lsym, rsym, bsym, tsym, carp) bind(C, name="lns3d_drop_cgrad2")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: ndof, nx, ny, optx, opty, xper, yper, lsym, rsym, bsym, tsym, carp
      type(c_ptr), value :: v, g1v, g11v, g12v, g22v
      real(c_double), value :: dx, dy
    end function
    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! a Fortran logical as the C ABI's 0/1
  integer(c_int) function lns3d_hip_flag(l)
    logical, intent(in) :: l
    lns3d_hip_flag = merge(1_c_int, 0_c_int, l)
  end function

  subroutine lns3d_hip_record(rc)
    integer(c_int), intent(in) :: rc
    lns3d_hip_last = rc
  end subroutine

  ! status of the last call (0: success; the codes of include/lns3d_amd.h)
  integer function lns3d_hip_status()
    lns3d_hip_status = int(lns3d_hip_last)
  end function

  ! message of the last failure
  function lns3d_hip_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = c_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [max(n, 1)])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

  ! metric(nx,ny,10) and vm(5,nx,ny) as src/gengrid.f90 and genmean.f90
  ! hold them; generates the coefficients and dtl on the device and factors
  ! both directions (src/impdrv3d.f90:248-269, 310-319).  Selectors whose
  ! boundary values come from the deck (top = 1, left = 1, 4 or 5, right = 4
  ! or 5) are not factored here: give those values (lns3d_hip_top,
  ! lns3d_hip_inflow), then call lns3d_hip_factor.
  subroutine lns3d_hip_init(prm, metric, vm)
    type(lns3d_params), intent(in) :: prm
    real(c_double), intent(in), target :: metric(*), vm(*)
    integer(c_int) :: rc
    logical :: needs_values
    if (c_associated(ctx)) call lns3d_hip_free()
    needs_values = prm%top == 1 .or. ((prm%left == 1 .or. prm%left == 4 .or. prm%left == 5 .or. &
                                       prm%right == 4 .or. prm%right == 5) .and. prm%xper == 0) .or. &
                   (prm%wall == 3 .and. prm%yper == 0)
    rc = c_ctx_create(prm, ctx)
    if (rc == 0 .and. prm%mattyp == 1) rc = c_set_sutherland(ctx, s0_given)
    if (rc == 0) rc = c_setup(ctx, c_loc(metric), c_loc(vm))
    if (rc == 0 .and. .not. needs_values) rc = c_factor(ctx)
    factor_deferred = needs_values
    call lns3d_hip_record(rc)
  end subroutine

  ! Sutherland's constant S0 = 110.4 K / Tref (src/input.f90:41-47) for a
  ! mattyp = 1 deck; call before lns3d_hip_init
  subroutine lns3d_hip_sutherland(s0)
    real(c_double), intent(in) :: s0
    s0_given = s0
  end subroutine

  ! forms the factors again (after any of the setters below)
  subroutine lns3d_hip_factor()
    integer(c_int) :: rc
    rc = c_factor(ctx)
    if (rc == 0) factor_deferred = .false.
    call lns3d_hip_record(rc)
  end subroutine

  ! sponge: the module state src/sponge.f90:20 (init_sponge) leaves, spg and
  ! spg2 (nx,ny), the physical x (nx,ny), wamp (nx), x0
  subroutine lns3d_hip_sponge(ispg, spg, spg2, x, wamp, x0)
    integer, intent(in) :: ispg
    real(c_double), intent(in), target :: spg(*), spg2(*), x(*), wamp(*)
    real(c_double), intent(in) :: x0
    call lns3d_hip_record(c_set_sponge(ctx, int(ispg, c_int), c_loc(spg), c_loc(spg2), c_loc(x), &
                                       c_loc(wamp), x0))
  end subroutine

  ! top = 1: the incoming wave from x (nx,ny), wamp (nx) and x0
  subroutine lns3d_hip_top(x, wamp, x0)
    real(c_double), intent(in), target :: x(*), wamp(*)
    real(c_double), intent(in) :: x0
    call lns3d_hip_record(c_set_top_forcing(ctx, c_loc(x), c_loc(wamp), x0))
  end subroutine

  ! left = 1, 4, 5 and right = 4, 5: x (nx,ny), x0, the eigenfunction
  ! profiles eig(10,ny) (rhor, rhoi, ur, ui, vr, vi, wr, wi, tr, ti) and the
  ! inflow namelist's lalpha
  subroutine lns3d_hip_inflow(x, x0, eig, lalpha)
    real(c_double), intent(in), target :: x(*), eig(*)
    real(c_double), intent(in) :: x0
    complex(c_double_complex), intent(in) :: lalpha
    real(c_double), target :: la(2)
    la(1) = real(lalpha, c_double)
    la(2) = aimag(lalpha)
    call lns3d_hip_record(c_set_inflow(ctx, c_loc(x), x0, c_loc(eig), c_loc(la)))
  end subroutine

  ! wall = 3: the bump's wall values uw(nx,5): u1w, u2w, u3w, tw, twp
  subroutine lns3d_hip_wall(uw)
    complex(c_double_complex), intent(in), target :: uw(*)
    call lns3d_hip_record(c_set_wall(ctx, c_loc(uw)))
  end subroutine

  ! smoother: eps_e and the damping profile buff (nx,ny) of init_buffer
  subroutine lns3d_hip_smoother(eps_e, buff)
    real(c_double), intent(in) :: eps_e
    real(c_double), intent(in), target :: buff(*)
    call lns3d_hip_record(c_set_smoother(ctx, eps_e, c_loc(buff)))
  end subroutine

  ! the deck's time scheme and corrector iterations (src/impdrv3d.f90:173-199):
  ! impl 1 backward Euler, 2 two-step backward, 3 alfa = 1/2; niter >= 1
  ! (impl 1 or 2).  Forms the factors again (unless the deck's boundary
  ! values are still to come: then lns3d_hip_factor after their setters); a
  ! new run starts at lstep 0.
  subroutine lns3d_hip_scheme(impl, niter)
    integer, intent(in) :: impl, niter
    integer(c_int) :: rc
    rc = c_set_time_scheme(ctx, int(impl, c_int))
    if (rc == 0) rc = c_set_corrector_iterations(ctx, int(niter, c_int))
    if (rc == 0 .and. .not. factor_deferred) rc = c_factor(ctx)
    call lns3d_hip_record(rc)
  end subroutine

  subroutine lns3d_hip_put(v)
    complex(c_double_complex), intent(in), target :: v(*)
    call lns3d_hip_record(c_state_upload(ctx, c_loc(v)))
  end subroutine

  subroutine lns3d_hip_get(v)
    complex(c_double_complex), intent(inout), target :: v(*)
    call lns3d_hip_record(c_state_download(ctx, c_loc(v)))
  end subroutine

  subroutine lns3d_hip_advance(nsteps, dv)
    integer, intent(in) :: nsteps
    complex(c_double_complex), intent(inout), target :: dv(*)
    call lns3d_hip_record(c_advance(ctx, int(nsteps, c_int), c_loc(dv)))
  end subroutine

  subroutine lns3d_hip_free()
    integer(c_int) :: rc
    if (c_associated(ctx)) then
      rc = c_ctx_destroy(ctx)
      ctx = c_null_ptr
    end if
  end subroutine

end module lns3d_hip

!-----------------------------------------------------------------------------
! drop-in externals (external names cpenta1bc_ ... under flang and gfortran)
!-----------------------------------------------------------------------------
subroutine cpenta1bc(nsys, n, nblk, mat, rhs, code)
  use lns3d_hip
  use iso_c_binding
  implicit none
  integer :: nsys, n, nblk, code
  complex(c_double_complex), target :: mat(*), rhs(*)
  call lns3d_hip_record(c_cpenta1bc(int(nsys, c_int), int(n, c_int), int(nblk, c_int), &
                                    c_loc(mat), c_loc(rhs), int(code, c_int)))
end subroutine

subroutine cpenta2bc(n, nsys, nblk, mat, rhs, code)
  use lns3d_hip
  use iso_c_binding
  implicit none
  integer :: n, nsys, nblk, code
  complex(c_double_complex), target :: mat(*), rhs(*)
  call lns3d_hip_record(c_cpenta2bc(int(n, c_int), int(nsys, c_int), int(nblk, c_int), &
                                    c_loc(mat), c_loc(rhs), int(code, c_int)))
end subroutine

subroutine cpenta1p(nsys, np, nblk, mat, rhs, per, per2, code)
  use lns3d_hip
  use iso_c_binding
  implicit none
  integer :: nsys, np, nblk, code
  complex(c_double_complex), target :: mat(*), rhs(*), per(*), per2(*)
  call lns3d_hip_record(c_cpenta1p(int(nsys, c_int), int(np, c_int), int(nblk, c_int), &
                                   c_loc(mat), c_loc(rhs), c_loc(per), c_loc(per2), &
                                   int(code, c_int)))
end subroutine

subroutine cpenta2p(np, nsys, nblk, mat, rhs, per, per2, code)
  use lns3d_hip
  use iso_c_binding
  implicit none
  integer :: np, nsys, nblk, code
  complex(c_double_complex), target :: mat(*), rhs(*), per(*), per2(*)
  call lns3d_hip_record(c_cpenta2p(int(np, c_int), int(nsys, c_int), int(nblk, c_int), &
                                   c_loc(mat), c_loc(rhs), c_loc(per), c_loc(per2), &
                                   int(code, c_int)))
end subroutine

!-----------------------------------------------------------------------------
! RHS, LHS assembly and derivative drop-ins with the reference's argument
! lists (lrhs3d.f90:2, lhs1f3d.f90:2, lhs2f3d.f90:2, cgrad.f90:2-3,
! cgrad2.f90:2-3).  lrhs3D, lhs1f3D and lhs2f3D read the module state the
! host forwarded with lns3d_set_module_state (c_set_module_state).
!-----------------------------------------------------------------------------
subroutine lrhs3D(vl, rl, vml, xl, yl)
  use lns3d_hip
  use iso_c_binding
  implicit none
  complex(c_double_complex), target :: vl(*), rl(*)
  real(c_double), target :: vml(*), xl(*), yl(*)
  call lns3d_hip_record(c_drop_lrhs3d(c_loc(vl), c_loc(rl), c_loc(vml), c_loc(xl), c_loc(yl)))
end subroutine

subroutine lhs1f3D(mat, Ah, Dh, Dhi, Vh11, ABhi, dtl, vl)
  use lns3d_hip
  use iso_c_binding
  implicit none
  complex(c_double_complex), target :: mat(*), vl(*)
  real(c_double), target :: Ah(*), Dh(*), Dhi(*), Vh11(*), ABhi(*), dtl(*)
  call lns3d_hip_record(c_drop_lhs1f3d(c_loc(mat), c_loc(Ah), c_loc(Dh), c_loc(Dhi), c_loc(Vh11), &
                                       c_loc(ABhi), c_loc(dtl), c_loc(vl)))
end subroutine

subroutine lhs2f3D(mat, Bh, Dh, Dhi, Vh22, ABhi, dtl, vl)
  use lns3d_hip
  use iso_c_binding
  implicit none
  complex(c_double_complex), target :: mat(*), vl(*)
  real(c_double), target :: Bh(*), Dh(*), Dhi(*), Vh22(*), ABhi(*), dtl(*)
  call lns3d_hip_record(c_drop_lhs2f3d(c_loc(mat), c_loc(Bh), c_loc(Dh), c_loc(Dhi), c_loc(Vh22), &
                                       c_loc(ABhi), c_loc(dtl), c_loc(vl)))
end subroutine

subroutine cgrad(ndof, nx, ny, v, g1v, g2v, dx, dy, optx, opty, xper, yper, lsym, rsym, bsym, tsym, carp)
  use lns3d_hip
  use iso_c_binding
  implicit none
  integer :: ndof, nx, ny, optx, opty
  logical :: xper, yper, lsym, rsym, bsym, tsym, carp
  complex(c_double_complex), target :: v(*), g1v(*), g2v(*)
  real(c_double) :: dx, dy
  call lns3d_hip_record(c_drop_cgrad(int(ndof, c_int), int(nx, c_int), int(ny, c_int), c_loc(v), c_loc(g1v), &
                                     c_loc(g2v), dx, dy, int(optx, c_int), int(opty, c_int), &
                                     lns3d_hip_flag(xper), lns3d_hip_flag(yper), lns3d_hip_flag(lsym), &
                                     lns3d_hip_flag(rsym), lns3d_hip_flag(bsym), lns3d_hip_flag(tsym), &
                                     lns3d_hip_flag(carp)))
end subroutine

subroutine cgrad2(ndof, nx, ny, v, g1v, g11v, g12v, g22v, dx, dy, optx, opty, xper, yper, lsym, rsym, &
                  bsym, tsym, carp)
  use lns3d_hip
  use iso_c_binding
  implicit none
  integer :: ndof, nx, ny, optx, opty
  logical :: xper, yper, lsym, rsym, bsym, tsym, carp
  complex(c_double_complex), target :: v(*), g1v(*), g11v(*), g12v(*), g22v(*)
  real(c_double) :: dx, dy
  call lns3d_hip_record(c_drop_cgrad2(int(ndof, c_int), int(nx, c_int), int(ny, c_int), c_loc(v), c_loc(g1v), &
                                      c_loc(g11v), c_loc(g12v), c_loc(g22v), dx, dy, int(optx, c_int), &
                                      int(opty, c_int), lns3d_hip_flag(xper), lns3d_hip_flag(yper), &
                                      lns3d_hip_flag(lsym), lns3d_hip_flag(rsym), lns3d_hip_flag(bsym), &
                                      lns3d_hip_flag(tsym), lns3d_hip_flag(carp)))
end subroutine

! status of the last drop-in call, callable from fixed-form code
integer function lns3d_hip_status_f()
  use lns3d_hip
  implicit none
  lns3d_hip_status_f = lns3d_hip_status()
end function
