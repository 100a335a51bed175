!=============================================================================!
! Direct-call harness over the reference's own Fortran routines
! (sscollis/lns3d src/).  Test infrastructure only: it is linked into
! oracle/_ref/libref.so by oracle/build_ref.sh and called from tests through
! ctypes.  It replaces the file-reading start-up of src/lns3d.f90:61-149
! (input, gengrid, genmean) by setters, then calls the reference routines
! exactly as src/impdrv3d.f90:130-356 does for impl=1, niter=1, iLHS=1.
!=============================================================================!
module refdrv_state
  use iso_c_binding
  implicit none
  complex, allocatable :: matx(:,:,:,:,:), maty(:,:,:,:,:)
  complex, allocatable :: lmatx(:,:,:,:,:), lr(:,:,:)
  complex, allocatable :: perx(:,:,:,:,:), per2x(:,:,:,:,:)
  complex, allocatable :: pery(:,:,:,:,:), per2y(:,:,:,:,:)
  real, allocatable    :: dtlh(:,:)
  ! impl = 2: the field at the start of the previous step, and the steps taken
  complex, allocatable :: vhist(:,:,:), v2old(:,:,:)
  integer :: lstep_h = 0
  ! forcing of the shift-invert driver's steps (src/si_eigdrv3d_new.f90:325-
  ! 358 with ebe3D, src/si_eigdrv3d.f90:319-345); unallocated: none
  complex, allocatable :: fsi(:,:,:)
  ! wall-clock seconds of the phases of the last ref_step: 1 residual
  ! (lrhs3D, rbe3D, rhsBC3D), 2 xi solve with its transposes, 3 eta solve,
  ! 4 update and itrBC3D, 5 forming and factoring the LHS (first step)
  real(c_double) :: phase_t(5) = 0
end module refdrv_state

!-----------------------------------------------------------------------------
! ip: 1 nx, 2 ny, 3 optx, 4 opty, 5 carp, 6 xper, 7 yper, 8 lsym, 9 rsym,
!     10 wall, 11 wallt, 12 top, 13 left, 14 right, 15 linear, 16 impl,
!     17 loctime, 18 mattyp, 19 ispg, 20 navier, 21 usecalcd, 22 idamp
! rp: 1 Ma, 2 Re, 3 Pr, 4 gamma, 5 kz, 6 omega, 7 cflmax, 8 dtmax, 9 eps_e,
!     10 x0
!-----------------------------------------------------------------------------
subroutine ref_init(ip, rp) bind(C, name="ref_init")
  use iso_c_binding
  use global
  use local2d
  use local3d
  use refdrv_state
  use pot, only: g1p
  implicit none
  integer(c_int), intent(in) :: ip(22)
  real(c_double), intent(in) :: rp(10)
  integer :: i, j

  nx = ip(1); ny = ip(2); nz = 1; ndof = 5
  optx = ip(3); opty = ip(4)
  carp = ip(5) /= 0; xper = ip(6) /= 0; yper = ip(7) /= 0
  lsym = ip(8) /= 0; rsym = ip(9) /= 0; bsym = .false.; tsym = .false.
  wall = ip(10); wallt = ip(11); top = ip(12); left = ip(13); right = ip(14)
  linear = ip(15); impl = ip(16); loctime = ip(17); mattyp = ip(18)
  ispg = ip(19); Navier = ip(20) /= 0; useCalcd = ip(21) /= 0
  Ma = rp(1); Re = rp(2); Pr = rp(3); gamma = rp(4); gamma1 = gamma - one
  kz = rp(5); omega = rp(6); cflmax = rp(7); dtmax = rp(8); eps_e = rp(9)
  x0 = rp(10)
  iLHS = 1; niter = 1; ires = 0; idamp = ip(22); general = .true.
  ! as src/input.f90:367-371 does after reading the deck: the complex
  ! analysis runs with linear = 1 (the only value src/lns3d.f90:153-167 lets
  ! reach impdrv3d)
  complex_analysis = (kz /= zero) .or. (linear == 2)
  if (linear == 2) linear = 1
  is = 1; ie = nx; js = 1; je = ny; updateLHS = .false.
  istep = 0; iter = 1; lstep = 0; alfa = one; time = zero

  dxi = one / float(nx-1)
  deta = one / float(ny-1)

  ! the damper (src/lns3d.f90:140 calls init_damper after the deck): damp is
  ! zero on the outflow column for idamp = 1.  genmtrx then calls calcp,
  ! whose gpr only the nonlinear forms read (src/genmtrx.f90:273, 646) but
  ! which reads the potential-flow gradient g1p, allocated only when that
  ! input was read (src/potential.f90:78): zeros stand in for it here
  if (idamp == 1) then
    if (allocated(damp)) deallocate(damp)
    call init_damper
    if (allocated(g1p)) deallocate(g1p)
    allocate(g1p(max(nx,ny), max(nx,ny)))
    g1p = zero
  end if

  if (allocated(x)) then
    deallocate(x, y, m1, m2, n1, n2, m11, m12, m22, n11, n12, n22)
    deallocate(m1m1, m1m2, m2m2, n1n1, n1n2, n2n2, m1n1, m1n2, m2n1, m2n2)
    deallocate(bnl, bnr, bnb, bnt, vm)
    deallocate(g1v, g2v, g11v, g12v, g22v, Ah, Bh, Dh, Vh11, Vh12, Vh22)
    deallocate(c1v, c2v, c11v, c12v, c22v, Dhi, ABhi)
  end if
  allocate(x(nx,ny), y(nx,ny), m1(nx,ny), m2(nx,ny), n1(nx,ny), n2(nx,ny))
  allocate(m11(nx,ny), m12(nx,ny), m22(nx,ny), n11(nx,ny), n12(nx,ny), n22(nx,ny))
  allocate(m1m1(nx,ny), m1m2(nx,ny), m2m2(nx,ny), n1n1(nx,ny), n1n2(nx,ny), n2n2(nx,ny))
  allocate(m1n1(nx,ny), m1n2(nx,ny), m2n1(nx,ny), m2n2(nx,ny))
  allocate(bnl(ny,2), bnr(ny,2), bnb(nx,2), bnt(nx,2), vm(ndof,nx,ny))
  call mlocal2d
  call mlocal3d
  x = zero; y = zero; vm = zero
  if (allocated(dtlh)) deallocate(dtlh)
  allocate(dtlh(nx,ny))
  dtlh = zero
  ! forcing state of an earlier case: itrBC3D (top = 1) allocates wamp itself
  ! at istep = 0 (src/itrbc3d.f90:145-160)
  if (allocated(spg)) deallocate(spg)
  if (allocated(spg2)) deallocate(spg2)
  if (allocated(wamp)) deallocate(wamp)
end subroutine ref_init

! grid and metrics (each nx*ny, Fortran order), then the products as in
! src/gengrid.f90:84-155
subroutine ref_set_grid(xl, yl, mt) bind(C, name="ref_set_grid")
  use iso_c_binding
  use global
  implicit none
  real(c_double), intent(in) :: xl(nx,ny), yl(nx,ny), mt(nx,ny,10)
  integer :: i, j
  x = xl; y = yl
  m1 = mt(:,:,1); m2 = mt(:,:,2); n1 = mt(:,:,3); n2 = mt(:,:,4)
  m11 = mt(:,:,5); m12 = mt(:,:,6); m22 = mt(:,:,7)
  n11 = mt(:,:,8); n12 = mt(:,:,9); n22 = mt(:,:,10)
  m1m1 = m1*m1; m1m2 = m1*m2; m2m2 = m2*m2
  n1n1 = n1*n1; n1n2 = n1*n2; n2n2 = n2*n2
  m1n1 = m1*n1; m1n2 = m1*n2; m2n1 = m2*n1; m2n2 = m2*n2
  do i = 1, nx
    bnb(i,1) = -n1(i,1) / sqrt( n1(i,1)**2 + n2(i,1)**2 )
    bnb(i,2) = -n2(i,1) / sqrt( n1(i,1)**2 + n2(i,1)**2 )
    bnt(i,1) = n1(i,ny) / sqrt( n1(i,ny)**2 + n2(i,ny)**2 )
    bnt(i,2) = n2(i,ny) / sqrt( n1(i,ny)**2 + n2(i,ny)**2 )
  end do
  do j = 1, ny
    bnl(j,1) = -m1(1,j) / sqrt( m1(1,j)**2 + m2(1,j)**2 )
    bnl(j,2) = -m2(1,j) / sqrt( m1(1,j)**2 + m2(1,j)**2 )
    bnr(j,1) = m1(nx,j) / sqrt( m1(nx,j)**2 + m2(nx,j)**2 )
    bnr(j,2) = m2(nx,j) / sqrt( m1(nx,j)**2 + m2(nx,j)**2 )
  end do
end subroutine ref_set_grid

subroutine ref_set_mean(vml) bind(C, name="ref_set_mean")
  use iso_c_binding
  use global
  implicit none
  real(c_double), intent(in) :: vml(ndof,nx,ny)
  vm = vml
end subroutine ref_set_mean

! sponge profiles spg, spg2 (nx*ny) and the inflow amplitude wamp (nx), the
! module arrays that src/sponge.f90:20 (init_sponge) and the forcing input
! fill; ispg and x0 come from ref_init
subroutine ref_set_sponge(spgl, spg2l, wampl) bind(C, name="ref_set_sponge")
  use iso_c_binding
  use global
  implicit none
  real(c_double), intent(in) :: spgl(nx,ny), spg2l(nx,ny), wampl(nx)
  if (allocated(spg)) deallocate(spg)
  if (allocated(spg2)) deallocate(spg2)
  if (allocated(wamp)) deallocate(wamp)
  allocate(spg(nx,ny), spg2(nx,ny), wamp(nx))
  spg = spgl; spg2 = spg2l; wamp = wampl
end subroutine ref_set_sponge

! Sutherland's constant (mattyp = 1): S0 nondimensional, src/input.f90:41-47
subroutine ref_set_sutherland(s) bind(C, name="ref_set_sutherland")
  use iso_c_binding
  use global
  implicit none
  real(c_double), value :: s
  S0 = s
end subroutine ref_set_sutherland

! the bump wall (wall = 3): the wall values src/genbump.f90:74-131 forms,
! uw(nx,5) complex: u1w, u2w, u3w, tw, twp
subroutine ref_set_bump(uw) bind(C, name="ref_set_bump")
  use iso_c_binding
  use global
  use bump
  implicit none
  complex(c_double_complex), intent(in) :: uw(nx,5)
  if (allocated(u1w)) deallocate(u1w, u2w, u3w, tw, twp)
  allocate(u1w(nx), u2w(nx), u3w(nx), tw(nx), twp(nx))
  u1w = uw(:,1); u2w = uw(:,2); u3w = uw(:,3); tw = uw(:,4); twp = uw(:,5)
end subroutine ref_set_bump

! eigenfunction inflow (left = 4, right = 4): the profiles src/gendist.f90
! reads from inflow_file (columns rhor, rhoi, ur, ui, vr, vi, wr, wi, tr, ti)
! and the inflow namelist's lalpha, lomega
subroutine ref_set_inflow(eig, la, lo) bind(C, name="ref_set_inflow")
  use iso_c_binding
  use global
  implicit none
  real(c_double), intent(in) :: eig(10,ny), la(2), lo(2)
  if (allocated(rhor)) deallocate(rhor, rhoi, ur, ui, vr, vi, wr, wi, tr, ti)
  allocate(rhor(ny), rhoi(ny), ur(ny), ui(ny), vr(ny), vi(ny), wr(ny), wi(ny), tr(ny), ti(ny))
  rhor = eig(1,:); rhoi = eig(2,:); ur = eig(3,:); ui = eig(4,:); vr = eig(5,:)
  vi = eig(6,:); wr = eig(7,:); wi = eig(8,:); tr = eig(9,:); ti = eig(10,:)
  lalpha = cmplx(la(1), la(2)); lomega = cmplx(lo(1), lo(2))
end subroutine ref_set_inflow

! smoother (src/smoother3d.f90 and the implicit damping terms): eps_e and the
! damping profile buff that src/buffer.f90 (init_buffer) forms from ibuff
subroutine ref_set_smoother(eps, buffl) bind(C, name="ref_set_smoother")
  use iso_c_binding
  use global
  use buffer
  implicit none
  real(c_double), intent(in) :: eps, buffl(nx,ny)
  eps_e = eps
  if (allocated(buff)) deallocate(buff)
  allocate(buff(nx,ny))
  buff = buffl
end subroutine ref_set_smoother

! coefficient generation (src/genmtrx.f90) and time step (src/temporal.f90)
! genmtrx with linear = 1 and wall = 3 calls genbump (src/genmtrx.f90:100),
! which forms the bump's wall values from a namelist file; here those values
! are given through ref_set_bump, as the library takes them, so genmtrx runs
! with the bump switched off (nothing else in it reads wall = 3)
subroutine ref_genmtrx() bind(C, name="ref_genmtrx")
  use global
  implicit none
  real, allocatable :: rl(:,:,:)
  integer :: wall_deck
  allocate(rl(ndof,nx,ny))
  rl = zero
  wall_deck = wall
  if (wall == 3) wall = 1
  call genmtrx(vm, kz, rl)
  wall = wall_deck
  deallocate(rl)
end subroutine ref_genmtrx

subroutine ref_set_time_step(dtl) bind(C, name="ref_set_time_step")
  use iso_c_binding
  use global
  use temporal
  use refdrv_state
  implicit none
  real(c_double), intent(out) :: dtl(nx,ny)
  if (allocated(volinv)) deallocate(volinv, sxj, syj, sxi, acoustic, viscous, syi)
  istep = 0
  call set_time_step(vm, dtlh)
  dtl = dtlh
end subroutine ref_set_time_step

subroutine ref_set_dtl(dtl) bind(C, name="ref_set_dtl")
  use iso_c_binding
  use global
  use refdrv_state
  implicit none
  real(c_double), intent(in) :: dtl(nx,ny)
  dtlh = dtl
end subroutine ref_set_dtl

! coefficient arrays out: Ah,Bh,Dh,Dhi (25 each), Vh11,Vh12,Vh22,ABhi (6 each)
subroutine ref_get_coeffs(a, b, d, di, v11, v12, v22, abi) bind(C, name="ref_get_coeffs")
  use iso_c_binding
  use global
  use local2d
  use local3d
  implicit none
  real(c_double), intent(out) :: a(ndof,ndof,nx,ny), b(ndof,ndof,nx,ny)
  real(c_double), intent(out) :: d(ndof,ndof,nx,ny), di(ndof,ndof,nx,ny)
  real(c_double), intent(out) :: v11(6,nx,ny), v12(6,nx,ny), v22(6,nx,ny), abi(6,nx,ny)
  a = Ah; b = Bh; d = Dh; di = Dhi
  v11 = Vh11; v12 = Vh12; v22 = Vh22; abi = ABhi
end subroutine ref_get_coeffs

subroutine ref_set_coeffs(a, b, d, di, v11, v12, v22, abi) bind(C, name="ref_set_coeffs")
  use iso_c_binding
  use global
  use local2d
  use local3d
  implicit none
  real(c_double), intent(in) :: a(ndof,ndof,nx,ny), b(ndof,ndof,nx,ny)
  real(c_double), intent(in) :: d(ndof,ndof,nx,ny), di(ndof,ndof,nx,ny)
  real(c_double), intent(in) :: v11(6,nx,ny), v12(6,nx,ny), v22(6,nx,ny), abi(6,nx,ny)
  Ah = a; Bh = b; Dh = d; Dhi = di
  Vh11 = v11; Vh12 = v12; Vh22 = v22; ABhi = abi
end subroutine ref_set_coeffs

! lrhs3D with the derivative arrays it leaves in module local3d
subroutine ref_lrhs3d(v, r, g) bind(C, name="ref_lrhs3d")
  use iso_c_binding
  use global
  use local3d
  implicit none
  complex(c_double_complex), intent(in)  :: v(ndof,nx,ny)
  complex(c_double_complex), intent(out) :: r(ndof,nx,ny)
  complex(c_double_complex), intent(out) :: g(ndof,nx,ny,5)
  complex, allocatable :: vl(:,:,:)
  allocate(vl(ndof,nx,ny))
  vl = v
  call lrhs3D(vl, r, vm, x, y)
  g(:,:,:,1) = c1v; g(:,:,:,2) = c2v
  g(:,:,:,3) = c11v; g(:,:,:,4) = c12v; g(:,:,:,5) = c22v
  deallocate(vl)
end subroutine ref_lrhs3d

subroutine ref_rbe3d(r, v, vold) bind(C, name="ref_rbe3d")
  use iso_c_binding
  use global
  use refdrv_state
  implicit none
  complex(c_double_complex), intent(inout) :: r(ndof,nx,ny)
  complex(c_double_complex), intent(in) :: v(ndof,nx,ny), vold(ndof,nx,ny)
  call rbe3D(r, v, vold, dtlh)
end subroutine ref_rbe3d

subroutine ref_rhsbc3d(r, v) bind(C, name="ref_rhsbc3d")
  use iso_c_binding
  use global
  implicit none
  complex(c_double_complex), intent(inout) :: r(ndof,nx,ny)
  complex(c_double_complex), intent(in) :: v(ndof,nx,ny)
  complex, allocatable :: vl(:,:,:)
  allocate(vl(ndof,nx,ny))
  vl = v
  call rhsBC3D(r, vl, vm)
  deallocate(vl)
end subroutine ref_rhsbc3d

subroutine ref_itrbc3d(v) bind(C, name="ref_itrbc3d")
  use iso_c_binding
  use global
  implicit none
  complex(c_double_complex), intent(inout) :: v(ndof,nx,ny)
  call itrBC3D(v, vm)
end subroutine ref_itrbc3d

! assembled (unfactored) LHS blocks, mat(5,ndof,ndof,nx,ny); dir = 1 or 2
subroutine ref_lhs(dir, alfa_in, mat) bind(C, name="ref_lhs")
  use iso_c_binding
  use global
  use local2d
  use local3d
  use refdrv_state
  implicit none
  integer(c_int), value :: dir
  real(c_double), value :: alfa_in
  complex(c_double_complex), intent(out) :: mat(5,ndof,ndof,nx,ny)
  complex, allocatable :: vl(:,:,:)
  allocate(vl(ndof,nx,ny))
  vl = zero
  alfa = alfa_in
  if (dir == 1) then
    call lhs1f3D(mat, Ah, Dh, Dhi, Vh11, ABhi, dtlh, vl)
  else
    call lhs2f3D(mat, Bh, Dh, Dhi, Vh22, ABhi, dtlh, vl)
  end if
  deallocate(vl)
end subroutine ref_lhs

! One implicit step following src/impdrv3d.f90:146-356 (impl=1, niter=1,
! USE_CPENTA1BC path for non-periodic x).  first /= 0 forms and factors the
! LHS (as at istep=1, iter=1).  On return v holds the new field and dv the
! solution increment r that was added.
subroutine ref_step(v, dv, first) bind(C, name="ref_step")
  use iso_c_binding
  use omp_lib, only: omp_get_wtime
  use global
  use local2d
  use local3d
  use refdrv_state
  implicit none
  complex(c_double_complex), intent(inout) :: v(ndof,nx,ny)
  complex(c_double_complex), intent(out) :: dv(ndof,nx,ny)
  integer(c_int), value :: first
  complex, allocatable :: vold(:,:,:), r(:,:,:)
  integer :: i, j, idof, jdof, it
  logical :: form
  real(c_double) :: tc0, tc1

  phase_t = 0
  allocate(vold(ndof,nx,ny), r(ndof,nx,ny))
  if (first /= 0) lstep_h = 0
  ! src/impdrv3d.f90:134-153: the time levels of impl = 2
  if (impl == 2) then
    if (.not. allocated(vhist)) allocate(vhist(ndof,nx,ny), v2old(ndof,nx,ny))
    if (size(vhist,2) /= nx .or. size(vhist,3) /= ny) then
      deallocate(vhist, v2old)
      allocate(vhist(ndof,nx,ny), v2old(ndof,nx,ny))
    end if
    if (lstep_h > 0) v2old = vhist
    vhist = v
  end if
  ! src/impdrv3d.f90:229-231: factors formed at the first step, and again
  ! at the second for impl = 2 (alfa 2/3)
  form = first /= 0 .or. (impl == 2 .and. lstep_h == 1) .or. updateLHS
  if (first /= 0) then
    if (allocated(matx)) deallocate(matx, maty, lmatx, lr)
    allocate(matx(5,ndof,ndof,nx,ny), maty(5,ndof,ndof,nx,ny))
    allocate(lmatx(5,ndof,ndof,ny,nx), lr(ndof,ny,nx))
    if (xper .or. yper) then
      if (allocated(perx)) deallocate(perx, per2x, pery, per2y)
      allocate(perx(2,ndof,ndof,nx,ny), per2x(2,ndof,ndof,nx,ny))
      allocate(pery(2,ndof,ndof,nx,ny), per2y(2,ndof,ndof,nx,ny))
    end if
  end if
  ! the step count of this run (src/impdrv3d.f90:130 do istep = 1, nstep):
  ! useCalcd alternates the D terms with istep - 1 + iter
  istep = lstep_h + 1; iter = 1
  vold = v
  call itrBC3D(v, vm)
  ! the corrector iterations, src/impdrv3d.f90:173-358 (factors at iter 1)
  do it = 1, niter
  iter = it
  if (it > 1 .and. .not. updateLHS) form = .false.
  tc0 = omp_get_wtime()
  ! src/impdrv3d.f90:179-199
  if (impl == 3) then
    call lrhs3D(pt5*(v+vold), r, vm, x, y)
    call rbe3D(r, v, vold, dtlh)
    alfa = pt5
  else if (impl == 2 .and. lstep_h > 0) then
    call lrhs3D(v, r, vm, x, y)
    call rtb3D(r, v, vold, v2old, dtlh)
    alfa = pt66
  else if (allocated(fsi)) then
    ! ebe3D of src/si_eigdrv3d.f90:333-342: r = vold - v - dtl (r - f)
    call lrhs3D(v, r, vm, x, y)
    do j = 1, ny
      do i = 1, nx
        do idof = 1, ndof
          r(idof,i,j) = vold(idof,i,j) - v(idof,i,j) - dtlh(i,j) * (r(idof,i,j) - fsi(idof,i,j))
        end do
      end do
    end do
    alfa = one
  else
    call lrhs3D(v, r, vm, x, y)
    call rbe3D(r, v, vold, dtlh)
    alfa = one
  end if
  call rhsBC3D(r, v, vm)
  tc1 = omp_get_wtime(); phase_t(1) = phase_t(1) + (tc1 - tc0); tc0 = tc1

  if (form) then
    call lhs1f3D(matx, Ah, Dh, Dhi, Vh11, ABhi, dtlh, v)
    if (xper) then
      call cpenta2p(nx, ny, ndof, matx, r, perx, per2x, 0)
    else
      ! the transposes run in parallel, as src/impdrv3d.f90:257-304 does
      !$omp parallel do private(i, idof, jdof)
      do j = 1, ny
        do i = 1, nx
          do idof = 1, ndof
            do jdof = 1, ndof
              lmatx(:,idof,jdof,j,i) = matx(:,idof,jdof,i,j)
            end do
          end do
        end do
      end do
      call cpenta1bc(ny, nx, ndof, lmatx, lr, 0)
    end if
    tc1 = omp_get_wtime(); phase_t(5) = phase_t(5) + (tc1 - tc0); tc0 = tc1
  end if
  if (xper) then
    call cpenta2p(nx, ny, ndof, matx, r, perx, per2x, 1)
  else
    !$omp parallel do private(i)
    do j = 1, ny
      do i = 1, nx
        lr(:,j,i) = r(:,i,j)
      end do
    end do
    call cpenta1bc(ny, nx, ndof, lmatx, lr, 1)
    !$omp parallel do private(i)
    do j = 1, ny
      do i = 1, nx
        r(:,i,j) = lr(:,j,i)
      end do
    end do
  end if
  tc1 = omp_get_wtime(); phase_t(2) = phase_t(2) + (tc1 - tc0); tc0 = tc1

  if (form) then
    call lhs2f3D(maty, Bh, Dh, Dhi, Vh22, ABhi, dtlh, v)
    if (yper) then
      call cpenta1p(nx, ny, ndof, maty, r, pery, per2y, 0)
    else
      call cpenta1bc(nx, ny, ndof, maty, r, 0)
    end if
    tc1 = omp_get_wtime(); phase_t(5) = phase_t(5) + (tc1 - tc0); tc0 = tc1
  end if
  if (yper) then
    call cpenta1p(nx, ny, ndof, maty, r, pery, per2y, 1)
  else
    call cpenta1bc(nx, ny, ndof, maty, r, 1)
  end if
  ! inviscid: rotate back to global coordinates on the body
  ! (src/impdrv3d.f90:333-336)
  if ((.not. Navier) .and. (.not. yper)) call bcfix(nx, ny, ndof, bnb, r)
  tc1 = omp_get_wtime(); phase_t(3) = phase_t(3) + (tc1 - tc0); tc0 = tc1

  ! the update in parallel, as src/impdrv3d.f90:340-348
  !$omp parallel do private(i)
  do j = 1, ny
    do i = 1, nx
      v(:,i,j) = v(:,i,j) + r(:,i,j)
      dv(:,i,j) = r(:,i,j)
    end do
  end do
  call itrBC3D(v, vm)
  tc1 = omp_get_wtime(); phase_t(4) = phase_t(4) + (tc1 - tc0); tc0 = tc1
  end do
  lstep_h = lstep_h + 1
  deallocate(vold, r)
end subroutine ref_step

! the &extras namelist's wall extent is:ie (src/lns3d.f90:42, 73) and
! updateLHS (src/global.f90:33)
subroutine ref_set_extras(is_l, ie_l, update) bind(C, name="ref_set_extras")
  use iso_c_binding
  use global
  implicit none
  integer(c_int), value :: is_l, ie_l, update
  is = is_l; ie = ie_l
  updateLHS = update /= 0
end subroutine ref_set_extras

! the phase times of the last ref_step (seconds; see phase_t)
subroutine ref_phase_times(t) bind(C, name="ref_phase_times")
  use iso_c_binding
  use refdrv_state
  implicit none
  real(c_double), intent(out) :: t(5)
  t = phase_t
end subroutine ref_phase_times

! OpenMP threads of the harness's parallel regions
subroutine ref_set_threads(n) bind(C, name="ref_set_threads")
  use iso_c_binding
  use omp_lib, only: omp_set_num_threads
  implicit none
  integer(c_int), value :: n
  call omp_set_num_threads(n)
end subroutine ref_set_threads

! the forcing f of the shift-invert driver's pseudo-time steps (set: the
! steps use ebe3D; clear: rbe3D)
subroutine ref_set_forcing(f, on) bind(C, name="ref_set_forcing")
  use iso_c_binding
  use global
  use refdrv_state
  implicit none
  complex(c_double_complex), intent(in) :: f(ndof,nx,ny)
  integer(c_int), value :: on
  if (allocated(fsi)) deallocate(fsi)
  if (on /= 0) then
    allocate(fsi(ndof,nx,ny))
    fsi = f
  end if
end subroutine ref_set_forcing

! the shift-invert driver's residual norm after a step
! (src/si_eigdrv3d_new.f90:359-374): r = rhsBC3D(ebe3D(lrhs3D(v))) with
! vold, and sqrt(sum |r|^2 / dtl^2)
subroutine ref_forced_residual(v, vold, r, nrm) bind(C, name="ref_forced_residual")
  use iso_c_binding
  use global
  use local2d
  use local3d
  use refdrv_state
  implicit none
  complex(c_double_complex), intent(in) :: v(ndof,nx,ny), vold(ndof,nx,ny)
  complex(c_double_complex), intent(out) :: r(ndof,nx,ny)
  real(c_double), intent(out) :: nrm
  complex, allocatable :: vl(:,:,:)
  integer :: i, j, idof
  allocate(vl(ndof,nx,ny))
  vl = v
  call lrhs3D(vl, r, vm, x, y)
  do j = 1, ny
    do i = 1, nx
      do idof = 1, ndof
        r(idof,i,j) = vold(idof,i,j) - vl(idof,i,j) - dtlh(i,j) * (r(idof,i,j) - fsi(idof,i,j))
      end do
    end do
  end do
  call rhsBC3D(r, vl, vm)
  nrm = 0.0
  do j = 1, ny
    do i = 1, nx
      do idof = 1, ndof
        nrm = nrm + real(r(idof,i,j) * conjg(r(idof,i,j))) / dtlh(i,j)**2
      end do
    end do
  end do
  nrm = sqrt(nrm)
  deallocate(vl)
end subroutine ref_forced_residual

! corrector iterations per step (the deck's niter, src/impdrv3d.f90:173)
subroutine ref_set_niter(n) bind(C, name="ref_set_niter")
  use iso_c_binding
  use global
  implicit none
  integer(c_int), value :: n
  niter = n
end subroutine ref_set_niter

! flush the Fortran unit 6 buffer (the reference's routines write progress
! lines to it)
subroutine ref_flush() bind(C, name="ref_flush")
  flush(6)
end subroutine ref_flush
