! Writes one file of each reference data format with the Fortran runtime's
! own unformatted sequential records, the record structure of the
! reference's writers and readers:
!   restart.R   src/impdrv3d.f90:393-399 (header, v, vold)
!   mean.dat    src/genmean.f90:28-33    (header, vm)
!   grid.dat    src/gengrid.f90:17-32    (nx, ny, nz; x, y, z)
!   metric.dat  src/gengrid.f90:76-80    (ten metric arrays)
! Values follow closed forms that tests/test_files.py restates.
program fortran_files
  implicit none
  integer, parameter :: nx = 7, ny = 5, nz = 1, ndof = 5
  integer :: i, j, k, e, lstep
  real(8) :: time, Re, Ma, Pr, gamma, cv
  complex(8) :: v(ndof, nx, ny), vold(ndof, nx, ny)
  real(8) :: vm(ndof, nx, ny), x(nx, ny), y(nx, ny), z(nx, ny)
  real(8) :: mt(nx, ny, 10)

  lstep = 1234; time = 5.25d0; Re = 2400d0; Ma = 0.1d0; Pr = 0.72d0; gamma = 1.4d0; cv = 716.5d0
  do j = 1, ny
    do i = 1, nx
      do e = 1, ndof
        v(e, i, j) = cmplx(e + 10 * i + 100 * j, -(e * i + j) / 8d0, kind=8)
        vold(e, i, j) = cmplx(-0.5d0 * e, i + 0.25d0 * j, kind=8)
        vm(e, i, j) = e * 1000d0 + i * 10d0 + j + 0.5d0
      end do
      x(i, j) = 1.5d0 * i + 0.25d0 * j
      y(i, j) = 0.75d0 * j - 0.125d0 * i
      z(i, j) = 0d0
      do k = 1, 10
        mt(i, j, k) = 100d0 * k + i + j / 16d0
      end do
    end do
  end do

  open(unit=10, file='restart.R', form='unformatted', status='unknown')
  write(10) lstep, time, nx, ny, nz, ndof, Re, Ma, Pr, gamma, cv
  write(10) v
  write(10) vold
  close(10)

  open(unit=10, file='mean.dat', form='unformatted', status='unknown')
  write(10) lstep, time, nx, ny, nz, ndof, Re, Ma, Pr, gamma, cv
  write(10) vm
  close(10)

  open(unit=10, file='grid.dat', form='unformatted', status='unknown')
  write(10) nx, ny, nz
  write(10) (((x(i, j), i = 1, nx), j = 1, ny), k = 1, nz), &
            (((y(i, j), i = 1, nx), j = 1, ny), k = 1, nz), &
            (((z(i, j), i = 1, nx), j = 1, ny), k = 1, nz)
  close(10)

  open(unit=10, file='metric.dat', form='unformatted', status='unknown')
  write(10) mt(:, :, 1), mt(:, :, 2), mt(:, :, 3), mt(:, :, 4), mt(:, :, 5), &
            mt(:, :, 6), mt(:, :, 7), mt(:, :, 8), mt(:, :, 9), mt(:, :, 10)
  close(10)
end program fortran_files
