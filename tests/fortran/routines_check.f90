! Calls the reference-signature routines of fortran/pf_routines.f90 as
! external procedures, the way src/var_diff.f90 does, on a 16 x 12 x 8 grid and
! prints what a test compares with numpy and the oracle:
!   line 1: max |inv_trans(f_trans(x)) - x| (f_trans scales by 1/(Nx Ny Nz), inv_trans does not)
!   line 2: f_trans(x) at (kz,ky,kx) = (2,3,4) (1-based), real and imaginary
!   line 3: aimag(kf(1..3)), dble(kf_sq(1..3)), dble(kf_4(1..3)) at (kz,ky,kx) = (5,7,9)
!   line 4: the first three ran_2(-7) values
program routines_check
  implicit none
  integer, parameter :: nx = 16, ny = 12, nz = 8
  integer :: ist(3), ien(3), fst(3), fen(3), i, j, k, idum
  real*8 :: x(nx, ny, nz), y(nx, ny, nz)
  double complex :: h(nz, ny, nx / 2 + 1)
  double complex, allocatable :: kf(:, :, :, :), kf_sq(:, :, :, :), kf_4(:, :, :, :)
  real*8, external :: ran_2
  ist = 1; ien = [nx, ny, nz]; fst = 1; fen = [nz, ny, nx / 2 + 1]
  do k = 1, nz
    do j = 1, ny
      do i = 1, nx
        x(i, j, k) = sin(0.3d0 * i + 0.7d0 * j * j) + cos(1.1d0 * k * i) * 0.5d0
      end do
    end do
  end do
  call f_trans(x, h, nx, ny, nz, ist, ien, fst, fen)
  call inv_trans(h, y, ist, ien, fst, fen)
  write (*, '(ES25.16E3)') maxval(abs(y - x))
  write (*, '(2ES25.16E3)') dble(h(2, 3, 4)), aimag(h(2, 3, 4))
  allocate (kf(3, nz, ny, nx / 2 + 1), kf_sq(3, nz, ny, nx / 2 + 1), kf_4(3, nz, ny, nx / 2 + 1))
  call k_space(nx, ny, nz, fst, fen, kf, kf_sq, kf_4)
  write (*, '(9ES25.16E3)') aimag(kf(:, 5, 7, 9)), dble(kf_sq(:, 5, 7, 9)), dble(kf_4(:, 5, 7, 9))
  idum = -7
  write (*, '(3ES25.16E3)') ran_2(idum), ran_2(idum), ran_2(idum)
end program
