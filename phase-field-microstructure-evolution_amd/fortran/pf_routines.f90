! Drop-in replacements for the reference's transform and setup routines, with
! the reference's own signatures, for a maintainer who keeps src/var_diff.f90
! and links these instead of src/f_trans.f90, src/inv_trans.f90,
! src/k_space.f90 and src/ran_2.f90 (single process: ist = fst = (1,1,1),
! ien = (Nx,Ny,Nz), fen = (Nz,Ny,Nx/2+1), the P3DFFT STRIDE1 Z-pencil).
! The transforms run on the GPU through libpf.so with host buffers.

! src/f_trans.f90:1-15: forward R2C, then / float(Nx*Ny*Nz)
subroutine f_trans(in, out, Nx, Ny, Nz, ist, ien, fst, fen)
  use iso_c_binding
  use pf_capi, only: pf_f_trans_z, pf_check
  implicit none
  integer Nx, Ny, Nz, ist(3), ien(3), fst(3), fen(3)
  real*8 in(ist(1):ien(1), ist(2):ien(2), ist(3):ien(3))
  double complex out(fst(1):fen(1), fst(2):fen(2), fst(3):fen(3))
  call pf_check(pf_f_trans_z(in, out, int(Nx, c_int), int(Ny, c_int), int(Nz, c_int)), 'f_trans')
end subroutine

! src/inv_trans.f90:1-13: inverse C2R, unscaled
subroutine inv_trans(in, out, ist, ien, fst, fen)
  use iso_c_binding
  use pf_capi, only: pf_inv_trans_z, pf_check
  implicit none
  integer ist(3), ien(3), fst(3), fen(3)
  double complex in(fst(1):fen(1), fst(2):fen(2), fst(3):fen(3))
  real*8 out(ist(1):ien(1), ist(2):ien(2), ist(3):ien(3))
  call pf_check(pf_inv_trans_z(in, out, int(ien(1) - ist(1) + 1, c_int), int(ien(2) - ist(2) + 1, c_int), &
                               int(ien(3) - ist(3) + 1, c_int)), 'inv_trans')
end subroutine

! src/k_space.f90:4-67: kf, kf_sq, kf_4 (3, z, y, x), bit-exact
subroutine k_space(Nx, Ny, Nz, fst, fen, kf, kf_sq, kf_4)
  use iso_c_binding
  use pf_capi, only: pf_k_space_z, pf_check
  implicit none
  integer Nx, Ny, Nz, fst(3), fen(3)
  double complex kf(3, fst(1):fen(1), fst(2):fen(2), fst(3):fen(3))
  double complex kf_sq(3, fst(1):fen(1), fst(2):fen(2), fst(3):fen(3))
  double complex kf_4(3, fst(1):fen(1), fst(2):fen(2), fst(3):fen(3))
  call pf_check(pf_k_space_z(int(Nx, c_int), int(Ny, c_int), int(Nz, c_int), kf, kf_sq, kf_4), 'k_space')
end subroutine

! src/ran_2.f90:1-40: the same generator, state kept by pf_capi
real*8 function ran_2(idum)
  use pf_capi, only: pf_ran_2 => ran_2
  implicit none
  integer idum
  ran_2 = pf_ran_2(idum)
end function
