! Reads a data_save.NNNN file the way a Fortran post-processor does (plain
! unformatted sequential reads of ist, ien and then phi, con; cf.
! src/paraview.f90:177-183) and prints sums of the arrays, so that a test can
! check the driver's record layout with a Fortran runtime's own reader.
! Usage: read_checkpoint <file> <nx> <ny> <nz> <var>
program read_checkpoint
  implicit none
  character(len=512) :: path, arg
  integer :: nx, ny, nz, var, ist(3), ien(3), u
  real(8), allocatable :: phi(:, :, :, :), con(:, :, :)
  call get_command_argument(1, path)
  call get_command_argument(2, arg); read (arg, *) nx
  call get_command_argument(3, arg); read (arg, *) ny
  call get_command_argument(4, arg); read (arg, *) nz
  call get_command_argument(5, arg); read (arg, *) var
  allocate (phi(var, nx, ny, nz), con(nx, ny, nz))
  open (newunit=u, file=trim(path), status='old', form='unformatted', action='read')
  read (u) ist, ien
  read (u) phi(1:var, 1:nx, ist(2):ien(2), ist(3):ien(3)), con(1:nx, ist(2):ien(2), ist(3):ien(3))
  close (u)
  write (*, '(6I8)') ist, ien
  write (*, '(3ES25.16)') sum(phi), sum(con), sum(phi(var, :, :, :) * con)
end program
