! paraview_hip.x <input_dir> <output_dir> <threshold>
!
! Drop-in for the reference's post-processor (src/paraview.f90): for every
! input_dir/step_NNNNNNN directory, read data_extract_mult.in and the
! data_save.NNNN files, evaluate the per-voxel quantities on the GPU
! (pf_analyze_state), write phi_sq_, con_, var_num_, phi_3_, grad_phi_3_ and
! boundary_<step>.vtk (ASCII STRUCTURED_POINTS, the reference's header), and
! print one line per step: step, precipitate boundary voxels, boundary voxels,
! precipitate fraction (format (A, 2(I8), F6.2)).  Single process.
program paraview_hip
  use iso_c_binding
  use iso_fortran_env, only: int64
  use pf_capi
  use pf_records
  implicit none
  character(len=512) :: input_dir, output_dir, arg, step_dir, listfile
  character(len=64) :: line
  character(len=64), allocatable :: steps(:)
  real(8) :: threshold, frac
  integer :: nsteps, u, ios, s, nx, ny, nz, np, var, dims(2), ierr
  real(8), allocatable, target :: phi(:, :, :, :), con(:, :, :)
  real(8), allocatable :: phi_sq(:), phi_3(:), grad3(:), bnd(:)
  integer(c_int), allocatable :: var_num(:)
  integer(c_long) :: counts(2)

  call get_command_argument(1, input_dir)
  call get_command_argument(2, output_dir)
  call get_command_argument(3, arg)
  read (arg, *) threshold

  listfile = trim(output_dir)//'/.steps_list'
  call execute_command_line('ls '//trim(input_dir)//' | grep "^step_" | cut -d"_" -f2 > '// &
                            trim(listfile), wait=.true., exitstat=ierr)
  if (ierr /= 0) then
    write (0, *) 'cannot list step directories of ', trim(input_dir)
    error stop 1
  end if
  open (newunit=u, file=trim(listfile), status='old', action='read')
  nsteps = 0
  do
    read (u, '(A)', iostat=ios) line
    if (ios /= 0) exit
    nsteps = nsteps + 1
  end do
  allocate (steps(nsteps))
  rewind (u)
  do s = 1, nsteps
    read (u, '(A)') steps(s)
    steps(s) = adjustl(steps(s))
  end do
  close (u, status='delete')

  write (*, '(A)') "# step, precipitate_voxels, boundary_voxels, precipitate_fraction"
  do s = 1, nsteps
    step_dir = trim(input_dir)//'/step_'//trim(steps(s))//'/'
    open (newunit=u, file=trim(step_dir)//'data_extract_mult.in', status='old', action='read')
    read (u, *) nx, ny, nz, np, var
    read (u, *) dims(1), dims(2)
    close (u)
    allocate (phi(var, nx, ny, nz), con(nx, ny, nz))
    call read_slabs(trim(step_dir), dims, nx, ny, nz, var, phi, con)
    allocate (phi_sq(nx * ny * nz), phi_3(nx * ny * nz), grad3(nx * ny * nz), bnd(nx * ny * nz))
    allocate (var_num(nx * ny * nz))
    call pf_check(pf_analyze_state(phi, int(nx, c_int), int(ny, c_int), int(nz, c_int), int(var, c_int), &
                                   threshold, phi_sq, phi_3, grad3, bnd, var_num, counts), 'pf_analyze_state')
    call write_vtk('phi_sq_', steps(s), nx, ny, nz, real_vals=phi_sq)
    call write_vtk('con_', steps(s), nx, ny, nz, real_vals=reshape(con, [nx * ny * nz]))
    call write_vtk('var_num_', steps(s), nx, ny, nz, int_vals=var_num)
    call write_vtk('phi_3_', steps(s), nx, ny, nz, real_vals=phi_3)
    call write_vtk('grad_phi_3_', steps(s), nx, ny, nz, real_vals=grad3)
    call write_vtk('boundary_', steps(s), nx, ny, nz, real_vals=bnd)
    if (counts(2) > 0) then
      frac = real(counts(1), 8) / real(counts(2), 8)
    else
      frac = 0.0d0
    end if
    write (*, '(A, 2(I8), F6.2)') trim(steps(s)), counts(1), counts(2), frac
    deallocate (phi, con, phi_sq, phi_3, grad3, bnd, var_num)
  end do

contains

  ! data_save.NNNN, NNNN = 1 .. dims(1)*dims(2), each holding ist(3), ien(3) and
  ! then phi(1:var, 1:nx, ist(2):ien(2), ist(3):ien(3)), con(1:nx, ...).
  subroutine read_slabs(dir, dims, nx, ny, nz, var, phi, con)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: dims(2), nx, ny, nz, var
    real(8), intent(inout) :: phi(var, nx, ny, nz), con(nx, ny, nz)
    real(8), allocatable, target :: lphi(:, :, :, :), lcon(:, :, :)
    integer, target :: b(6)
    integer :: f, u, ny_l, nz_l
    integer(int64) :: pos
    logical :: ok
    character(len=4) :: num
    do f = 1, dims(1) * dims(2)
      write (num, '(i4.4)') f
      open (newunit=u, file=dir//'data_save.'//num, status='old', access='stream', form='unformatted', &
            action='read')
      pos = 1
      call get_record(u, pos, c_loc(b), 24_int64, c_null_ptr, 0_int64, ok)
      if (.not. ok .or. b(1) /= 1 .or. b(4) /= nx .or. b(2) < 1 .or. b(5) > ny .or. b(3) < 1 .or. &
          b(6) > nz .or. b(5) < b(2) .or. b(6) < b(3)) then
        write (0, *) dir//'data_save.'//num, ': bad bounds record'
        error stop 1
      end if
      ny_l = b(5) - b(2) + 1
      nz_l = b(6) - b(3) + 1
      allocate (lphi(var, nx, ny_l, nz_l), lcon(nx, ny_l, nz_l))
      call get_record(u, pos, c_loc(lphi), 8_int64 * size(lphi, kind=int64), c_loc(lcon), &
                      8_int64 * size(lcon, kind=int64), ok)
      if (.not. ok) then
        write (0, *) dir//'data_save.'//num, ': phi/con record is truncated or has the wrong length'
        error stop 1
      end if
      close (u)
      phi(:, :, b(2):b(5), b(3):b(6)) = lphi
      con(:, b(2):b(5), b(3):b(6)) = lcon
      deallocate (lphi, lcon)
    end do
  end subroutine

  subroutine write_vtk(prefix, step, nx, ny, nz, real_vals, int_vals)
    character(len=*), intent(in) :: prefix, step
    integer, intent(in) :: nx, ny, nz
    real(8), intent(in), optional :: real_vals(:)
    integer(c_int), intent(in), optional :: int_vals(:)
    integer :: u
    open (newunit=u, file=trim(output_dir)//'/'//prefix//trim(step)//'.vtk', status='replace')
    write (u, '(a26)') "# vtk DataFile Version 3.0"
    write (u, '(a11)') "Sample Data"
    write (u, '(a5)') "ASCII"
    write (u, '(a25)') "DATASET STRUCTURED_POINTS"
    write (u, '(a10,2x,3i6)') "DIMENSIONS", nx, ny, nz
    write (u, '(a24)') "ASPECT_RATIO 1.0 1.0 1.0"
    write (u, '(a18)') "ORIGIN 0.0 0.0 0.0"
    write (u, '(a10,2x,i12)') "POINT_DATA", nx * ny * nz
    write (u, '(a22)') "SCALARS Field double 1"
    write (u, '(a20)') "LOOKUP_TABLE default"
    if (present(real_vals)) write (u, '(4ES25.16E3)') real_vals
    if (present(int_vals)) write (u, '(16I5)') int_vals
    close (u)
  end subroutine

end program paraview_hip
