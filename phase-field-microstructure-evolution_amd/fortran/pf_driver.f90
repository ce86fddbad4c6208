! Host-side pieces of the var_diff drop-in: input.txt parsing, the gamma matrix,
! initial conditions, precipitate insertion and the output files.  Single
! process, so the reference's pencil bounds are ist = (1,1,1), ien = (Nx,Ny,Nz).
!
! Behaviour follows src/var_diff.f90 (line numbers cited per routine); the code
! is this project's own.
module pf_driver
  use iso_c_binding
  use iso_fortran_env, only: int64
  use pf_capi
  use pf_records
  implicit none
  private

  ! Everything input.txt carries (src/var_diff.f90:106-121).
  type, public :: run_config
    integer :: nrun, ppt_int, ppt_rad(3), n_step, ifreq
    integer :: i_ppt, j_ppt, k_ppt, iseed
    integer :: grn_rad, num_ppt
    real(8) :: c_mat, c_ppt
    integer :: nx, ny, nz, var
    real(8) :: dx, dy, dz, t_step
    real(8) :: grad_coeff_phi, grad_coeff_c, d_s, d_l, d_gb, mob_phi
    real(8) :: sigma_1, sigma_2
    real(8) :: am, ap, con_0_mat, con_0_ppt, gb_force
    integer :: ndim, dims(2)
    character(len=10) :: start
  end type

  public :: read_input, params_of, build_gamma, write_gamma
  public :: ic_polycryst, ic_plane, ic_circle, read_restart, insert_precipitates
  public :: place_precipitates, state_extrema
  public :: write_checkpoint, write_summary_row, write_timing_row, write_error_header, write_error_row

contains

  ! List-directed reads, one record per read, as src/var_diff.f90:104-121.
  ! nprocs = 1, so ndim = 1 gives dims = (1, 1).
  subroutine read_input(path, c)
    character(len=*), intent(in) :: path
    type(run_config), intent(out) :: c
    integer :: u
    open (newunit=u, file=path, status='old', action='read')
    read (u, *) c%nrun, c%ppt_int
    read (u, *) c%ppt_rad(1), c%ppt_rad(2), c%ppt_rad(3), c%n_step, c%ifreq
    read (u, *) c%i_ppt, c%j_ppt, c%k_ppt, c%iseed
    read (u, *) c%grn_rad, c%num_ppt, c%c_mat, c%c_ppt
    read (u, *) c%nx, c%ny, c%nz, c%var, c%dx, c%dy, c%dz, c%t_step
    read (u, *) c%grad_coeff_phi, c%grad_coeff_c, c%d_s, c%d_l, c%d_gb, c%mob_phi
    read (u, *) c%sigma_1, c%sigma_2
    read (u, *) c%am, c%ap, c%con_0_mat, c%con_0_ppt, c%gb_force
    read (u, *) c%ndim
    if (c%ndim == 1) then
      c%dims(1) = 1
      c%dims(2) = 1
    else if (c%ndim == 2) then
      read (u, *) c%dims(1), c%dims(2)
    end if
    read (u, *) c%start
    close (u)
  end subroutine

  function params_of(c) result(p)
    type(run_config), intent(in) :: c
    type(pf_params) :: p
    p%t_step = c%t_step
    p%grad_coeff_phi = c%grad_coeff_phi
    p%grad_coeff_c = c%grad_coeff_c
    p%D_S = c%d_s
    p%D_L = c%d_l
    p%D_gb = c%d_gb
    p%mob_phi = c%mob_phi
    p%sigma_1 = c%sigma_1
    p%sigma_2 = c%sigma_2
    p%am = c%am
    p%ap = c%ap
    p%con_0_mat = c%con_0_mat
    p%con_0_ppt = c%con_0_ppt
  end function

  ! gamma(i,j): sigma_1 between matrix grains, sigma_2 between a grain and the
  ! last (precipitate) order parameter, 0 on the diagonal (src/var_diff.f90:206-218).
  subroutine build_gamma(var, s1, s2, g)
    integer, intent(in) :: var
    real(8), intent(in) :: s1, s2
    real(8), intent(out) :: g(var, var)
    integer :: i, j
    g = 0.0d0
    do j = 1, var
      do i = 1, var
        if (i /= j .and. i <= var - 1 .and. j <= var - 1) g(i, j) = s1
        if (j == var .and. i < var) g(i, j) = s2
        if (i == var .and. j < var) g(i, j) = s2
      end do
    end do
  end subroutine

  subroutine write_gamma(u, var, g)
    integer, intent(in) :: u, var
    real(8), intent(in) :: g(var, var)
    integer :: i, j
    do i = 1, var
      write (u, *) i, (g(i, j), j=1, var)
    end do
    flush (u)
  end subroutine

  ! 'polycryst': Nx*Ny*Nz lines "ii jj kk ivar" set phi(ivar,ii,jj,kk) = 1
  ! (src/var_diff.f90:232-252).  The reference opens a fixed absolute path; the
  ! path is an argument here.
  subroutine ic_polycryst(c, fname, phi)
    type(run_config), intent(in) :: c
    character(len=*), intent(in) :: fname
    real(8), intent(inout) :: phi(c%var, c%nx, c%ny, c%nz)
    integer :: u, n, ii, jj, kk, ivar
    open (newunit=u, file=fname, status='old', action='read')
    do n = 1, c%nx * c%ny * c%nz
      read (u, *) ii, jj, kk, ivar
      if (ii >= 1 .and. ii <= c%nx .and. jj >= 1 .and. jj <= c%ny .and. &
          kk >= 1 .and. kk <= c%nz) then
        if (ivar < 1 .or. ivar > c%var) then
          write (0, *) trim(fname), ': grain index ', ivar, ' outside 1..', c%var
          error stop 1
        end if
        phi(ivar, ii, jj, kk) = 1.d0
      end if
    end do
    close (u)
  end subroutine

  ! Marks an ellipsoid (semi-axes ppt_rad, centre ci,cj,ck) as phase 3, clearing
  ! phases 1 and 2; the extent test is done in default REAL as in the reference.
  subroutine mark_ellipsoid(c, ci, cj, ck, phi)
    type(run_config), intent(in) :: c
    integer, intent(in) :: ci, cj, ck
    real(8), intent(inout) :: phi(c%var, c%nx, c%ny, c%nz)
    integer :: i, j, k, ii, jj, kk
    real(8) :: t1, t2, t3, s
    if (c%var < 3) then
      write (0, *) 'this initial condition writes order parameters 1..3 and needs var >= 3'
      error stop 1
    end if
    do k = -c%ppt_rad(3), c%ppt_rad(3)
      do j = -c%ppt_rad(2), c%ppt_rad(2)
        do i = -c%ppt_rad(1), c%ppt_rad(1)
          t1 = (real(i) / real(c%ppt_rad(1)))**2
          t2 = (real(j) / real(c%ppt_rad(2)))**2
          t3 = (real(k) / real(c%ppt_rad(3)))**2
          s = t1 + t2 + t3
          if (s <= 1.0) then
            ii = i + ci
            jj = j + cj
            kk = k + ck
            if (ii >= 1 .and. ii <= c%nx .and. jj >= 1 .and. jj <= c%ny .and. &
                kk >= 1 .and. kk <= c%nz) then
              phi(3, ii, jj, kk) = 1.0
              phi(1, ii, jj, kk) = 0.d0
              phi(2, ii, jj, kk) = 0.d0
            end if
          end if
        end do
      end do
    end do
  end subroutine

  ! 'plane' (src/var_diff.f90:254-355): grain 2 in the band j1 <= y < j_ppt,
  ! grain 1 outside, and two ellipsoidal precipitates centred at j_ppt and at
  ! Ny - j_ppt + 1.
  subroutine ic_plane(c, phi)
    type(run_config), intent(inout) :: c
    real(8), intent(inout) :: phi(c%var, c%nx, c%ny, c%nz)
    integer :: j1, jj
    j1 = c%ny - c%j_ppt + 1
    write (*, *) 'j1=', j1
    phi(2, :, :, :) = 1.0
    do jj = 1, c%ny
      if (jj < j1) then
        phi(:, :, jj, :) = 0.0
        phi(1, :, jj, :) = 1.d0
      end if
    end do
    do jj = 1, c%ny
      if (jj >= c%j_ppt) then
        phi(:, :, jj, :) = 0.0
        phi(1, :, jj, :) = 1.d0
      end if
    end do
    call mark_ellipsoid(c, c%i_ppt, c%j_ppt, c%k_ppt, phi)
    c%j_ppt = c%ny - c%j_ppt + 1
    call mark_ellipsoid(c, c%i_ppt, c%j_ppt, c%k_ppt, phi)
  end subroutine

  ! 'circle' (src/var_diff.f90:357-449): grain 1 everywhere, a spherical grain 2
  ! of radius grn_rad at the centre, then num_ppt precipitates (phase 3, c =
  ! con_0_ppt) of squared radius sum(ppt_rad**2) at ran_2 positions, rejecting a
  ! position whose sphere already holds phase var > 0.
  subroutine ic_circle(c, phi, con, iseed)
    type(run_config), intent(in) :: c
    real(8), intent(inout) :: phi(c%var, c%nx, c%ny, c%nz), con(c%nx, c%ny, c%nz)
    integer, intent(inout) :: iseed
    integer :: i, j, k, ii, jj, kk, nppt, rad_ppt, radius, ix, jy, kz, cnt
    real(8) :: t1, t2, t3, s
    if (c%var < 3) then
      write (0, *) 'this initial condition writes order parameters 1..3 and needs var >= 3'
      error stop 1
    end if
    phi(1, :, :, :) = 1.0
    do k = -c%grn_rad, c%grn_rad
      do j = -c%grn_rad, c%grn_rad
        do i = -c%grn_rad, c%grn_rad
          t1 = (real(i) / real(c%grn_rad))**2
          t2 = (real(j) / real(c%grn_rad))**2
          t3 = (real(k) / real(c%grn_rad))**2
          s = t1 + t2 + t3
          if (s <= 1.0) then
            ii = i + c%nx / 2
            jj = j + c%ny / 2
            kk = k + c%nz / 2
            if (ii >= 1 .and. ii <= c%nx .and. jj >= 1 .and. jj <= c%ny .and. &
                kk >= 1 .and. kk <= c%nz) then
              phi(2, ii, jj, kk) = 1.d0
              phi(1, ii, jj, kk) = 0.d0
              phi(3, ii, jj, kk) = 0.d0
            end if
          end if
        end do
      end do
    end do
    nppt = 1
    rad_ppt = int((c%ppt_rad(1)**2 + c%ppt_rad(2)**2 + c%ppt_rad(3)**2) * 1.0)
    do while (nppt <= c%num_ppt)
      ix = int(ran_2(iseed) * c%nx + 1.0)
      jy = int(ran_2(iseed) * c%ny + 1.0)
      kz = int(ran_2(iseed) * c%nz + 1.0)
      cnt = 0
      do k = 1, c%nz
        do j = 1, c%ny
          do i = 1, c%nx
            radius = (i - ix)**2 + (j - jy)**2 + (k - kz)**2
            if (radius <= rad_ppt .and. phi(c%var, i, j, k) > 0.0) cnt = cnt + 1
          end do
        end do
      end do
      if (cnt > 0) cycle
      do k = 1, c%nz
        do j = 1, c%ny
          do i = 1, c%nx
            radius = (i - ix)**2 + (j - jy)**2 + (k - kz)**2
            if (radius <= rad_ppt) then
              phi(3, i, j, k) = 1.0
              phi(2, i, j, k) = 0.0
              phi(1, i, j, k) = 0.0
              con(i, j, k) = c%con_0_ppt
            end if
          end do
        end do
      end do
      nppt = nppt + 1
      write (*, *) 'nppt=', nppt
    end do
  end subroutine

  ! nrun /= 1: unformatted record "ist, ien" then "phi, con" (src/var_diff.f90:451-461).
  subroutine read_restart(fname, c, phi, con)
    character(len=*), intent(in) :: fname
    type(run_config), intent(in) :: c
    real(8), intent(out), target :: phi(c%var, c%nx, c%ny, c%nz), con(c%nx, c%ny, c%nz)
    integer, target :: b(6)
    integer :: u
    integer(int64) :: pos
    logical :: ok
    open (newunit=u, file=fname, status='old', access='stream', form='unformatted', action='read')
    pos = 1
    call get_record(u, pos, c_loc(b), 24_int64, c_null_ptr, 0_int64, ok)
    if (.not. ok) then
      write (0, *) 'restart file ', trim(fname), ': first record is not ist(3), ien(3)'
      error stop 1
    end if
    if (any(b(1:3) /= 1) .or. b(4) /= c%nx .or. b(5) /= c%ny .or. b(6) /= c%nz) then
      write (0, *) 'restart file ', trim(fname), ' holds bounds ', b, &
        ' for a ', c%nx, c%ny, c%nz, ' grid'
      error stop 1
    end if
    call get_record(u, pos, c_loc(phi), 8_int64 * size(phi, kind=int64), &
                    c_loc(con), 8_int64 * size(con, kind=int64), ok)
    if (.not. ok) then
      write (0, *) 'restart file ', trim(fname), ': phi/con record is truncated or has the wrong length'
      error stop 1
    end if
    close (u)
  end subroutine

  ! Precipitates at step == ppt_int (src/var_diff.f90:479-554): a ran_2 position
  ! is rejected when sum_{i<var} phi_i^2 > 0.8 there or when its sphere already
  ! holds phase var > 0; otherwise the sphere becomes phase var.  Rejected
  ! positions are redrawn and do not count towards num_ppt.  centres receives
  ! the accepted positions (3, num_ppt), in order.  The reference redraws
  ! without bound, so it never ends when no site qualifies (sharp grain
  ! boundaries, or no room left); here that stops with an error after
  ! kMaxInsertDraws positions.
  subroutine insert_precipitates(c, phi, iseed, centres)
    type(run_config), intent(in) :: c
    real(8), intent(inout) :: phi(c%var, c%nx, c%ny, c%nz)
    integer, intent(inout) :: iseed
    integer, intent(out) :: centres(3, max(c%num_ppt, 1))
    integer, parameter :: kMaxInsertDraws = 1000000
    integer :: i, j, k, ivar, nppt, rad_ppt, radius, ix, jy, kz, cnt, draws
    real(8) :: fsum, fi
    nppt = 1
    draws = 0
    rad_ppt = c%ppt_rad(1)**2 + c%ppt_rad(2)**2 + c%ppt_rad(3)**2
    do while (nppt <= c%num_ppt)
      do
        draws = draws + 1
        if (draws > kMaxInsertDraws) then
          write (0, '(A, I0, A, I0, A)') 'precipitate insertion: no admissible site for precipitate ', nppt, &
            ' after ', kMaxInsertDraws, ' drawn positions'
          error stop 1
        end if
        ix = int(ran_2(iseed) * c%nx + 1.0)
        jy = int(ran_2(iseed) * c%ny + 1.0)
        kz = int(ran_2(iseed) * c%nz + 1.0)
        fsum = 0.0
        if (ix >= 1 .and. ix <= c%nx .and. jy >= 1 .and. jy <= c%ny .and. &
            kz >= 1 .and. kz <= c%nz) then
          do ivar = 1, c%var - 1
            fi = phi(ivar, ix, jy, kz)
            fsum = fsum + fi**2
          end do
        end if
        if (fsum > 0.8) cycle
        cnt = 0
        do k = 1, c%nz
          do j = 1, c%ny
            do i = 1, c%nx
              radius = (i - ix)**2 + (j - jy)**2 + (k - kz)**2
              if (radius <= rad_ppt .and. phi(c%var, i, j, k) > 0.0) cnt = cnt + 1
            end do
          end do
        end do
        if (cnt > 0) cycle
        exit
      end do
      centres(:, nppt) = [ix, jy, kz]
      call place_precipitates(c, phi, centres(:, nppt:nppt), 1)
      nppt = nppt + 1
      write (*, *) 'nppt=', nppt
    end do
  end subroutine

  ! The spheres of radius^2 <= sum ppt_rad^2 around n centres become phase var
  ! (the assignment of src/var_diff.f90:531-546).
  subroutine place_precipitates(c, phi, centres, n)
    type(run_config), intent(in) :: c
    real(8), intent(inout) :: phi(c%var, c%nx, c%ny, c%nz)
    integer, intent(in) :: n, centres(3, n)
    integer :: i, j, k, ivar, m, rad_ppt, radius
    rad_ppt = c%ppt_rad(1)**2 + c%ppt_rad(2)**2 + c%ppt_rad(3)**2
    do m = 1, n
      do k = 1, c%nz
        do j = 1, c%ny
          do i = 1, c%nx
            radius = (i - centres(1, m))**2 + (j - centres(2, m))**2 + (k - centres(3, m))**2
            if (radius <= rad_ppt) then
              phi(c%var, i, j, k) = 1.0
              do ivar = 1, c%var - 1
                phi(ivar, i, j, k) = 0.0
              end do
            end if
          end do
        end do
      end do
    end do
  end subroutine

  ! The extrema the reference computes at the end of a step and prints in the
  ! next reported row (src/var_diff.f90:1076-1094), with its initial values and
  ! comparisons: ext = (c_max, c_min, phi_min, phi_3_max).
  subroutine state_extrema(c, phi, con, ext)
    type(run_config), intent(in) :: c
    real(8), intent(in) :: phi(c%var, c%nx, c%ny, c%nz), con(c%nx, c%ny, c%nz)
    real(8), intent(out) :: ext(4)
    real(8) :: phi_min, phi_3, c_max, c_min, phi_sq
    integer :: i, j, k, ivar
    phi_min = 1.0
    phi_3 = 0.0
    c_max = 0.0
    c_min = 1.0
    do k = 1, c%nz
      do j = 1, c%ny
        do i = 1, c%nx
          if (phi(c%var, i, j, k) > phi_3) phi_3 = phi(c%var, i, j, k)
          if (con(i, j, k) <= c_min) c_min = con(i, j, k)
          if (con(i, j, k) > c_max) c_max = con(i, j, k)
          phi_sq = 0.0
          do ivar = 1, c%var
            phi_sq = phi_sq + phi(ivar, i, j, k)**2
          end do
          if (phi_sq < phi_min) phi_min = phi_sq
        end do
      end do
    end do
    ext = [c_max, c_min, phi_min, phi_3]
  end subroutine

  ! output_dir/step_NNNNNNN/{data_extract_mult.in, data_save.0001}
  ! (src/var_diff.f90:1120-1142).
  subroutine write_checkpoint(outdir, c, step, phi, con)
    character(len=*), intent(in) :: outdir
    type(run_config), intent(in) :: c
    integer, intent(in) :: step
    real(8), intent(in), target :: phi(c%var, c%nx, c%ny, c%nz), con(c%nx, c%ny, c%nz)
    character(len=512) :: dname
    integer, target :: b(6)
    integer :: u, ierr
    write (dname, '(A, a5, i7.7)') trim(outdir), 'step_', step
    call execute_command_line('mkdir -p '//trim(dname), exitstat=ierr)
    open (newunit=u, file=trim(dname)//'/data_extract_mult.in')
    write (u, *) c%nx, c%ny, c%nz, 1, c%var
    write (u, *) c%dims
    close (u)
    b = [1, 1, 1, c%nx, c%ny, c%nz]
    open (newunit=u, file=trim(dname)//'/data_save.0001', access='stream', form='unformatted', &
          status='replace')
    call put_record(u, c_loc(b), 24_int64, c_null_ptr, 0_int64)
    call put_record(u, c_loc(phi), 8_int64 * size(phi, kind=int64), &
                    c_loc(con), 8_int64 * size(con, kind=int64))
    flush (u)
    close (u)
  end subroutine

  ! summary.dat row, format (i6, 8e15.7) (src/var_diff.f90:1070-1071, 1158).
  subroutine write_summary_row(u, d)
    integer, intent(in) :: u
    type(pf_diag), intent(in) :: d
    write (u, '(i6, 8e15.7)') d%step, d%f_int, d%f_ch, d%f_eta, d%e_tot, d%c_max, d%c_min, &
      d%phi_min, d%phi_3_max
    flush (u)
  end subroutine

  ! timing_data.dat row, format (I8, 16(A, F8.6)) (src/var_diff.f90:1150-1153).
  subroutine write_timing_row(u, step, loop_s, step_s, run_s)
    integer, intent(in) :: u, step
    real(8), intent(in) :: loop_s(14), step_s, run_s
    integer :: i
    write (u, '(I8, 16(A, F8.6))') step, (' ', loop_s(i), i=1, 14), ' ', step_s, ' ', run_s
  end subroutine

  ! error_vs_fp64.dat: one row per checkpoint with, per field (c, eta_1 ..
  ! eta_var), the largest absolute difference from the FP64 run and the root mean
  ! square difference over the grid.
  subroutine write_error_header(u, var)
    integer, intent(in) :: u, var
    integer :: i
    write (u, '(A)', advance='no') '# step, max|c - c64|, rms(c - c64)'
    do i = 1, var
      write (u, '(A, I0, A, I0, A)', advance='no') ', max|eta', i, ' - eta64|, rms(eta', i, ' - eta64)'
    end do
    write (u, '(A)') ''
    flush (u)
  end subroutine

  subroutine write_error_row(u, step, c, phi, con, phi64, con64)
    integer, intent(in) :: u, step
    type(run_config), intent(in) :: c
    real(8), intent(in) :: phi(c%var, c%nx, c%ny, c%nz), con(c%nx, c%ny, c%nz)
    real(8), intent(in) :: phi64(c%var, c%nx, c%ny, c%nz), con64(c%nx, c%ny, c%nz)
    real(8) :: emax(0:c%var), ess(0:c%var), d, npts
    integer :: i, j, k, iv
    emax = 0.0d0
    ess = 0.0d0
    do k = 1, c%nz
      do j = 1, c%ny
        do i = 1, c%nx
          d = abs(con(i, j, k) - con64(i, j, k))
          emax(0) = max(emax(0), d)
          ess(0) = ess(0) + d * d
          do iv = 1, c%var
            d = abs(phi(iv, i, j, k) - phi64(iv, i, j, k))
            emax(iv) = max(emax(iv), d)
            ess(iv) = ess(iv) + d * d
          end do
        end do
      end do
    end do
    npts = real(c%nx, 8) * real(c%ny, 8) * real(c%nz, 8)
    write (u, '(I8)', advance='no') step
    do iv = 0, c%var
      write (u, '(2(A, ES12.5))', advance='no') ' ', emax(iv), ' ', sqrt(ess(iv) / npts)
    end do
    write (u, '(A)') ''
    flush (u)
  end subroutine

end module pf_driver
