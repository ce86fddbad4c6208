! var_diff_hip.x <input.txt> <output_dir/>
!
! Drop-in for the reference program var_diff.x (src/var_diff.f90): the same
! command line, input grammar, initial conditions, precipitate insertion and
! output files, with the timestep (loops 1-14, energies, extrema) executed by
! libpf.so on the GPU.  Single process: nprocs = 1, dims = (1, 1) unless given.
!
! The steps between two host events (a precipitate insertion, a checkpoint, the
! end of the run) go to the library in one pf_step call, which replays them as
! captured graphs; time_step.dat then carries each step's share of that call's
! wall time.
!
! Environment:
!   PF_MODE            fp64_faithful (default) | fp64_reduced | fp32_mixed
!   PF_POLYCRYST_FILE  grain-map file for start = 'polycryst' (default mc_ivar_fin)
!   PF_DRY_RUN         if set: write the initial state as output_dir/step_0000000
!                      and stop before any GPU work
!   PF_TIMING          if set: per-stage GPU times in the loop columns of
!                      timing_data.dat (steps are then issued one at a time with
!                      timing events; otherwise those columns are 0)
!   PF_ERROR_LOG       if set and PF_MODE is not fp64_faithful: a second run of
!                      the same input in fp64_faithful form alongside, and at
!                      every checkpoint the per-field max and RMS difference from
!                      it in output_dir/error_vs_fp64.dat
program var_diff_hip
  use iso_c_binding
  use pf_capi
  use pf_driver
  implicit none

  type(run_config) :: c
  type(pf_params) :: p
  type(pf_diag), allocatable :: d(:), d64(:)
  type(c_ptr) :: h, h64
  character(len=256) :: input_file_path, output_dir, env
  real(8), allocatable :: phi(:, :, :, :), con(:, :, :), gamma(:, :)
  real(8), allocatable :: phi64(:, :, :, :), con64(:, :, :)
  integer, allocatable :: centres(:, :)
  real(8) :: run_time, t_step_s, chunk_s, loop_s(14), ms(16), ms_prev(16), ext(4)
  integer(c_long) :: cnt(16)
  integer :: step, last, n, s, mode, iseed, u93, u95, u97, u98, uerr, elen, estat
  integer(8) :: t0, t1, rate
  logical :: timing, shadow, have_ext
  ! steps per pf_step call at most (the diagnostics rows of a call are held on
  ! the device)
  integer, parameter :: kMaxChunk = 1000

  call get_command_argument(1, input_file_path)
  call get_command_argument(2, output_dir)

  open (newunit=u93, file=trim(output_dir)//'time_step.dat')
  write (u93, '(A)') '# step number, iteration duration, run time elapsed'
  open (newunit=u95, file=trim(output_dir)//'timing_data.dat')
  write (u95, '(A)') '# step number, loop 1 avg, loop 2 avg, loop 3 avg, loop 4 avg, loop 5 avg, '// &
    'loop 6 avg, loop 7 avg, loop 8 avg, loop 9 avg, loop 10 avg, loop 11 avg, loop 12 avg, '// &
    'loop 13 avg, loop 14 avg, loop duration, run time elapsed'
  open (newunit=u97, file=trim(output_dir)//'gamma_matrix.dat')
  open (newunit=u98, file=trim(output_dir)//'summary.dat')
  write (u98, '(A)') '# step number, f_int_tot, f_ch_tot, f_eta_tot, e_tot, c_max_glob, '// &
    'c_min_glob, phi_min_glob, phi_3_max'

  call read_input(trim(input_file_path), c)
  write (*, *) c%dims(1), c%dims(2)
  write (*, *) 1, 1, 1, c%nx, c%ny, c%nz
  write (*, *) 1, 1, 1, c%nz, c%ny, c%nx / 2 + 1

  allocate (phi(c%var, c%nx, c%ny, c%nz), con(c%nx, c%ny, c%nz), gamma(c%var, c%var))
  allocate (centres(3, max(c%num_ppt, 1)))
  phi = 0.d0
  con = c%c_mat
  iseed = c%iseed

  call build_gamma(c%var, c%sigma_1, c%sigma_2, gamma)
  call write_gamma(u97, c%var, gamma)

  if (c%nrun == 1) then
    if (c%start == 'polycryst') then
      call get_environment_variable('PF_POLYCRYST_FILE', env, elen, estat)
      if (estat /= 0) env = 'mc_ivar_fin'
      call ic_polycryst(c, trim(env), phi)
    end if
    if (c%start == 'plane') call ic_plane(c, phi)
    if (c%start == 'circle') call ic_circle(c, phi, con, iseed)
  else
    call read_restart('data_rerun.0001', c, phi, con)
  end if
  write (*, *) 'precipitates assigned'

  call get_environment_variable('PF_DRY_RUN', env, elen, estat)
  if (estat == 0) then
    call write_checkpoint(trim(output_dir), c, 0, phi, con)
    stop
  end if

  mode = PF_MODE_FP64_FAITHFUL
  call get_environment_variable('PF_MODE', env, elen, estat)
  if (estat == 0) then
    select case (trim(env))
    case ('fp64_faithful')
      mode = PF_MODE_FP64_FAITHFUL
    case ('fp64_reduced')
      mode = PF_MODE_FP64_REDUCED
    case ('fp32_mixed')
      mode = PF_MODE_FP32_MIXED
    case default
      write (0, *) 'unknown PF_MODE ', trim(env)
      error stop 1
    end select
  end if
  call get_environment_variable('PF_TIMING', env, elen, estat)
  timing = estat == 0
  call get_environment_variable('PF_ERROR_LOG', env, elen, estat)
  shadow = estat == 0 .and. mode /= PF_MODE_FP64_FAITHFUL
  if (c%ifreq < 1) then
    write (0, *) 'ifreq must be >= 1'
    error stop 1
  end if

  p = params_of(c)
  call pf_check(pf_create(p, int(c%nx, c_int), int(c%ny, c_int), int(c%nz, c_int), &
                          int(c%var, c_int), int(mode, c_int), h), 'pf_create')
  call pf_check(pf_put_state(h, phi, con), 'pf_put_state')
  if (timing) call pf_check(pf_set_timing(h, 1_c_int), 'pf_set_timing')
  if (shadow) then
    allocate (phi64(c%var, c%nx, c%ny, c%nz), con64(c%nx, c%ny, c%nz))
    call pf_check(pf_create(p, int(c%nx, c_int), int(c%ny, c_int), int(c%nz, c_int), &
                            int(c%var, c_int), int(PF_MODE_FP64_FAITHFUL, c_int), h64), 'pf_create')
    call pf_check(pf_put_state(h64, phi, con), 'pf_put_state')
    open (newunit=uerr, file=trim(output_dir)//'error_vs_fp64.dat')
    call write_error_header(uerr, c%var)
  end if
  allocate (d(min(c%n_step, kMaxChunk)), d64(min(c%n_step, kMaxChunk)))
  ms_prev = 0.0d0

  run_time = 0.0d0
  call system_clock(count_rate=rate)
  step = 1
  do while (step <= c%n_step)
    call system_clock(t0)
    have_ext = .false.
    if (step == c%ppt_int) then
      call pf_check(pf_get_state(h, phi, con), 'pf_get_state')
      ! a reported row at ppt_int carries the extrema of the state before the
      ! insertion: the reference computed them at the end of the previous step
      if (mod(step, 2) == 0) then
        call state_extrema(c, phi, con, ext)
        have_ext = .true.
      end if
      call insert_precipitates(c, phi, iseed, centres)
      call pf_check(pf_put_state(h, phi, con), 'pf_put_state')
      if (shadow) then
        call pf_check(pf_get_state(h64, phi64, con64), 'pf_get_state')
        call place_precipitates(c, phi64, centres, c%num_ppt)
        call pf_check(pf_put_state(h64, phi64, con64), 'pf_put_state')
      end if
    end if
    ! this call runs steps step .. last: up to the step before the next
    ! insertion, the next checkpoint or the end
    last = min(c%n_step, step + kMaxChunk - 1, ((step + c%ifreq - 1) / c%ifreq) * c%ifreq)
    if (c%ppt_int > step) last = min(last, c%ppt_int - 1)
    if (timing) last = step
    n = last - step + 1
    call pf_check(pf_step(h, int(n, c_int), 2_c_int, d), 'pf_step')
    if (shadow) call pf_check(pf_step(h64, int(n, c_int), 0_c_int, d64), 'pf_step')
    call system_clock(t1)
    chunk_s = real(t1 - t0, 8) / real(rate, 8)

    loop_s = 0.0d0
    if (timing) then
      ! per-stage GPU times of this step, in the reference's loop columns:
      ! 6-9 nonlinear terms, 1-5 energies/diagnostics, 10-12 mobility, 13 the
      ! spectral update, 14 the transforms
      call pf_check(int(min(0, pf_stage_times(h, ms, cnt, 16_c_int)), c_int), 'pf_stage_times')
      loop_s(6) = (ms(1) - ms_prev(1)) * 1.0d-3
      loop_s(5) = (ms(9) - ms_prev(9)) * 1.0d-3
      loop_s(10) = (ms(10) - ms_prev(10)) * 1.0d-3
      loop_s(13) = (ms(5) - ms_prev(5)) * 1.0d-3
      loop_s(14) = (sum(ms(2:4)) + sum(ms(6:8)) - sum(ms_prev(2:4)) - sum(ms_prev(6:8))) * 1.0d-3
      loop_s(1) = (ms(11) - ms_prev(11)) * 1.0d-3
      ms_prev = ms
    end if
    do s = step, last
      t_step_s = chunk_s / real(n, 8)
      run_time = run_time + t_step_s
      write (u93, *) s, t_step_s, run_time
      if (mod(s, 2) == 0) then
        if (have_ext .and. s == step) then
          d(1)%c_max = ext(1)
          d(1)%c_min = ext(2)
          d(1)%phi_min = ext(3)
          d(1)%phi_3_max = ext(4)
        end if
        call write_summary_row(u98, d(s - step + 1))
      end if
      call write_timing_row(u95, s, loop_s, t_step_s, run_time)
    end do

    if (mod(last, c%ifreq) == 0) then
      call pf_check(pf_get_state(h, phi, con), 'pf_get_state')
      call write_checkpoint(trim(output_dir), c, last, phi, con)
      if (shadow) then
        call pf_check(pf_get_state(h64, phi64, con64), 'pf_get_state')
        call write_error_row(uerr, last, c, phi, con, phi64, con64)
      end if
    end if
    step = last + 1
  end do

  call pf_check(pf_destroy(h), 'pf_destroy')
  if (shadow) then
    call pf_check(pf_destroy(h64), 'pf_destroy')
    close (uerr)
  end if
  close (u93)
  close (u95)
  close (u97)
  close (u98)
end program var_diff_hip
