! iso_c_binding interface to libpf.so (include/pf.h).
!
! Every entry point returns 0 on success and a negative code otherwise; the
! message is available from pf_error_message().  pf_check() aborts the program
! with that message, which is how the host reproduces the reference's
! die-on-error behaviour (P3DFFT/MPI abort inside the library there).
module pf_capi
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: PF_OK = 0
  integer(c_int), parameter, public :: PF_MODE_FP64_FAITHFUL = 0
  integer(c_int), parameter, public :: PF_MODE_FP64_REDUCED = 1
  integer(c_int), parameter, public :: PF_MODE_FP32_MIXED = 2
  integer(c_int), parameter, public :: PF_LAYOUT_ZFAST = 1

  ! struct pf_params: names as in input.txt (src/var_diff.f90:106-113)
  type, bind(C), public :: pf_params
    real(c_double) :: t_step, grad_coeff_phi, grad_coeff_c, D_S, D_L, D_gb, mob_phi
    real(c_double) :: sigma_1, sigma_2, am, ap, con_0_mat, con_0_ppt
  end type

  ! struct pf_diag: one summary.dat row (state at the start of `step`)
  type, bind(C), public :: pf_diag
    integer(c_int) :: step, valid
    real(c_double) :: f_int, f_ch, f_eta, e_tot, c_max, c_min, phi_min, phi_3_max, c_mean
  end type

  type, bind(C), public :: pf_ran2_state
    integer(c_int) :: idum2, iy
    integer(c_int) :: iv(32)
  end type

  public :: pf_create, pf_destroy, pf_put_state, pf_get_state, pf_step, pf_diagnose
  public :: pf_set_timing, pf_stage_times, pf_stage_name, pf_f_trans, pf_inv_trans
  public :: pf_f_trans_z, pf_inv_trans_z, pf_k_space_z
  public :: pf_ran2_init, pf_ran2, pf_k_space, pf_gamma, pf_last_error, pf_analyze_state
  public :: pf_error_message, pf_check, ran_2

  interface
    integer(c_int) function pf_create(p, nx, ny, nz, var, mode, h) bind(C, name="pf_create")
      import :: c_int, c_ptr, pf_params
      type(pf_params), intent(in) :: p
      integer(c_int), value :: nx, ny, nz, var, mode
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function pf_destroy(h) bind(C, name="pf_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function pf_put_state(h, phi, con) bind(C, name="pf_put_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: phi(*), con(*)
    end function
    integer(c_int) function pf_get_state(h, phi, con) bind(C, name="pf_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: phi(*), con(*)
    end function
    integer(c_int) function pf_step(h, nsteps, diag_every, out) bind(C, name="pf_step")
      import :: c_int, c_ptr, pf_diag
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps, diag_every
      type(pf_diag), intent(out) :: out(*)
    end function
    integer(c_int) function pf_diagnose(h, out) bind(C, name="pf_diagnose")
      import :: c_int, c_ptr, pf_diag
      type(c_ptr), value :: h
      type(pf_diag), intent(out) :: out
    end function
    integer(c_int) function pf_set_timing(h, on) bind(C, name="pf_set_timing")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function
    integer(c_int) function pf_stage_times(h, ms, cnt, n) bind(C, name="pf_stage_times")
      import :: c_int, c_ptr, c_double, c_long
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ms(*)
      integer(c_long), intent(out) :: cnt(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function pf_stage_name(stage, buf, len) bind(C, name="pf_stage_name")
      import :: c_int, c_char
      integer(c_int), value :: stage, len
      character(kind=c_char), intent(out) :: buf(*)
    end function
    integer(c_int) function pf_f_trans(in, out, nx, ny, nz) bind(C, name="pf_f_trans")
      import :: c_int, c_double
      real(c_double), intent(in) :: in(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int), value :: nx, ny, nz
    end function
    integer(c_int) function pf_inv_trans(in, out, nx, ny, nz) bind(C, name="pf_inv_trans")
      import :: c_int, c_double
      real(c_double), intent(in) :: in(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int), value :: nx, ny, nz
    end function
    ! The same three entry points with complex spectra (C sees interleaved doubles)
    integer(c_int) function pf_f_trans_z(in, out, nx, ny, nz) bind(C, name="pf_f_trans")
      import :: c_int, c_double, c_double_complex
      real(c_double), intent(in) :: in(*)
      complex(c_double_complex), intent(out) :: out(*)
      integer(c_int), value :: nx, ny, nz
    end function
    integer(c_int) function pf_inv_trans_z(in, out, nx, ny, nz) bind(C, name="pf_inv_trans")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex), intent(in) :: in(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int), value :: nx, ny, nz
    end function
    integer(c_int) function pf_k_space_z(nx, ny, nz, kf, kf_sq, kf_4) bind(C, name="pf_k_space")
      import :: c_int, c_double_complex
      integer(c_int), value :: nx, ny, nz
      complex(c_double_complex), intent(out) :: kf(*), kf_sq(*), kf_4(*)
    end function
    subroutine pf_ran2_init(st) bind(C, name="pf_ran2_init")
      import :: pf_ran2_state
      type(pf_ran2_state), intent(out) :: st
    end subroutine
    real(c_double) function pf_ran2(st, idum) bind(C, name="pf_ran2")
      import :: c_double, c_int, pf_ran2_state
      type(pf_ran2_state), intent(inout) :: st
      integer(c_int), intent(inout) :: idum
    end function
    integer(c_int) function pf_k_space(nx, ny, nz, kf, kf_sq, kf_4) bind(C, name="pf_k_space")
      import :: c_int, c_double
      integer(c_int), value :: nx, ny, nz
      real(c_double), intent(out) :: kf(*), kf_sq(*), kf_4(*)
    end function
    integer(c_int) function pf_gamma(var, s1, s2, g) bind(C, name="pf_gamma")
      import :: c_int, c_double
      integer(c_int), value :: var
      real(c_double), value :: s1, s2
      real(c_double), intent(out) :: g(*)
    end function
    integer(c_int) function pf_analyze_state(phi, nx, ny, nz, var, threshold, phi_sq, phi_3, &
                                             grad_phi_3, boundary, var_num, counts) &
      bind(C, name="pf_analyze_state")
      import :: c_int, c_double, c_long
      real(c_double), intent(in) :: phi(*)
      integer(c_int), value :: nx, ny, nz, var
      real(c_double), value :: threshold
      real(c_double), intent(out) :: phi_sq(*), phi_3(*), grad_phi_3(*), boundary(*)
      integer(c_int), intent(out) :: var_num(*)
      integer(c_long), intent(out) :: counts(2)
    end function
    type(c_ptr) function pf_last_error() bind(C, name="pf_last_error")
      import :: c_ptr
    end function
  end interface

  ! generator state behind ran_2 (the reference keeps it in SAVE variables)
  type(pf_ran2_state), save :: ran2_state
  logical, save :: ran2_ready = .false.

contains

  function pf_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = pf_last_error()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate (character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

  subroutine pf_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= PF_OK) then
      write (0, '(A,A,I0,A,A)') trim(what), ' failed (', rc, '): ', pf_error_message()
      error stop 1
    end if
  end subroutine

  ! ran_2(idum): uniform deviate with the generator of src/ran_2.f90; idum <= 0
  ! (re)initialises the sequence.
  real(8) function ran_2(idum)
    integer, intent(inout) :: idum
    integer(c_int) :: id
    if (.not. ran2_ready) then
      call pf_ran2_init(ran2_state)
      ran2_ready = .true.
    end if
    id = int(idum, c_int)
    ran_2 = pf_ran2(ran2_state, id)
    idum = int(id)
  end function

end module pf_capi
