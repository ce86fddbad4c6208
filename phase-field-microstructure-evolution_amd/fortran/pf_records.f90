! Fortran unformatted sequential records written and read byte by byte over a
! stream unit, so that files of any size come out in the layout the reference's
! runtimes (Intel, GNU) use for data_save.NNNN and data_rerun.NNNN
! (src/var_diff.f90:453-457, 1134-1138): a 4-byte length marker before and
! after each record; a record longer than 2147483639 bytes is split into
! subrecords, the leading marker negative when more subrecords follow and the
! trailing marker negative when one precedes.  (A one-piece 4-byte marker cannot
! describe the 2.28 GB phi+con record of a 256^3 x 17 state.)
module pf_records
  use iso_c_binding
  use iso_fortran_env, only: int8, int32, int64
  implicit none
  private
  public :: put_record, get_record

  integer(int64), parameter :: kMaxSub = 2147483639_int64

contains

  ! Write one record holding na bytes at a followed by nb bytes at b.
  subroutine put_record(u, a, na, b, nb)
    integer, intent(in) :: u
    type(c_ptr), intent(in) :: a, b
    integer(int64), intent(in) :: na, nb
    integer(int8), pointer :: pa(:), pb(:)
    integer(int64) :: total, done, sub, off, take
    integer(int32) :: head, tail
    total = na + nb
    nullify (pa, pb)
    if (na > 0) call c_f_pointer(a, pa, [na])
    if (nb > 0) call c_f_pointer(b, pb, [nb])
    done = 0
    do
      sub = min(kMaxSub, total - done)
      head = int(sub, int32)
      if (done + sub < total) head = -head
      tail = int(sub, int32)
      if (done > 0) tail = -tail
      write (u) head
      off = done
      do while (off < done + sub)
        if (off < na) then
          take = min(na, done + sub) - off
          write (u) pa(off + 1:off + take)
        else
          take = done + sub - off
          write (u) pb(off - na + 1:off - na + take)
        end if
        off = off + take
      end do
      write (u) tail
      done = done + sub
      if (done >= total) exit
    end do
  end subroutine

  ! Read the record starting at stream position pos (1-based) into na bytes at a
  ! followed by nb bytes at b; pos advances past it.  ok = .false. when the
  ! markers are inconsistent or the record length is not na + nb.
  subroutine get_record(u, pos, a, na, b, nb, ok)
    integer, intent(in) :: u
    integer(int64), intent(inout) :: pos
    type(c_ptr), intent(in) :: a, b
    integer(int64), intent(in) :: na, nb
    logical, intent(out) :: ok
    integer(int8), pointer :: pa(:), pb(:)
    integer(int64) :: total, done, sub, off, take
    integer(int32) :: head, tail
    integer :: ios
    total = na + nb
    nullify (pa, pb)
    if (na > 0) call c_f_pointer(a, pa, [na])
    if (nb > 0) call c_f_pointer(b, pb, [nb])
    done = 0
    ok = .false.
    do
      read (u, pos=pos, iostat=ios) head
      if (ios /= 0) return
      sub = abs(int(head, int64))
      if (done + sub > total) return
      off = done
      do while (off < done + sub)
        if (off < na) then
          take = min(na, done + sub) - off
          read (u, iostat=ios) pa(off + 1:off + take)
        else
          take = done + sub - off
          read (u, iostat=ios) pb(off - na + 1:off - na + take)
        end if
        if (ios /= 0) return
        off = off + take
      end do
      read (u, iostat=ios) tail
      if (ios /= 0 .or. abs(int(tail, int64)) /= sub) return
      pos = pos + 8 + sub
      done = done + sub
      if (head >= 0) exit
    end do
    ok = done == total
  end subroutine

end module pf_records
