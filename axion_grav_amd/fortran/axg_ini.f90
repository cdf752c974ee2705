! Parameter-file reader with the semantics of the reference's inifile.f90 (119-446), written
! afresh: lines 'name = value' of at most 120 characters; blank lines skipped; a line equal
! to 'END' ends the file; lines starting with '#' or 'COMMENT' and lines without '=' are
! ignored; a value wrapped in single quotes loses them; the first occurrence of a key wins.
! A key read without a default must be present ("key not found" stop); integers must consist
! of '-+0123456789', logicals of '10TF' and read as T/F.
module axg_ini
  implicit none
  private
  public :: ini_open, ini_close, ini_string, ini_int, ini_double, ini_logical, ini_has

  integer, parameter :: max_line = 120, max_items = 4096, lkey = 128, lval = 1024
  character(len=lkey), allocatable :: keys(:)
  character(len=lval), allocatable :: vals(:)
  integer :: nitems = 0

contains

  subroutine ini_open(path, ok)
    character(*), intent(in) :: path
    logical, intent(out) :: ok
    character(len=max_line) :: line
    character(len=max_line) :: s, name, v
    integer :: u, ios, eq, lp
    ok = .false.
    if (allocated(keys)) deallocate(keys, vals)
    allocate(keys(max_items), vals(max_items))
    nitems = 0
    open(newunit=u, file=path, form='formatted', status='old', action='read', iostat=ios)
    if (ios /= 0) return
    do
       read(u, '(a)', iostat=ios) line
       if (ios /= 0) exit
       if (line == 'END') exit
       if (line == '') cycle
       s = trim(adjustl(line))
       eq = index(s, '=')
       if (eq == 0) cycle
       if (s(1:1) == '#' .or. s(1:7) == 'COMMENT') cycle
       name = trim(s(1:eq-1))
       v = adjustl(s(eq+1:))
       lp = len_trim(v)
       if (lp > 1) then
          if (v(1:1) == "'" .and. v(lp:lp) == "'") v = v(2:lp-1)
       end if
       if (nitems < max_items) then
          nitems = nitems + 1
          keys(nitems) = name
          vals(nitems) = v
       end if
    end do
    close(u)
    ok = .true.
  end subroutine ini_open

  subroutine ini_close()
    if (allocated(keys)) deallocate(keys, vals)
    nitems = 0
  end subroutine ini_close

  function lookup(key) result(v)
    character(*), intent(in) :: key
    character(len=lval) :: v
    integer :: i
    v = ''
    do i = 1, nitems
       if (keys(i) == key) then
          v = vals(i)
          return
       end if
    end do
  end function lookup

  logical function ini_has(key)
    character(*), intent(in) :: key
    ini_has = lookup(key) /= ''
  end function ini_has

  subroutine not_found(key)
    character(*), intent(in) :: key
    write(*, *) 'key not found : '//key
    stop 1
  end subroutine not_found

  function ini_string(key) result(v)
    character(*), intent(in) :: key
    character(len=lval) :: v
    v = lookup(key)
  end function ini_string

  integer function ini_int(key, default)
    character(*), intent(in) :: key
    integer, intent(in), optional :: default
    character(len=lval) :: s
    integer :: ios
    s = lookup(key)
    if (s == '') then
       if (.not. present(default)) call not_found(key)
       ini_int = default
       return
    end if
    if (verify(trim(s), '-+0123456789') /= 0) then
       write(*, *) 'error reading integer for key: '//key
       stop 1
    end if
    read(s, *, iostat=ios) ini_int
    if (ios /= 0) then
       write(*, *) 'error reading integer for key: '//key
       stop 1
    end if
  end function ini_int

  double precision function ini_double(key, default)
    character(*), intent(in) :: key
    double precision, intent(in), optional :: default
    character(len=lval) :: s
    integer :: ios
    s = lookup(key)
    if (s == '') then
       if (.not. present(default)) call not_found(key)
       ini_double = default
       return
    end if
    read(s, *, iostat=ios) ini_double
    if (ios /= 0) then
       write(*, *) 'error reading double for key: '//key
       stop 1
    end if
  end function ini_double

  logical function ini_logical(key, default)
    character(*), intent(in) :: key
    logical, intent(in), optional :: default
    character(len=lval) :: s
    s = lookup(key)
    if (s == '') then
       if (.not. present(default)) call not_found(key)
       ini_logical = default
       return
    end if
    if (verify(trim(s), '10TF') /= 0 .or. (s(1:1) /= 'T' .and. s(1:1) /= 'F')) then
       write(*, *) 'error reading logical for key: '//key
       stop 1
    end if
    ini_logical = s(1:1) == 'T'
  end function ini_logical

end module axg_ini
