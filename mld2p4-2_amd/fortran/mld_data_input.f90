! Input-deck reading of the drivers, with the reference's conventions
! (tests/pdegen/data_input.f90, tests/fileread/data_input.f90): lines
! starting with '%' are skipped, a value is the first blank-delimited token of
! the next line, anything after '!' is a comment; Fortran 'd' exponents are
! accepted.  read_data(v, unit) for a string, an integer, a real, or a whole
! line of reals; read_data_opt keeps v when the deck has no further line.
module mld_data_input
  use mld_d_prec_mod, only : psb_dpk_, psb_ipk_
  implicit none

  interface read_data
    module procedure read_char, read_int, read_double, read_double_array
  end interface read_data

contains

  ! next non-comment line with the '!' comment removed; blank at end of input
  subroutine next_line(unit, line)
    integer, intent(in) :: unit
    character(len=*), intent(out) :: line
    integer :: ios, k
    do
      read(unit, '(a)', iostat=ios) line
      if (ios /= 0) then
        line = ' '
        return
      end if
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '%') cycle
      exit
    end do
    k = index(line, '!')
    if (k > 0) line(k:) = ' '
  end subroutine next_line

  subroutine first_token(line, tok)
    character(len=*), intent(in) :: line
    character(len=*), intent(out) :: tok
    integer :: k
    tok = adjustl(line)
    k = index(trim(tok), ' ')
    if (k > 0) tok = tok(1:k-1)
  end subroutine first_token

  subroutine fix_exponent(s)
    character(len=*), intent(inout) :: s
    integer :: k
    do k = 1, len_trim(s)
      if (s(k:k) == 'd' .or. s(k:k) == 'D') s(k:k) = 'e'
    end do
  end subroutine fix_exponent

  subroutine read_char(v, unit)
    character(len=*), intent(out) :: v
    integer, intent(in) :: unit
    character(len=1024) :: line
    call next_line(unit, line)
    call first_token(line, v)
  end subroutine read_char

  subroutine read_int(v, unit)
    integer(psb_ipk_), intent(out) :: v
    integer, intent(in) :: unit
    character(len=1024) :: line
    character(len=64) :: t
    call next_line(unit, line)
    call first_token(line, t)
    read(t, *) v
  end subroutine read_int

  subroutine read_double(v, unit)
    real(psb_dpk_), intent(out) :: v
    integer, intent(in) :: unit
    character(len=1024) :: line
    character(len=64) :: t
    call next_line(unit, line)
    call first_token(line, t)
    call fix_exponent(t)
    read(t, *) v
  end subroutine read_double

  subroutine read_double_array(v, unit)
    real(psb_dpk_), intent(out) :: v(:)
    integer, intent(in) :: unit
    character(len=1024) :: line
    call next_line(unit, line)
    call fix_exponent(line)
    read(line, *) v
  end subroutine read_double_array

  ! optional trailing entries of a deck (engine settings with no reference
  ! counterpart): v unchanged at the end of input
  subroutine read_data_opt(v, unit)
    character(len=*), intent(inout) :: v
    integer, intent(in) :: unit
    character(len=1024) :: line
    call next_line(unit, line)
    if (len_trim(line) > 0) call first_token(line, v)
  end subroutine read_data_opt

end module mld_data_input
