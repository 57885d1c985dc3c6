! MatrixMarket input and output for the Fortran host: the routines the
! reference's file-reading driver takes from PSBLAS's utilities
! (tests/fileread/mld_df_sample.f90:200-208: mm_mat_read, mm_array_read),
! with their argument lists, on the serial CSR type of mld_d_prec_mod.
!   mm_mat_read   coordinate real|integer|pattern, general|symmetric|
!                 skew-symmetric (the other triangle is generated), any
!                 entry order, duplicates summed; rows and columns 1-based
!   mm_array_read array real general, column-major, into b(:) (first column)
!                 or b(:,:)
!   mm_mat_write / mm_array_write: coordinate real general / array real
!                 general, 17 significant digits
! Unit: iunit when given, else a free unit; filename '-' or absent: standard
! input (read) / output (write).  info: 0, or 2 unknown header, 3 I/O error.
module mld_mm_io
  use mld_d_prec_mod
  implicit none

  interface mm_array_read
    module procedure mm_vet1_read, mm_vet2_read
  end interface mm_array_read
  interface mm_array_write
    module procedure mm_vet1_write
  end interface mm_array_write

contains

  subroutine open_in(u, info, iunit, filename)
    integer, intent(out) :: u
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: iunit
    character(len=*), optional, intent(in) :: filename
    integer :: ios
    info = 0
    u = 5
    if (present(filename)) then
      if (trim(filename) /= '-') then
        if (present(iunit)) then
          u = iunit
          open(unit=u, file=filename, status='old', action='read', iostat=ios)
        else
          open(newunit=u, file=filename, status='old', action='read', iostat=ios)
        end if
        if (ios /= 0) info = 3
      end if
    else if (present(iunit)) then
      u = iunit
    end if
  end subroutine open_in

  subroutine open_out(u, info, iunit, filename)
    integer, intent(out) :: u
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: iunit
    character(len=*), optional, intent(in) :: filename
    integer :: ios
    info = 0
    u = 6
    if (present(filename)) then
      if (trim(filename) /= '-') then
        if (present(iunit)) then
          u = iunit
          open(unit=u, file=filename, status='replace', action='write', iostat=ios)
        else
          open(newunit=u, file=filename, status='replace', action='write', iostat=ios)
        end if
        if (ios /= 0) info = 3
      end if
    else if (present(iunit)) then
      u = iunit
    end if
  end subroutine open_out

  subroutine close_unit(u)
    integer, intent(in) :: u
    if (u /= 5 .and. u /= 6) close(u)
  end subroutine close_unit

  ! lower-cased header tokens: object, format, field, symmetry
  subroutine read_header(u, fmt, field, symm, info)
    integer, intent(in) :: u
    character(len=*), intent(out) :: fmt, field, symm
    integer(psb_ipk_), intent(out) :: info
    character(len=1024) :: line
    character(len=32) :: mm, obj
    integer :: ios
    info = 0
    read(u, '(a)', iostat=ios) line
    if (ios /= 0) then
      info = 3
      return
    end if
    mm = ' '; obj = ' '; fmt = ' '; field = ' '; symm = ' '
    read(line, *, iostat=ios) mm, obj, fmt, field, symm
    call lower(mm); call lower(obj); call lower(fmt); call lower(field); call lower(symm)
    if (ios /= 0 .or. trim(mm) /= '%%matrixmarket' .or. trim(obj) /= 'matrix') info = 2
  end subroutine read_header

  ! the first line that is not a comment
  subroutine next_data_line(u, line, info)
    integer, intent(in) :: u
    character(len=*), intent(out) :: line
    integer(psb_ipk_), intent(out) :: info
    integer :: ios
    info = 0
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) then
        info = 3
        return
      end if
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '%') cycle
      exit
    end do
  end subroutine next_data_line

  subroutine lower(s)
    character(len=*), intent(inout) :: s
    integer :: k
    do k = 1, len_trim(s)
      if (s(k:k) >= 'A' .and. s(k:k) <= 'Z') s(k:k) = achar(iachar(s(k:k)) + 32)
    end do
  end subroutine lower

  subroutine mm_mat_read(a, info, iunit, filename)
    type(psb_dspmat_type), intent(inout) :: a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: iunit
    character(len=*), optional, intent(in) :: filename
    character(len=32) :: fmt, field, symm
    character(len=1024) :: line
    integer :: u, ios
    integer(psb_ipk_) :: m, n, i, j
    integer(psb_epk_) :: nnz, k, nz, cap
    integer(psb_ipk_), allocatable :: ia(:), ja(:)
    real(psb_dpk_), allocatable :: val(:)
    real(psb_dpk_) :: v
    logical :: sym, skew, pat
    call open_in(u, info, iunit, filename)
    if (info /= 0) return
    call read_header(u, fmt, field, symm, info)
    if (info == 0 .and. trim(fmt) /= 'coordinate') info = 2
    if (info == 0 .and. .not. (trim(field) == 'real' .or. trim(field) == 'integer' .or. trim(field) == 'pattern')) &
         & info = 2
    if (info == 0 .and. .not. (trim(symm) == 'general' .or. trim(symm) == 'symmetric' .or. &
         & trim(symm) == 'skew-symmetric')) info = 2
    if (info == 0) call next_data_line(u, line, info)
    if (info == 0) then
      read(line, *, iostat=ios) m, n, nnz
      if (ios /= 0) info = 3
    end if
    if (info /= 0) then
      call close_unit(u)
      return
    end if
    sym = (trim(symm) /= 'general'); skew = (trim(symm) == 'skew-symmetric'); pat = (trim(field) == 'pattern')
    cap = nnz
    if (sym) cap = 2*nnz
    allocate(ia(max(cap, 1_psb_epk_)), ja(max(cap, 1_psb_epk_)), val(max(cap, 1_psb_epk_)))
    nz = 0
    do k = 1, nnz
      if (pat) then
        read(u, *, iostat=ios) i, j
        v = 1.0_psb_dpk_
      else
        read(u, *, iostat=ios) i, j, v
      end if
      if (ios /= 0 .or. i < 1 .or. i > m .or. j < 1 .or. j > n) then
        info = 3
        exit
      end if
      nz = nz + 1
      ia(nz) = i; ja(nz) = j; val(nz) = v
      if (sym .and. i /= j) then
        nz = nz + 1
        ia(nz) = j; ja(nz) = i; val(nz) = v
        if (skew) val(nz) = -v
      end if
    end do
    call close_unit(u)
    if (info /= 0) return
    call psb_spasb_coo(a, m, nz, ia, ja, val, info)
    a%n = n
  end subroutine mm_mat_read

  subroutine read_array(b2, info, iunit, filename)
    real(psb_dpk_), allocatable, intent(inout) :: b2(:,:)
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: iunit
    character(len=*), optional, intent(in) :: filename
    character(len=32) :: fmt, field, symm
    character(len=1024) :: line
    integer :: u, ios
    integer(psb_ipk_) :: m, n, i, j
    call open_in(u, info, iunit, filename)
    if (info /= 0) return
    call read_header(u, fmt, field, symm, info)
    if (info == 0 .and. (trim(fmt) /= 'array' .or. .not. (trim(field) == 'real' .or. trim(field) == 'integer') &
         & .or. trim(symm) /= 'general')) info = 2
    if (info == 0) call next_data_line(u, line, info)
    if (info == 0) then
      read(line, *, iostat=ios) m, n
      if (ios /= 0) info = 3
    end if
    if (info == 0) then
      if (allocated(b2)) deallocate(b2)
      allocate(b2(m, n))
      do j = 1, n
        do i = 1, m
          call next_data_line(u, line, info)
          if (info /= 0) exit
          read(line, *, iostat=ios) b2(i, j)
          if (ios /= 0) then
            info = 3
            exit
          end if
        end do
        if (info /= 0) exit
      end do
    end if
    call close_unit(u)
  end subroutine read_array

  subroutine mm_vet2_read(b, info, iunit, filename)
    real(psb_dpk_), allocatable, intent(inout) :: b(:,:)
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: iunit
    character(len=*), optional, intent(in) :: filename
    call read_array(b, info, iunit, filename)
  end subroutine mm_vet2_read

  subroutine mm_vet1_read(b, info, iunit, filename)
    real(psb_dpk_), allocatable, intent(inout) :: b(:)
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: iunit
    character(len=*), optional, intent(in) :: filename
    real(psb_dpk_), allocatable :: b2(:,:)
    call read_array(b2, info, iunit, filename)
    if (info /= 0) return
    if (allocated(b)) deallocate(b)
    allocate(b(size(b2, 1)))
    b = b2(:, 1)
  end subroutine mm_vet1_read

  subroutine mm_mat_write(a, mtitle, info, iunit, filename)
    type(psb_dspmat_type), intent(in) :: a
    character(len=*), intent(in) :: mtitle
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: iunit
    character(len=*), optional, intent(in) :: filename
    integer :: u
    integer(psb_ipk_) :: i
    integer(psb_epk_) :: k
    call open_out(u, info, iunit, filename)
    if (info /= 0) return
    write(u, '(a)') '%%MatrixMarket matrix coordinate real general'
    write(u, '(a)') '% '//trim(mtitle)
    write(u, '(i0,1x,i0,1x,i0)') a%m, a%n, a%get_nzeros()
    do i = 1, a%m
      do k = a%irp(i), a%irp(i+1) - 1
        write(u, '(i0,1x,i0,1x,es25.16e3)') i, a%ja(k), a%val(k)
      end do
    end do
    call close_unit(u)
  end subroutine mm_mat_write

  subroutine mm_vet1_write(b, vtitle, info, iunit, filename)
    real(psb_dpk_), intent(in) :: b(:)
    character(len=*), intent(in) :: vtitle
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: iunit
    character(len=*), optional, intent(in) :: filename
    integer :: u
    integer(psb_ipk_) :: i
    call open_out(u, info, iunit, filename)
    if (info /= 0) return
    write(u, '(a)') '%%MatrixMarket matrix array real general'
    write(u, '(a)') '% '//trim(vtitle)
    write(u, '(i0,1x,i0)') size(b), 1
    do i = 1, size(b)
      write(u, '(es25.16e3)') b(i)
    end do
    call close_unit(u)
  end subroutine mm_vet1_write

end module mld_mm_io
