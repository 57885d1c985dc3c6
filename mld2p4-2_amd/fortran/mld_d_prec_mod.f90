! Fortran 2003 host of the MI355X engine: the reference's preconditioner
! API (type mld_dprec_type with init / set / hierarchy_build /
! smoothers_build / build / apply / free / descr,
! mlprec/mld_d_prec_type.f90:84-142) and the psb_krylov call site
! (tests/pdegen/mld_d_pde3d.f90:753-755), over the C-ABI of
! include/mld2p4_amd.h.  PSBLAS is not available, so a minimal layer of
! the PSBLAS types the API mentions is provided here (serial descriptor,
! CSR matrix, vector), with PSBLAS's 1-based CSR storage.
module mld_d_prec_mod
  use iso_c_binding
  use mld_hip_cbind
  implicit none

  integer, parameter :: psb_dpk_ = c_double, psb_ipk_ = c_int, psb_epk_ = c_int64_t
  integer(psb_ipk_), parameter :: psb_success_ = 0

  ! ---------------------------------------------------------------- PSBLAS-shaped types
  type psb_desc_type
    integer(psb_ipk_) :: n = 0
  contains
    procedure :: get_local_rows => desc_rows
    procedure :: get_local_cols => desc_rows
    procedure :: get_global_rows => desc_rows
  end type psb_desc_type

  type psb_dspmat_type   ! CSR, 1-based (psb_d_csr_sparse_mat layout)
    integer(psb_ipk_) :: m = 0, n = 0
    integer(psb_epk_), allocatable :: irp(:)
    integer(psb_ipk_), allocatable :: ja(:)
    real(psb_dpk_), allocatable :: val(:)
  contains
    procedure :: get_nrows => mat_nrows
    procedure :: get_nzeros => mat_nzeros
    procedure :: csmv => mat_csmv
  end type psb_dspmat_type

  ! ---------------------------------------------------------------- the preconditioner
  type mld_dprec_type
    type(c_ptr) :: h = c_null_ptr
    integer(psb_ipk_) :: device = 0
    integer(psb_ipk_) :: n = 0
  contains
    procedure, pass(prec) :: init => mld_dprecinit
    procedure, pass(prec) :: seti => mld_dprecseti
    procedure, pass(prec) :: setc => mld_dprecsetc
    procedure, pass(prec) :: setr => mld_dprecsetr
    generic :: set => seti, setc, setr
    procedure, pass(prec) :: set_precision => mld_dprec_set_precision
    procedure, pass(prec) :: hierarchy_build => mld_d_hierarchy_bld
    procedure, pass(prec) :: smoothers_build => mld_d_smoothers_bld
    procedure, pass(prec) :: build => mld_dprecbld
    procedure, pass(prec) :: hierarchy_load => mld_d_hierarchy_load
    procedure, pass(prec) :: apply2v => mld_dprecaply2
    procedure, pass(prec) :: apply1v => mld_dprecaply1
    procedure, pass(prec) :: apply_dev => mld_dprecaply_dev
    generic :: apply => apply2v, apply1v
    procedure, pass(prec) :: free => mld_dprecfree
    procedure, pass(prec) :: descr => mld_dprecdescr
    procedure, pass(prec) :: get_nlevs => mld_dprec_get_nlevs
    procedure, pass(prec) :: get_complexity => mld_dprec_get_complexity
    procedure, pass(prec) :: dump => mld_dprec_dump
  end type mld_dprec_type

contains

  integer function desc_rows(desc)
    class(psb_desc_type), intent(in) :: desc
    desc_rows = desc%n
  end function desc_rows
  integer function mat_nrows(a)
    class(psb_dspmat_type), intent(in) :: a
    mat_nrows = a%m
  end function mat_nrows
  integer(psb_epk_) function mat_nzeros(a)
    class(psb_dspmat_type), intent(in) :: a
    mat_nzeros = 0
    if (allocated(a%irp)) mat_nzeros = a%irp(a%m+1) - 1
  end function mat_nzeros

  ! y = alpha*A*x + beta*y (host, for residual checks in drivers)
  subroutine mat_csmv(a, alpha, x, beta, y)
    class(psb_dspmat_type), intent(in) :: a
    real(psb_dpk_), intent(in) :: alpha, beta, x(:)
    real(psb_dpk_), intent(inout) :: y(:)
    integer(psb_ipk_) :: i
    integer(psb_epk_) :: k
    real(psb_dpk_) :: acc
    do i = 1, a%m
      acc = 0.0_psb_dpk_
      do k = a%irp(i), a%irp(i+1) - 1
        acc = acc + a%val(k)*x(a%ja(k))
      end do
      if (beta == 0.0_psb_dpk_) then
        y(i) = alpha*acc
      else
        y(i) = beta*y(i) + alpha*acc
      end if
    end do
  end subroutine mat_csmv

  ! Assemble a CSR matrix from COO triplets (psb_spins + psb_spasb with
  ! dupl_add): rows and columns 1-based, columns sorted within rows.
  subroutine psb_spasb_coo(a, m, nz, ia, ja, val, info)
    type(psb_dspmat_type), intent(inout) :: a
    integer(psb_ipk_), intent(in) :: m
    integer(psb_epk_), intent(in) :: nz
    integer(psb_ipk_), intent(in) :: ia(:), ja(:)
    real(psb_dpk_), intent(in) :: val(:)
    integer(psb_ipk_), intent(out) :: info
    integer(psb_epk_), allocatable :: cnt(:), pos(:)
    integer(psb_ipk_), allocatable :: jt(:)
    real(psb_dpk_), allocatable :: vt(:)
    integer(psb_epk_) :: k, q, s, e, w
    integer(psb_ipk_) :: i, jj
    real(psb_dpk_) :: vv
    info = 0
    a%m = m; a%n = m
    allocate(cnt(m+1), pos(m+1), jt(nz), vt(nz))
    cnt = 0
    do k = 1, nz
      cnt(ia(k)+1) = cnt(ia(k)+1) + 1
    end do
    cnt(1) = 1
    do i = 1, m
      cnt(i+1) = cnt(i+1) + cnt(i)
    end do
    pos = cnt
    do k = 1, nz
      q = pos(ia(k)); pos(ia(k)) = q + 1
      jt(q) = ja(k); vt(q) = val(k)
    end do
    ! insertion sort of each row by column, summing duplicates
    if (allocated(a%irp)) deallocate(a%irp, a%ja, a%val)
    allocate(a%irp(m+1), a%ja(nz), a%val(nz))
    w = 0
    a%irp(1) = 1
    do i = 1, m
      s = cnt(i); e = cnt(i+1) - 1
      do k = s + 1, e
        jj = jt(k); vv = vt(k); q = k - 1
        do while (q >= s)
          if (jt(q) <= jj) exit
          jt(q+1) = jt(q); vt(q+1) = vt(q); q = q - 1
        end do
        jt(q+1) = jj; vt(q+1) = vv
      end do
      do k = s, e
        if (w >= a%irp(i) .and. k > s) then
          if (a%ja(w) == jt(k)) then
            a%val(w) = a%val(w) + vt(k)
            cycle
          end if
        end if
        w = w + 1
        a%ja(w) = jt(k); a%val(w) = vt(k)
      end do
      a%irp(i+1) = w + 1
    end do
  end subroutine psb_spasb_coo

  ! wall-clock seconds (psb_wtime)
  real(psb_dpk_) function psb_wtime()
    integer(c_int64_t) :: cnt, rate
    call system_clock(cnt, rate)
    psb_wtime = real(cnt, psb_dpk_)/real(rate, psb_dpk_)
  end function psb_wtime

  ! drop stored zeros (a%clean_zeros), rows keep their order
  subroutine psb_clean_zeros(a)
    type(psb_dspmat_type), intent(inout) :: a
    integer(psb_ipk_) :: i
    integer(psb_epk_) :: k, w, s
    w = 0
    do i = 1, a%m
      s = a%irp(i)
      a%irp(i) = w + 1
      do k = s, a%irp(i+1) - 1
        if (a%val(k) /= 0.0_psb_dpk_) then
          w = w + 1
          a%ja(w) = a%ja(k); a%val(w) = a%val(k)
        end if
      end do
    end do
    a%irp(a%m+1) = w + 1
  end subroutine psb_clean_zeros

  ! ---------------------------------------------------------------- error text
  subroutine report(prec, info, where)
    class(mld_dprec_type), intent(in) :: prec
    integer(psb_ipk_), intent(in) :: info
    character(len=*), intent(in) :: where
    character(kind=c_char, len=256) :: buf
    integer :: k, e
    if (info == 0) return
    e = mldhip_get_error(prec%h, buf, 256)
    k = index(buf, c_null_char)
    if (k == 0) k = 257
    write(0, '(a,": info=",i0," ",a)') where, info, buf(1:k-1)
  end subroutine report

  ! ---------------------------------------------------------------- init / set
  ! mlprec/impl/mld_dprecinit.F90:86-218
  subroutine mld_dprecinit(prec, ptype, info, device)
    class(mld_dprec_type), intent(inout) :: prec
    character(len=*), intent(in) :: ptype
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: device
    if (present(device)) prec%device = device
    if (.not. c_associated(prec%h)) then
      info = mldhip_create(prec%h, prec%device)
      if (info /= 0) return
    end if
    info = mldhip_init(prec%h, cstr(ptype))
    call report(prec, info, 'init')
  end subroutine mld_dprecinit

  subroutine opt_args(ilev, ilmax, pos, il, im, ps)
    integer(psb_ipk_), optional, intent(in) :: ilev, ilmax
    character(len=*), optional, intent(in) :: pos
    integer(c_int), intent(out) :: il, im
    character(len=8), intent(out) :: ps
    il = -1; im = -1; ps = ' '
    if (present(ilev)) il = ilev
    if (present(ilmax)) im = ilmax
    if (present(pos)) ps = pos
  end subroutine opt_args

  ! mlprec/impl/mld_dcprecset.F90:78-453
  subroutine mld_dprecseti(prec, what, val, info, ilev, ilmax, pos)
    class(mld_dprec_type), intent(inout) :: prec
    character(len=*), intent(in) :: what
    integer(psb_ipk_), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: ilev, ilmax
    character(len=*), optional, intent(in) :: pos
    integer(c_int) :: il, im
    character(len=8) :: ps
    call opt_args(ilev, ilmax, pos, il, im, ps)
    info = mldhip_seti(prec%h, cstr(what), val, il, im, cstr(ps))
  end subroutine mld_dprecseti

  ! mlprec/impl/mld_dcprecset.F90:494-556
  subroutine mld_dprecsetc(prec, what, string, info, ilev, ilmax, pos)
    class(mld_dprec_type), intent(inout) :: prec
    character(len=*), intent(in) :: what, string
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: ilev, ilmax
    character(len=*), optional, intent(in) :: pos
    integer(c_int) :: il, im
    character(len=8) :: ps
    call opt_args(ilev, ilmax, pos, il, im, ps)
    info = mldhip_setc(prec%h, cstr(what), cstr(string), il, im, cstr(ps))
  end subroutine mld_dprecsetc

  ! mlprec/impl/mld_dcprecset.F90:597-693
  subroutine mld_dprecsetr(prec, what, val, info, ilev, ilmax, pos)
    class(mld_dprec_type), intent(inout) :: prec
    character(len=*), intent(in) :: what
    real(psb_dpk_), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: ilev, ilmax
    character(len=*), optional, intent(in) :: pos
    integer(c_int) :: il, im
    character(len=8) :: ps
    call opt_args(ilev, ilmax, pos, il, im, ps)
    info = mldhip_setr(prec%h, cstr(what), val, il, im, cstr(ps))
  end subroutine mld_dprecsetr

  ! FP64 (default) or FP32 hierarchy on the device ('DOUBLE' / 'MIXED')
  subroutine mld_dprec_set_precision(prec, kind, info)
    class(mld_dprec_type), intent(inout) :: prec
    character(len=*), intent(in) :: kind
    integer(psb_ipk_), intent(out) :: info
    if (kind == 'MIXED' .or. kind == 'mixed') then
      info = mldhip_set_precision(prec%h, 1_c_int)
    else
      info = mldhip_set_precision(prec%h, 0_c_int)
    end if
  end subroutine mld_dprec_set_precision

  ! ---------------------------------------------------------------- build
  ! mlprec/impl/mld_d_hierarchy_bld.f90:66-490: 1-based CSR passed 0-based.
  subroutine mld_d_hierarchy_bld(a, desc_a, prec, info)
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    class(mld_dprec_type), intent(inout) :: prec
    integer(psb_ipk_), intent(out) :: info
    integer(c_int64_t), allocatable :: irp0(:)
    integer(c_int), allocatable :: ja0(:)
    if (.not. c_associated(prec%h)) then
      info = 3111
      return
    end if
    allocate(irp0(a%m+1), ja0(size(a%ja)))
    irp0 = a%irp - 1
    ja0 = a%ja - 1
    prec%n = a%m
    info = mldhip_hierarchy_build(prec%h, a%m, irp0, ja0, a%val)
    call report(prec, info, 'hierarchy_build')
    if (desc_a%n /= a%m .and. desc_a%n /= 0) info = -1
  end subroutine mld_d_hierarchy_bld

  ! mlprec/impl/mld_d_smoothers_bld.f90 (+ upload of the hierarchy to HBM)
  subroutine mld_d_smoothers_bld(a, desc_a, prec, info)
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    class(mld_dprec_type), intent(inout) :: prec
    integer(psb_ipk_), intent(out) :: info
    info = mldhip_smoothers_build(prec%h)
    call report(prec, info, 'smoothers_build')
  end subroutine mld_d_smoothers_bld

  ! The hierarchy of a prec%dump (levels 2.. from <prefix>_p000_lLLL_
  ! {ac,p,r}.mtx) over the fine matrix a: the state after hierarchy_build.
  ! smoothers = .true.: the smoother files of every level too (a dump from
  ! level 1 with smoother and solver data), the state after smoothers_build.
  subroutine mld_d_hierarchy_load(a, desc_a, prec, prefix, info, smoothers)
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    class(mld_dprec_type), intent(inout) :: prec
    character(len=*), intent(in) :: prefix
    integer(psb_ipk_), intent(out) :: info
    logical, optional, intent(in) :: smoothers
    integer(c_int64_t), allocatable :: irp0(:)
    integer(c_int), allocatable :: ja0(:)
    integer(c_int) :: sm
    if (.not. c_associated(prec%h)) then
      info = 3111
      return
    end if
    sm = 0
    if (present(smoothers)) then
      if (smoothers) sm = 1
    end if
    allocate(irp0(a%m+1), ja0(size(a%ja)))
    irp0 = a%irp - 1
    ja0 = a%ja - 1
    prec%n = a%m
    info = mldhip_load_hierarchy(prec%h, trim(prefix)//c_null_char, a%m, irp0, ja0, a%val, sm)
    call report(prec, info, 'hierarchy_load')
    if (desc_a%n /= a%m .and. desc_a%n /= 0) info = -1
  end subroutine mld_d_hierarchy_load

  ! mlprec/impl/mld_dprecbld.f90:135,143
  subroutine mld_dprecbld(a, desc_a, prec, info)
    type(psb_dspmat_type), intent(in) :: a
    type(psb_desc_type), intent(in) :: desc_a
    class(mld_dprec_type), intent(inout) :: prec
    integer(psb_ipk_), intent(out) :: info
    call prec%hierarchy_build(a, desc_a, info)
    if (info == 0) call prec%smoothers_build(a, desc_a, info)
  end subroutine mld_dprecbld

  ! ---------------------------------------------------------------- apply
  ! mld_dprecaply2_vect (mlprec/impl/mld_dprecaply.f90:305-442): y = op(M^-1) x
  subroutine mld_dprecaply2(prec, x, y, desc_data, info, trans)
    class(mld_dprec_type), intent(inout) :: prec
    real(psb_dpk_), intent(inout) :: x(:)
    real(psb_dpk_), intent(inout) :: y(:)
    type(psb_desc_type), intent(in) :: desc_data
    integer(psb_ipk_), intent(out) :: info
    character(len=1), optional, intent(in) :: trans
    character(kind=c_char) :: t
    t = 'N'
    if (present(trans)) t = trans
    if (.not. c_associated(prec%h)) then
      info = 3112
      return
    end if
    info = mldhip_apply_host(prec%h, x, y, t)
    call report(prec, info, 'apply')
  end subroutine mld_dprecaply2

  ! In-place form (mld_dprecaply1, mlprec/impl/mld_dprecaply.f90:445-575)
  subroutine mld_dprecaply1(prec, x, desc_data, info, trans)
    class(mld_dprec_type), intent(inout) :: prec
    real(psb_dpk_), intent(inout) :: x(:)
    type(psb_desc_type), intent(in) :: desc_data
    integer(psb_ipk_), intent(out) :: info
    character(len=1), optional, intent(in) :: trans
    real(psb_dpk_), allocatable :: w(:)
    allocate(w(size(x)))
    call mld_dprecaply2(prec, x, w, desc_data, info, trans)
    if (info == 0) x = w
  end subroutine mld_dprecaply1

  ! Device-resident vectors (c_ptr to HBM, length n)
  subroutine mld_dprecaply_dev(prec, x, y, info, trans)
    class(mld_dprec_type), intent(inout) :: prec
    type(c_ptr), intent(in) :: x, y
    integer(psb_ipk_), intent(out) :: info
    character(len=1), optional, intent(in) :: trans
    character(kind=c_char) :: t
    t = 'N'
    if (present(trans)) t = trans
    info = mldhip_apply(prec%h, x, y, t)
    if (info == 0) info = mldhip_synchronize(prec%h)
  end subroutine mld_dprecaply_dev

  subroutine mld_dprecfree(prec, info)
    class(mld_dprec_type), intent(inout) :: prec
    integer(psb_ipk_), intent(out) :: info
    info = 0
    if (c_associated(prec%h)) info = mldhip_destroy(prec%h)
    prec%h = c_null_ptr
  end subroutine mld_dprecfree

  integer function mld_dprec_get_nlevs(prec)
    class(mld_dprec_type), intent(in) :: prec
    integer(c_int) :: nl, info
    info = mldhip_get_nlev(prec%h, nl)
    mld_dprec_get_nlevs = nl
  end function mld_dprec_get_nlevs

  real(psb_dpk_) function mld_dprec_get_complexity(prec)
    class(mld_dprec_type), intent(in) :: prec
    real(c_double) :: c
    integer(c_int) :: info
    info = mldhip_get_complexity(prec%h, c)
    mld_dprec_get_complexity = c
  end function mld_dprec_get_complexity

  ! prec%dump (mlprec/mld_d_prec_type.f90:698-729): MatrixMarket files of
  ! levels istart..iend in the reference's names (<prefix>_p000_lLLL_ac.mtx,
  ! _r, _p; <prefix>_sm[2a]_p000_lLLL_nd/_diag/_lower/_upper)
  subroutine mld_dprec_dump(prec, info, istart, iend, prefix, head, ac, rp, smoother, solver, global_num)
    class(mld_dprec_type), intent(in) :: prec
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: istart, iend
    character(len=*), intent(in), optional :: prefix, head
    logical, optional, intent(in) :: ac, rp, smoother, solver, global_num
    integer(c_int) :: il1, iln, flags
    character(len=256) :: pf, hd
    il1 = 0; iln = 0; flags = 0; pf = ' '; hd = ' '
    if (present(istart)) il1 = istart
    if (present(iend)) iln = iend
    if (present(prefix)) pf = prefix
    if (present(head)) hd = head
    if (present(ac)) then
      if (ac) flags = ior(flags, 1)
    end if
    if (present(rp)) then
      if (rp) flags = ior(flags, 2)
    end if
    if (present(smoother)) then
      if (smoother) flags = ior(flags, 4)
    end if
    if (present(solver)) then
      if (solver) flags = ior(flags, 8)
    end if
    ! global_num: one process, so global and local numbering coincide
    info = mldhip_dump(prec%h, il1, iln, cstr(pf), cstr(hd), flags)
    call report(prec, info, 'dump')
  end subroutine mld_dprec_dump

  ! prec%descr (mlprec/impl/mld_dfile_prec_descr.f90:118-180), abridged
  subroutine mld_dprecdescr(prec, iout)
    class(mld_dprec_type), intent(in) :: prec
    integer, optional, intent(in) :: iout
    integer :: u, l, nl, info
    integer(c_int64_t) :: li(10)
    u = 6
    if (present(iout)) u = iout
    nl = prec%get_nlevs()
    write(u, '(a,i0,a,f8.4)') 'Multilevel preconditioner (MI355X engine): levels ', nl, &
         & ', operator complexity ', prec%get_complexity()
    do l = 0, nl - 1
      info = mldhip_get_level_info(prec%h, l, li)
      write(u, '(a,i0,a,i0,a,i0,a,i0,a,i0,a,i0)') '  level ', l + 1, ': n=', li(1), ' nnz=', li(2), &
           & ' sweeps=', li(7), '/', li(8), ' solver=', li(9)
    end do
  end subroutine mld_dprecdescr

  ! ---------------------------------------------------------------- Krylov
  ! psb_krylov(method, a, prec, b, x, eps, desc_a, info, itmax, iter, err,
  ! itrace, irst, istop): the whole iteration runs on the device with the
  ! matrix given to hierarchy_build; b and x are host arrays.
  subroutine psb_krylov(method, a, prec, b, x, eps, desc_a, info, itmax, iter, err, itrace, irst, istop)
    character(len=*), intent(in) :: method
    type(psb_dspmat_type), intent(in) :: a
    class(mld_dprec_type), intent(inout) :: prec
    real(psb_dpk_), intent(in) :: b(:)
    real(psb_dpk_), intent(inout) :: x(:)
    real(psb_dpk_), intent(in) :: eps
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), optional, intent(in) :: itmax, itrace, irst, istop
    integer(psb_ipk_), optional, intent(out) :: iter
    real(psb_dpk_), optional, intent(out) :: err
    integer(c_int) :: itmax_, istop_, it
    real(c_double) :: e
    itmax_ = 1000
    istop_ = 2
    if (present(itmax)) itmax_ = itmax
    if (present(istop)) istop_ = istop
    info = mldhip_krylov_host(prec%h, cstr(method), b, x, eps, itmax_, istop_, it, e)
    call report(prec, info, 'psb_krylov')
    if (present(iter)) iter = it
    if (present(err)) err = e
  end subroutine psb_krylov

end module mld_d_prec_mod
