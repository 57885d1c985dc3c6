! 3-D PDE driver over the MI355X engine, reading the reference's input deck
! layout (tests/pdegen/runs/mld_pde3d.inp, parsed in the order of
! get_parms, tests/pdegen/mld_d_pde3d.f90:838-970) from standard input,
! building the 7-point Poisson matrix of mld_d_gen_pde3d (:98-412; right-hand
! side all ones), setting the preconditioner with the same calls as
! :627-716, and solving with psb_krylov (:753-755).  Prints iterations,
! error estimate and phase times like the reference (:788-809).
program mld_d_pde3d
  use iso_c_binding
  use mld_d_prec_mod
  implicit none

  type(psb_dspmat_type) :: a
  type(psb_desc_type) :: desc_a
  type(mld_dprec_type) :: prec
  real(psb_dpk_), allocatable :: b(:), x(:), r(:)
  character(len=40) :: afmt, kmethd, descr, ptype, smther, restr, prol, solve, smther2, restr2, prol2, solve2
  character(len=40) :: mlcycle, aggr_prol, par_aggr_alg, aggr_ord, aggr_filter, csolve, csbsolve, cmat, prec_kind
  integer(psb_ipk_) :: idim, istopc, itmax, itrace, irst, jsweeps, novr, fill, jsweeps2, novr2, fill2
  integer(psb_ipk_) :: outer_sweeps, maxlevs, csize, thrvsz, cfill, cjswp, info, iter, k, dev
  real(psb_dpk_) :: eps, thr, thr2, mncrratio, athres, cthres, err, t0, t1, tprec, tsolve, rn, bn
  real(psb_dpk_), allocatable :: athresv(:)

  call read_data_s(afmt); call read_data_i(idim)
  call read_data_s(kmethd); call read_data_i(istopc); call read_data_i(itmax); call read_data_i(itrace)
  call read_data_i(irst); call read_data_r(eps)
  call read_data_s(descr); call read_data_s(ptype)
  call read_data_s(smther); call read_data_i(jsweeps); call read_data_i(novr); call read_data_s(restr)
  call read_data_s(prol); call read_data_s(solve); call read_data_i(fill); call read_data_r(thr)
  call read_data_s(smther2); call read_data_i(jsweeps2); call read_data_i(novr2); call read_data_s(restr2)
  call read_data_s(prol2); call read_data_s(solve2); call read_data_i(fill2); call read_data_r(thr2)
  call read_data_s(mlcycle); call read_data_i(outer_sweeps); call read_data_i(maxlevs); call read_data_i(csize)
  call read_data_s(aggr_prol); call read_data_s(par_aggr_alg); call read_data_s(aggr_ord); call read_data_s(aggr_filter)
  call read_data_r(mncrratio); call read_data_i(thrvsz)
  allocate(athresv(max(1, thrvsz)))
  if (thrvsz > 0) then
    call read_data_rv(athresv, thrvsz)
  else
    call read_data_s(restr)   ! thresholds line, ignored
  end if
  call read_data_r(athres)
  call read_data_s(csolve); call read_data_s(csbsolve); call read_data_s(cmat); call read_data_i(cfill)
  call read_data_r(cthres); call read_data_i(cjswp)
  prec_kind = 'DOUBLE'
  call read_data_opt(prec_kind)

  call cpu_time(t0)
  call gen_pde3d(idim, a, b)
  desc_a%n = a%m
  allocate(x(a%m), r(a%m))
  x = 0.0_psb_dpk_
  call cpu_time(t1)
  write(*, '("Overall matrix creation time : ",es12.5)') t1 - t0

  ! device from MLD_HIP_DEVICE (default 0; -1 builds on the host only)
  call get_environment_variable('MLD_HIP_DEVICE', restr, status=k)
  dev = 0
  if (k == 0) read(restr, *) dev
  call prec%init(ptype, info, device=dev)
  if (info /= 0) stop 1
  select case (trim(ptype))
  case ('JACOBI', 'L1-JACOBI', 'GS', 'FWGS', 'FBGS')
    call prec%set('smoother_sweeps', jsweeps, info)
  case ('BJAC')
    call prec%set('smoother_sweeps', jsweeps, info)
    call prec%set('sub_solve', solve, info)
    call prec%set('sub_fillin', fill, info)
    call prec%set('sub_iluthrs', thr, info)
  case ('ML')
    call prec%set('ml_cycle', mlcycle, info)
    call prec%set('outer_sweeps', outer_sweeps, info)
    if (csize > 0) call prec%set('min_coarse_size', csize, info)
    if (mncrratio > 1) call prec%set('min_cr_ratio', mncrratio, info)
    if (maxlevs > 0) call prec%set('max_levs', maxlevs, info)
    if (athres >= 0.0_psb_dpk_) call prec%set('aggr_thresh', athres, info)
    if (thrvsz > 0) then
      do k = 1, thrvsz
        call prec%set('aggr_thresh', athresv(k), info, ilev=k+1)
      end do
    end if
    call prec%set('aggr_prol', aggr_prol, info)
    call prec%set('par_aggr_alg', par_aggr_alg, info)
    call prec%set('aggr_ord', aggr_ord, info)
    call prec%set('aggr_filter', aggr_filter, info)
    call prec%set('smoother_type', smther, info)
    call prec%set('smoother_sweeps', jsweeps, info)
    select case (trim(smther))
    case ('GS', 'BWGS', 'FBGS', 'JACOBI', 'L1-JACOBI')
    case default
      call prec%set('sub_solve', solve, info)
      call prec%set('sub_fillin', fill, info)
      call prec%set('sub_iluthrs', thr, info)
    end select
    if (trim(smther2) /= 'NONE') then
      call prec%set('smoother_type', smther2, info, pos='post')
      call prec%set('smoother_sweeps', jsweeps2, info, pos='post')
      select case (trim(smther2))
      case ('GS', 'BWGS', 'FBGS', 'JACOBI', 'L1-JACOBI')
      case default
        call prec%set('sub_solve', solve2, info, pos='post')
        call prec%set('sub_fillin', fill2, info, pos='post')
      end select
    end if
    call prec%set('coarse_solve', csolve, info)
    if (trim(csolve) == 'BJAC') call prec%set('coarse_subsolve', csbsolve, info)
    call prec%set('coarse_mat', cmat, info)
    call prec%set('coarse_fillin', cfill, info)
    call prec%set('coarse_iluthrs', cthres, info)
    call prec%set('coarse_sweeps', cjswp, info)
  end select
  call prec%set_precision(trim(prec_kind), info)

  call cpu_time(t0)
  call prec%hierarchy_build(a, desc_a, info)
  if (info == 0) call prec%smoothers_build(a, desc_a, info)
  call cpu_time(t1)
  if (info /= 0) then
    write(0, *) 'preconditioner build failed, info=', info
    stop 2
  end if
  tprec = t1 - t0
  call prec%descr()

  call cpu_time(t0)
  call psb_krylov(kmethd, a, prec, b, x, eps, desc_a, info, itmax=itmax, iter=iter, err=err, &
       & itrace=itrace, irst=irst, istop=istopc)
  call cpu_time(t1)
  tsolve = t1 - t0
  if (info /= 0) then
    write(0, *) 'solve failed, info=', info
    stop 3
  end if
  r = b
  call a%csmv(-1.0_psb_dpk_, x, 1.0_psb_dpk_, r)
  rn = sqrt(sum(r*r)); bn = sqrt(sum(b*b))
  write(*, '("Computed solution on ",i0," process(es)")') 1
  write(*, '("Number of levels             : ",i0)') prec%get_nlevs()
  write(*, '("Iterations to convergence    : ",i0)') iter
  write(*, '("Error estimate on exit       : ",es12.5)') err
  write(*, '("True relative residual       : ",es12.5)') rn / bn
  write(*, '("Time to buildprec            : ",es12.5)') tprec
  write(*, '("Time to solve system         : ",es12.5)') tsolve
  write(*, '("Time per iteration           : ",es12.5)') tsolve / max(iter, 1)
  call prec%free(info)

contains

  ! read_data (tests/pdegen/data_input.f90:41-56): skip '%' comment lines,
  ! take the first blank-delimited token, drop a trailing '!' comment.
  subroutine next_token(tok)
    character(len=*), intent(out) :: tok
    character(len=256) :: line
    integer :: ios, k
    do
      read(*, '(a)', iostat=ios) line
      if (ios /= 0) then
        tok = ' '
        return
      end if
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '%') cycle
      exit
    end do
    k = index(line, '!')
    if (k > 0) line = line(1:k-1)
    read(line, '(a)') tok
    k = index(trim(tok), ' ')
    if (k > 0) tok = tok(1:k-1)
  end subroutine next_token
  subroutine read_data_s(s)
    character(len=*), intent(out) :: s
    call next_token(s)
  end subroutine read_data_s
  subroutine read_data_opt(s)
    character(len=*), intent(inout) :: s
    character(len=40) :: t
    call next_token(t)
    if (len_trim(t) > 0) s = t
  end subroutine read_data_opt
  subroutine read_data_i(v)
    integer(psb_ipk_), intent(out) :: v
    character(len=64) :: t
    call next_token(t)
    read(t, *) v
  end subroutine read_data_i
  subroutine read_data_r(v)
    real(psb_dpk_), intent(out) :: v
    character(len=64) :: t
    integer :: k
    call next_token(t)
    k = index(t, 'd'); if (k > 0) t(k:k) = 'e'
    k = index(t, 'D'); if (k > 0) t(k:k) = 'e'
    read(t, *) v
  end subroutine read_data_r
  subroutine read_data_rv(v, n)
    real(psb_dpk_), intent(out) :: v(:)
    integer(psb_ipk_), intent(in) :: n
    character(len=256) :: line
    integer :: ios
    do
      read(*, '(a)', iostat=ios) line
      if (ios /= 0) return
      line = adjustl(line)
      if (len_trim(line) == 0 .or. line(1:1) == '%') cycle
      exit
    end do
    read(line, *) v(1:n)
  end subroutine read_data_rv

  ! mld_d_gen_pde3d with a1=a2=a3=1, b=c=0; b = 1
  subroutine gen_pde3d(idim, a, b)
    integer(psb_ipk_), intent(in) :: idim
    type(psb_dspmat_type), intent(out) :: a
    real(psb_dpk_), allocatable, intent(out) :: b(:)
    integer(psb_ipk_), allocatable :: ia(:), ja(:)
    real(psb_dpk_), allocatable :: val(:)
    integer(psb_ipk_) :: m, ix, iy, iz, g, info
    integer(psb_epk_) :: nz
    real(psb_dpk_) :: deltah, sq, off, cen
    m = idim*idim*idim
    deltah = 1.0_psb_dpk_/(idim + 2)
    sq = deltah*deltah
    off = -1.0_psb_dpk_/sq - 0.0_psb_dpk_/(2.0_psb_dpk_*deltah)
    cen = 2.0_psb_dpk_*(1.0_psb_dpk_ + 1.0_psb_dpk_ + 1.0_psb_dpk_)/sq + 0.0_psb_dpk_
    allocate(ia(7*m), ja(7*m), val(7*m), b(m))
    nz = 0
    do ix = 1, idim
      do iy = 1, idim
        do iz = 1, idim
          g = (ix-1)*idim*idim + (iy-1)*idim + iz
          if (ix > 1) call put(g, g - idim*idim, off)
          if (iy > 1) call put(g, g - idim, off)
          if (iz > 1) call put(g, g - 1, off)
          call put(g, g, cen)
          if (iz < idim) call put(g, g + 1, off)
          if (iy < idim) call put(g, g + idim, off)
          if (ix < idim) call put(g, g + idim*idim, off)
        end do
      end do
    end do
    b = 1.0_psb_dpk_
    call psb_spasb_coo(a, m, nz, ia, ja, val, info)
  contains
    subroutine put(i, j, v)
      integer(psb_ipk_), intent(in) :: i, j
      real(psb_dpk_), intent(in) :: v
      nz = nz + 1
      ia(nz) = i; ja(nz) = j; val(nz) = v
    end subroutine put
  end subroutine gen_pde3d

end program mld_d_pde3d
