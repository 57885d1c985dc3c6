! PDE drivers over the MI355X engine (the reference's tests/pdegen/mld_d_pde2d.f90
! and mld_d_pde3d.f90): read the input deck layout of tests/pdegen/runs/
! mld_pde{2,3}d.inp (the order of get_parms, mld_d_pde3d.f90:838-970) from
! standard input, build the 5- or 7-point Poisson matrix of mld_d_gen_pde2d
! (mld_d_pde2d.f90:96-370) or mld_d_gen_pde3d (mld_d_pde3d.f90:98-412)
! (right-hand side all ones), set the preconditioner with the same calls as
! mld_d_pde3d.f90:627-716, solve with psb_krylov (:753-755) and print
! iterations, error estimate and phase times like the reference (:788-809).
module mld_pde_driver
  use iso_c_binding
  use mld_d_prec_mod
  use mld_data_input
  implicit none
contains

  subroutine pde_driver(ndim)
    integer, intent(in) :: ndim

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

    call read_data(afmt, 5); call read_data(idim, 5)
    call read_data(kmethd, 5); call read_data(istopc, 5); call read_data(itmax, 5); call read_data(itrace, 5)
    call read_data(irst, 5); call read_data(eps, 5)
    call read_data(descr, 5); call read_data(ptype, 5)
    call read_data(smther, 5); call read_data(jsweeps, 5); call read_data(novr, 5); call read_data(restr, 5)
    call read_data(prol, 5); call read_data(solve, 5); call read_data(fill, 5); call read_data(thr, 5)
    call read_data(smther2, 5); call read_data(jsweeps2, 5); call read_data(novr2, 5); call read_data(restr2, 5)
    call read_data(prol2, 5); call read_data(solve2, 5); call read_data(fill2, 5); call read_data(thr2, 5)
    call read_data(mlcycle, 5); call read_data(outer_sweeps, 5); call read_data(maxlevs, 5); call read_data(csize, 5)
    call read_data(aggr_prol, 5); call read_data(par_aggr_alg, 5); call read_data(aggr_ord, 5); call read_data(aggr_filter, 5)
    call read_data(mncrratio, 5); call read_data(thrvsz, 5)
    allocate(athresv(max(1, thrvsz)))
    if (thrvsz > 0) then
      call read_data(athresv(1:thrvsz), 5)
    else
      call read_data(restr, 5)   ! thresholds line, ignored
    end if
    call read_data(athres, 5)
    call read_data(csolve, 5); call read_data(csbsolve, 5); call read_data(cmat, 5); call read_data(cfill, 5)
    call read_data(cthres, 5); call read_data(cjswp, 5)
    prec_kind = 'DOUBLE'
    call read_data_opt(prec_kind, 5)

    t0 = psb_wtime()
    if (ndim == 2) then
      call gen_pde2d(idim, a, b)
    else
      call gen_pde3d(idim, a, b)
    end if
    desc_a%n = a%m
    allocate(x(a%m), r(a%m))
    x = 0.0_psb_dpk_
    t1 = psb_wtime()
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

    t0 = psb_wtime()
    call prec%hierarchy_build(a, desc_a, info)
    if (info == 0) call prec%smoothers_build(a, desc_a, info)
    t1 = psb_wtime()
    if (info /= 0) then
      write(0, *) 'preconditioner build failed, info=', info
      stop 2
    end if
    tprec = t1 - t0
    call prec%descr()

    t0 = psb_wtime()
    call psb_krylov(kmethd, a, prec, b, x, eps, desc_a, info, itmax=itmax, iter=iter, err=err, &
         & itrace=itrace, irst=irst, istop=istopc)
    t1 = psb_wtime()
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
  end subroutine pde_driver

  ! mld_d_gen_pde2d with a1=a2=1, b1=b2=c=0 (mld_d_pde2d.f90:96-370); b = 1.
  ! The mesh size is formed as the reference writes it, 1.e0/(idim+2) with a
  ! default-real (single precision) literal (mld_d_pde2d.f90:163), then
  ! squared in double precision.
  subroutine gen_pde2d(idim, a, b)
    integer(psb_ipk_), intent(in) :: idim
    type(psb_dspmat_type), intent(out) :: a
    real(psb_dpk_), allocatable, intent(out) :: b(:)
    integer(psb_ipk_), allocatable :: ia(:), ja(:)
    real(psb_dpk_), allocatable :: val(:)
    integer(psb_ipk_) :: m, ix, iy, g, info
    integer(psb_epk_) :: nz
    real(psb_dpk_) :: deltah, sq, off, cen
    m = idim*idim
    deltah = 1.e0/(idim + 2)
    sq = deltah*deltah
    off = -1.0_psb_dpk_/sq - 0.0_psb_dpk_/(2.0_psb_dpk_*deltah)
    cen = 2.0_psb_dpk_*(1.0_psb_dpk_ + 1.0_psb_dpk_)/sq + 0.0_psb_dpk_
    allocate(ia(5*m), ja(5*m), val(5*m), b(m))
    nz = 0
    do ix = 1, idim
      do iy = 1, idim
        g = (ix-1)*idim + iy
        if (ix > 1) call put(g, g - idim, off)
        if (iy > 1) call put(g, g - 1, off)
        call put(g, g, cen)
        if (iy < idim) call put(g, g + 1, off)
        if (ix < idim) call put(g, g + idim, off)
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
  end subroutine gen_pde2d

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

end module mld_pde_driver
