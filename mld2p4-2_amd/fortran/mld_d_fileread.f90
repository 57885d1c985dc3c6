! File-reading driver over the MI355X engine: the reference's
! tests/fileread/mld_df_sample.f90 on one process.  Reads the deck layout of
! tests/fileread/runs/mld_dfs.inp from standard input (order of get_parms,
! :584-660), the matrix and optional right-hand side, initial guess and
! reference solution from MatrixMarket files (:196-280; right-hand side all
! ones and zero guess when absent), drops stored zeros (:287), sets the
! preconditioner with the same calls (:363-451), builds it (:456-466), solves
! with psb_krylov (:487-490) and prints the reference's report (:494-553),
! including ||x - xref|| when a reference solution is given.
! Deck entries with no meaning on one process are read and ignored:
! partition (BLOCK / GRAPH) and storage format.  Only the MM file format.
! Two optional trailing lines (no reference counterpart): the device value
! type (DOUBLE or MIXED) and a prefix for prec%dump of the whole hierarchy.
program mld_d_fileread
  use iso_c_binding
  use mld_d_prec_mod
  use mld_data_input
  use mld_mm_io
  implicit none

  type(psb_dspmat_type) :: a
  type(psb_desc_type) :: desc_a
  type(mld_dprec_type) :: prec
  real(psb_dpk_), allocatable :: b(:), x(:), r(:), xref(:), g(:)
  character(len=256) :: mtrx_file, rhs_file, guess_file, sol_file, dump_prefix
  character(len=40) :: filefmt, afmt, part, kmethd, descr, ptype, smther, restr, prol, solve
  character(len=40) :: smther2, restr2, prol2, solve2, mlcycle, aggr_prol, par_aggr_alg, aggr_ord, aggr_filter
  character(len=40) :: csolve, csbsolve, cmat, prec_kind, dummy
  integer(psb_ipk_) :: istopc, itmax, itrace, irst, jsweeps, novr, fill, jsweeps2, novr2, fill2
  integer(psb_ipk_) :: outer_sweeps, maxlevs, csize, thrvsz, cfill, cjswp, info, iter, k, dev
  real(psb_dpk_) :: eps, thr, thr2, mncrratio, athres, cthres, err, t1, t2, thier, tprec, tslv
  real(psb_dpk_) :: resmx, resmxp, xdiffn2, xdiffni, xn2, xni
  real(psb_dpk_), allocatable :: athresv(:)
  logical :: have_ref

  ! ---- deck (get_parms, :584-660)
  call read_data(mtrx_file, 5); call read_data(rhs_file, 5); call read_data(guess_file, 5)
  call read_data(sol_file, 5); call read_data(filefmt, 5); call read_data(afmt, 5); call read_data(part, 5)
  call read_data(kmethd, 5); call read_data(istopc, 5); call read_data(itmax, 5); call read_data(itrace, 5)
  call read_data(irst, 5); call read_data(eps, 5)
  call read_data(descr, 5); call read_data(ptype, 5)
  call read_data(smther, 5); call read_data(jsweeps, 5); call read_data(novr, 5); call read_data(restr, 5)
  call read_data(prol, 5); call read_data(solve, 5); call read_data(fill, 5); call read_data(thr, 5)
  call read_data(smther2, 5); call read_data(jsweeps2, 5); call read_data(novr2, 5); call read_data(restr2, 5)
  call read_data(prol2, 5); call read_data(solve2, 5); call read_data(fill2, 5); call read_data(thr2, 5)
  call read_data(mlcycle, 5); call read_data(outer_sweeps, 5); call read_data(maxlevs, 5); call read_data(csize, 5)
  call read_data(aggr_prol, 5); call read_data(par_aggr_alg, 5); call read_data(aggr_ord, 5)
  call read_data(aggr_filter, 5); call read_data(mncrratio, 5); call read_data(thrvsz, 5)
  allocate(athresv(max(1, thrvsz)))
  if (thrvsz > 0) then
    call read_data(athresv(1:thrvsz), 5)
  else
    call read_data(dummy, 5)
  end if
  call read_data(athres, 5)
  call read_data(csolve, 5); call read_data(csbsolve, 5); call read_data(cmat, 5); call read_data(cfill, 5)
  call read_data(cthres, 5); call read_data(cjswp, 5)
  prec_kind = 'DOUBLE'
  call read_data_opt(prec_kind, 5)
  dump_prefix = 'NONE'
  call read_data_opt(dump_prefix, 5)

  ! ---- matrix and vectors (:196-287)
  t1 = psb_wtime()
  if (trim(filefmt) /= 'MM') then
    write(0, '("Wrong choice for fileformat ",a," (MM only)")') trim(filefmt)
    stop 1
  end if
  call mm_mat_read(a, info, filename=mtrx_file)
  if (info /= 0) then
    write(0, '("Error while reading input matrix ",a," info=",i0)') trim(mtrx_file), info
    stop 1
  end if
  if (trim(rhs_file) /= 'NONE') then
    call mm_array_read(b, info, filename=rhs_file)
    if (info /= 0) stop 1
  end if
  if (.not. allocated(b) .or. size(b) /= a%m) then
    write(*, '("Generating an rhs...")')
    if (allocated(b)) deallocate(b)
    allocate(b(a%m))
    b = 1.0_psb_dpk_
  end if
  allocate(x(a%m), r(a%m))
  x = 0.0_psb_dpk_
  if (trim(guess_file) /= 'NONE') then
    call mm_array_read(g, info, filename=guess_file)
    if (info /= 0) stop 1
    if (size(g) == a%m) x = g
  end if
  have_ref = .false.
  if (trim(sol_file) /= 'NONE') then
    call mm_array_read(xref, info, filename=sol_file)
    if (info /= 0) stop 1
    have_ref = (size(xref) == a%m)
  end if
  if (.not. have_ref) write(*, '("No reference solution...")')
  call psb_clean_zeros(a)
  desc_a%n = a%m
  t2 = psb_wtime() - t1
  write(*, '("Time to read and partition matrix, rhs(, guess, ref sol) : ",es12.5)') t2

  ! ---- preconditioner (:363-451)
  call get_environment_variable('MLD_HIP_DEVICE', dummy, status=k)
  dev = 0
  if (k == 0) read(dummy, *) dev
  call prec%init(ptype, info, device=dev)
  if (info /= 0) stop 1
  select case (trim(ptype))
  case ('NONE', 'NOPREC')
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
        call prec%set('sub_iluthrs', thr2, info, pos='post')
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

  t1 = psb_wtime()
  call prec%hierarchy_build(a, desc_a, info)
  thier = psb_wtime() - t1
  if (info /= 0) then
    write(0, *) 'mld_hierarchy_bld failed, info=', info
    stop 2
  end if
  t1 = psb_wtime()
  call prec%smoothers_build(a, desc_a, info)
  tprec = psb_wtime() - t1
  if (info /= 0) then
    write(0, *) 'mld_smoothers_bld failed, info=', info
    stop 2
  end if
  write(*, '(" ")')
  write(*, '("Preconditioner: ",a)') trim(descr)
  write(*, '("Preconditioner time: ",es12.5)') thier + tprec
  call prec%descr()
  if (trim(dump_prefix) /= 'NONE') then
    call prec%dump(info, istart=1, prefix=trim(dump_prefix), ac=.true., rp=.true., smoother=.true., solver=.true.)
    if (info /= 0) stop 4
  end if

  ! ---- solve (:487-490) and report (:494-553)
  t1 = psb_wtime()
  call psb_krylov(kmethd, a, prec, b, x, eps, desc_a, info, itmax=itmax, iter=iter, err=err, &
       & itrace=itrace, irst=irst, istop=istopc)
  tslv = psb_wtime() - t1
  if (info /= 0) then
    write(0, *) 'solve failed, info=', info
    stop 3
  end if
  r = b
  call a%csmv(-1.0_psb_dpk_, x, 1.0_psb_dpk_, r)
  resmx = sqrt(sum(r*r))
  resmxp = maxval(abs(r))
  if (have_ref) then
    xref = xref - x
    xdiffn2 = sqrt(sum(xref*xref))
    xdiffni = maxval(abs(xref))
    xref = xref + x
    xn2 = sqrt(sum(xref*xref))
    xni = maxval(abs(xref))
  end if
  write(*, '("Matrix: ",a)') trim(mtrx_file)
  write(*, '("Computed solution on ",i8," processors")') 1
  write(*, '("Krylov method                      : ",a)') trim(kmethd)
  write(*, '("Preconditioner                     : ",a)') trim(descr)
  write(*, '("Iterations to convergence          : ",i12)') iter
  write(*, '("Relative error estimate on exit    : ",es12.5)') err
  write(*, '("Number of levels in hierarchy      : ",i12)') prec%get_nlevs()
  write(*, '("Time to build hierarchy            : ",es12.5)') thier
  write(*, '("Time to build smoothers            : ",es12.5)') tprec
  write(*, '("Total time for preconditioner      : ",es12.5)') tprec + thier
  write(*, '("Time to solve system               : ",es12.5)') tslv
  write(*, '("Time per iteration                 : ",es12.5)') tslv / max(iter, 1)
  write(*, '("Total time                         : ",es12.5)') tslv + tprec + thier
  write(*, '("Residual 2-norm                    : ",es12.5)') resmx
  write(*, '("Residual inf-norm                  : ",es12.5)') resmxp
  if (have_ref) then
    write(*, '(" ")')
    write(*, '(2x,a10,9x,a8,4x,a20,5x,a8)') '||X-XREF||', '||XREF||', '||X-XREF||/||XREF||', '(2-norm)'
    write(*, '(1x,3(e12.6,6x))') xdiffn2, xn2, xdiffn2/xn2
    write(*, '(" ")')
    write(*, '(2x,a10,9x,a8,4x,a20,4x,a10)') '||X-XREF||', '||XREF||', '||X-XREF||/||XREF||', '(inf-norm)'
    write(*, '(1x,3(e12.6,6x))') xdiffni, xni, xdiffni/xni
  end if
  call prec%free(info)
end program mld_d_fileread
