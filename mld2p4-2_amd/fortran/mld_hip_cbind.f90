! iso_c_binding interfaces of the engine's C-ABI (include/mld2p4_amd.h).
! The pattern follows the reference's own C binding of its third-party
! solvers (mlprec/mld_d_umf_solver.F90:89-121): bind(c) functions returning
! integer(c_int), value scalars, an opaque c_ptr handle.
module mld_hip_cbind
  use iso_c_binding
  implicit none

  interface
    function mldhip_create(p, device) bind(c, name='mldhip_create') result(info)
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: p
      integer(c_int), value :: device
      integer(c_int) :: info
    end function mldhip_create
    function mldhip_destroy(p) bind(c, name='mldhip_destroy') result(info)
      import :: c_ptr, c_int
      type(c_ptr), value :: p
      integer(c_int) :: info
    end function mldhip_destroy
    function mldhip_init(p, ptype) bind(c, name='mldhip_init') result(info)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: ptype
      integer(c_int) :: info
    end function mldhip_init
    function mldhip_seti(p, what, val, ilev, ilmax, pos) bind(c, name='mldhip_seti') result(info)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: what, pos
      integer(c_int), value :: val, ilev, ilmax
      integer(c_int) :: info
    end function mldhip_seti
    function mldhip_setc(p, what, val, ilev, ilmax, pos) bind(c, name='mldhip_setc') result(info)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: what, val, pos
      integer(c_int), value :: ilev, ilmax
      integer(c_int) :: info
    end function mldhip_setc
    function mldhip_setr(p, what, val, ilev, ilmax, pos) bind(c, name='mldhip_setr') result(info)
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: what, pos
      real(c_double), value :: val
      integer(c_int), value :: ilev, ilmax
      integer(c_int) :: info
    end function mldhip_setr
    function mldhip_set_precision(p, kind) bind(c, name='mldhip_set_precision') result(info)
      import :: c_ptr, c_int
      type(c_ptr), value :: p
      integer(c_int), value :: kind
      integer(c_int) :: info
    end function mldhip_set_precision
    function mldhip_hierarchy_build(p, n, irp, ja, val) bind(c, name='mldhip_hierarchy_build') result(info)
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: p
      integer(c_int), value :: n
      integer(c_int64_t) :: irp(*)
      integer(c_int) :: ja(*)
      real(c_double) :: val(*)
      integer(c_int) :: info
    end function mldhip_hierarchy_build
    function mldhip_smoothers_build(p) bind(c, name='mldhip_smoothers_build') result(info)
      import :: c_ptr, c_int
      type(c_ptr), value :: p
      integer(c_int) :: info
    end function mldhip_smoothers_build
    function mldhip_apply(p, x, y, trans) bind(c, name='mldhip_apply') result(info)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: p, x, y
      character(kind=c_char), value :: trans
      integer(c_int) :: info
    end function mldhip_apply
    function mldhip_apply_host(p, x, y, trans) bind(c, name='mldhip_apply_host') result(info)
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: p
      real(c_double) :: x(*), y(*)
      character(kind=c_char), value :: trans
      integer(c_int) :: info
    end function mldhip_apply_host
    function mldhip_krylov_host(p, method, b, x, eps, itmax, istop, iter, err) &
         & bind(c, name='mldhip_krylov_host') result(info)
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: method
      real(c_double) :: b(*), x(*)
      real(c_double), value :: eps
      integer(c_int), value :: itmax, istop
      integer(c_int) :: iter
      real(c_double) :: err
      integer(c_int) :: info
    end function mldhip_krylov_host
    function mldhip_krylov(p, method, b, x, eps, itmax, istop, iter, err) &
         & bind(c, name='mldhip_krylov') result(info)
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: p, b, x
      character(kind=c_char), dimension(*) :: method
      real(c_double), value :: eps
      integer(c_int), value :: itmax, istop
      integer(c_int) :: iter
      real(c_double) :: err
      integer(c_int) :: info
    end function mldhip_krylov
    function mldhip_synchronize(p) bind(c, name='mldhip_synchronize') result(info)
      import :: c_ptr, c_int
      type(c_ptr), value :: p
      integer(c_int) :: info
    end function mldhip_synchronize
    function mldhip_get_nlev(p, nlev) bind(c, name='mldhip_get_nlev') result(info)
      import :: c_ptr, c_int
      type(c_ptr), value :: p
      integer(c_int) :: nlev
      integer(c_int) :: info
    end function mldhip_get_nlev
    function mldhip_get_level_info(p, lev, linfo) bind(c, name='mldhip_get_level_info') result(info)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: p
      integer(c_int), value :: lev
      integer(c_int64_t) :: linfo(10)
      integer(c_int) :: info
    end function mldhip_get_level_info
    function mldhip_get_complexity(p, c) bind(c, name='mldhip_get_complexity') result(info)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: p
      real(c_double) :: c
      integer(c_int) :: info
    end function mldhip_get_complexity
    function mldhip_get_error(p, buf, len) bind(c, name='mldhip_get_error') result(info)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: buf
      integer(c_int), value :: len
      integer(c_int) :: info
    end function mldhip_get_error
    ! HIP runtime, for device-resident vectors (the role of PSBLAS's device
    ! vector molds)
    function hipMalloc(ptr, sz) bind(c, name='hipMalloc') result(e)
      import :: c_ptr, c_size_t, c_int
      type(c_ptr) :: ptr
      integer(c_size_t), value :: sz
      integer(c_int) :: e
    end function hipMalloc
    function hipFree(ptr) bind(c, name='hipFree') result(e)
      import :: c_ptr, c_int
      type(c_ptr), value :: ptr
      integer(c_int) :: e
    end function hipFree
    function hipMemcpy(dst, src, sz, kind) bind(c, name='hipMemcpy') result(e)
      import :: c_ptr, c_size_t, c_int
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: sz
      integer(c_int), value :: kind
      integer(c_int) :: e
    end function hipMemcpy
    function mldhip_dump(p, istart, iend, prefix, head, flags) bind(c, name='mldhip_dump') result(info)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: p
      integer(c_int), value :: istart, iend, flags
      character(kind=c_char), dimension(*) :: prefix, head
      integer(c_int) :: info
    end function mldhip_dump
    function hipDeviceSynchronize() bind(c, name='hipDeviceSynchronize') result(e)
      import :: c_int
      integer(c_int) :: e
    end function hipDeviceSynchronize
    ! a hierarchy built outside the engine (the reference's precv after its
    ! own hierarchy_build / smoothers_build): 0-based level index and CSR;
    ! optional arrays are passed as c_null_ptr through the c_ptr arguments
    function mldhip_upload_level(p, lev, n, irp, ja, val, nf, pirp, pja, pval, rirp, rja, rval) &
         & bind(c, name='mldhip_upload_level') result(info)
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: p
      integer(c_int), value :: lev, n, nf
      integer(c_int64_t) :: irp(*)
      integer(c_int) :: ja(*)
      real(c_double) :: val(*)
      type(c_ptr), value :: pirp, pja, pval, rirp, rja, rval
      integer(c_int) :: info
    end function mldhip_upload_level
    function mldhip_upload_smoother(p, lev, which, solver, ndirp, ndja, ndval, dv, lirp, lja, lval, &
         & uirp, uja, uval) bind(c, name='mldhip_upload_smoother') result(info)
      import :: c_ptr, c_int
      type(c_ptr), value :: p
      integer(c_int), value :: lev, which, solver
      type(c_ptr), value :: ndirp, ndja, ndval, dv, lirp, lja, lval, uirp, uja, uval
      integer(c_int) :: info
    end function mldhip_upload_smoother
    function mldhip_finalize_hierarchy(p, nlev) bind(c, name='mldhip_finalize_hierarchy') result(info)
      import :: c_ptr, c_int
      type(c_ptr), value :: p
      integer(c_int), value :: nlev
      integer(c_int) :: info
    end function mldhip_finalize_hierarchy
    function mldhip_load_hierarchy(p, prefix, n, irp, ja, val, smoothers) &
         & bind(c, name='mldhip_load_hierarchy') result(info)
      import :: c_ptr, c_int, c_int64_t, c_double, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: prefix
      integer(c_int), value :: n, smoothers
      integer(c_int64_t) :: irp(*)
      integer(c_int) :: ja(*)
      real(c_double) :: val(*)
      integer(c_int) :: info
    end function mldhip_load_hierarchy
    ! mld_dmlprec_aply_vect(alpha, prec, x, beta, y, ...): device vectors
    function mldhip_mlprec_aply(p, alpha, x, beta, y, trans) bind(c, name='mldhip_mlprec_aply') result(info)
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: p, x, y
      real(c_double), value :: alpha, beta
      character(kind=c_char), value :: trans
      integer(c_int) :: info
    end function mldhip_mlprec_aply
  end interface

  integer(c_int), parameter :: hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2

contains

  function cstr(s) result(r)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s)+1) :: r
    r = trim(s)//c_null_char
  end function cstr

end module mld_hip_cbind
