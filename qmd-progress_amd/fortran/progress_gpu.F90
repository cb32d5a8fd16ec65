! Fortran module layer over libprogress_gpu.so (iso_c_binding).
!
! Callers keep the reference's module and procedure names and Fortran
! signatures (`use bml`, `use prg_sp2_mod`, ...): src/prg_sp2_mod.F90,
! src/prg_normalize_mod.F90, src/prg_homolumo_mod.F90,
! src/prg_densitymatrix_mod.F90, src/prg_implicit_fermi_mod.F90 and
! src/prg_response_mod.F90.  The `bml` module here is the minimal dense
! subset those callers use (BML itself is not linked); its
! handle wraps the library's C matrix handle.

module bml
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: bml_dp = kind(1.0d0)
  character(len=*), parameter, public :: bml_matrix_dense = "dense"
  character(len=*), parameter, public :: bml_matrix_ellpack = "ellpack"
  character(len=*), parameter, public :: bml_element_real = "real"

  type, public :: bml_matrix_t
    type(c_ptr) :: ptr = c_null_ptr
  end type bml_matrix_t

  public :: bml_zero_matrix, bml_identity_matrix, bml_read_matrix, bml_deallocate
  public :: bml_get_n, bml_import_from_dense, bml_export_to_dense, bml_scale, bml_trace
  public :: bml_copy, bml_write_matrix, prg_gpu_last_error
  public :: bml_copy_new, bml_get_m, bml_threshold, bml_trace_mult, bml_fnorm2, bml_gershgorin

  interface
    function c_bml_zero_matrix(t, p, n, m, d) bind(C, name="bml_zero_matrix") result(r)
      import :: c_int, c_ptr
      integer(c_int), value :: t, p, n, m, d
      type(c_ptr) :: r
    end function
    function c_bml_identity_matrix(t, p, n, m, d) bind(C, name="bml_identity_matrix") result(r)
      import :: c_int, c_ptr
      integer(c_int), value :: t, p, n, m, d
      type(c_ptr) :: r
    end function
    function c_bml_import(t, p, o, n, m, a, thr, d) bind(C, name="bml_import_from_dense") result(r)
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: t, p, o, n, m, d
      type(c_ptr), value :: a
      real(c_double), value :: thr
      type(c_ptr) :: r
    end function
    function c_bml_export(a, o) bind(C, name="bml_export_to_dense") result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
      integer(c_int), value :: o
      type(c_ptr) :: r
    end function
    subroutine c_bml_read(a, f) bind(C, name="bml_read_bml_matrix")
      import :: c_ptr, c_char
      type(c_ptr), value :: a
      character(kind=c_char) :: f(*)
    end subroutine
    subroutine c_bml_write(a, f) bind(C, name="bml_write_bml_matrix")
      import :: c_ptr, c_char
      type(c_ptr), value :: a
      character(kind=c_char) :: f(*)
    end subroutine
    subroutine c_bml_deallocate(a) bind(C, name="bml_deallocate")
      import :: c_ptr
      type(c_ptr) :: a
    end subroutine
    function c_bml_get_n(a) bind(C, name="bml_get_N") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: a
      integer(c_int) :: r
    end function
    subroutine c_bml_scale(s, a, b) bind(C, name="bml_scale")
      import :: c_ptr
      type(c_ptr), value :: s, a, b
    end subroutine
    function c_bml_trace(a) bind(C, name="bml_trace") result(r)
      import :: c_ptr, c_double
      type(c_ptr), value :: a
      real(c_double) :: r
    end function
    subroutine c_bml_copy(a, b) bind(C, name="bml_copy")
      import :: c_ptr
      type(c_ptr), value :: a, b
    end subroutine
    function c_bml_copy_new(a) bind(C, name="bml_copy_new") result(r)
      import :: c_ptr
      type(c_ptr), value :: a
      type(c_ptr) :: r
    end function
    function c_bml_get_m(a) bind(C, name="bml_get_M") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: a
      integer(c_int) :: r
    end function
    subroutine c_bml_threshold(a, thr) bind(C, name="bml_threshold")
      import :: c_ptr, c_double
      type(c_ptr), value :: a
      real(c_double), value :: thr
    end subroutine
    function c_bml_trace_mult(a, b) bind(C, name="bml_trace_mult") result(r)
      import :: c_ptr, c_double
      type(c_ptr), value :: a, b
      real(c_double) :: r
    end function
    function c_bml_fnorm2(a, b) bind(C, name="bml_fnorm2") result(r)
      import :: c_ptr, c_double
      type(c_ptr), value :: a, b
      real(c_double) :: r
    end function
    function c_bml_gershgorin(a) bind(C, name="bml_gershgorin") result(r)
      import :: c_ptr
      type(c_ptr), value :: a
      type(c_ptr) :: r
    end function
    subroutine c_free(p) bind(C, name="free")
      import :: c_ptr
      type(c_ptr), value :: p
    end subroutine
    function c_last_error_code() bind(C, name="prg_last_error_code") result(r)
      import :: c_int
      integer(c_int) :: r
    end function
  end interface

contains

  integer function type_code(matrix_type)
    character(len=*), intent(in) :: matrix_type
    type_code = 1
    if (trim(matrix_type) == "ellpack") type_code = 2
  end function type_code

  !> Status of the last library call (0 = success).
  integer function prg_gpu_last_error()
    prg_gpu_last_error = c_last_error_code()
  end function prg_gpu_last_error

  subroutine bml_zero_matrix(matrix_type, element_type, element_precision, n, m, a, distrib_mode)
    character(len=*), intent(in) :: matrix_type, element_type
    integer, intent(in) :: element_precision, n, m
    type(bml_matrix_t), intent(inout) :: a
    character(len=*), optional, intent(in) :: distrib_mode
    if (c_associated(a%ptr)) call c_bml_deallocate(a%ptr)
    a%ptr = c_bml_zero_matrix(type_code(matrix_type), 2, n, m, 0)
  end subroutine bml_zero_matrix

  subroutine bml_identity_matrix(matrix_type, element_type, element_precision, n, m, a, &
       distrib_mode)
    character(len=*), intent(in) :: matrix_type, element_type
    integer, intent(in) :: element_precision, n, m
    type(bml_matrix_t), intent(inout) :: a
    character(len=*), optional, intent(in) :: distrib_mode
    if (c_associated(a%ptr)) call c_bml_deallocate(a%ptr)
    a%ptr = c_bml_identity_matrix(type_code(matrix_type), 2, n, m, 0)
  end subroutine bml_identity_matrix

  subroutine bml_read_matrix(a, filename)
    type(bml_matrix_t), intent(inout) :: a
    character(len=*), intent(in) :: filename
    call c_bml_read(a%ptr, trim(filename)//c_null_char)
  end subroutine bml_read_matrix

  subroutine bml_write_matrix(a, filename)
    type(bml_matrix_t), intent(in) :: a
    character(len=*), intent(in) :: filename
    call c_bml_write(a%ptr, trim(filename)//c_null_char)
  end subroutine bml_write_matrix

  subroutine bml_deallocate(a)
    type(bml_matrix_t), intent(inout) :: a
    if (c_associated(a%ptr)) call c_bml_deallocate(a%ptr)
    a%ptr = c_null_ptr
  end subroutine bml_deallocate

  integer function bml_get_n(a)
    type(bml_matrix_t), intent(in) :: a
    bml_get_n = c_bml_get_n(a%ptr)
  end function bml_get_n

  subroutine bml_import_from_dense(matrix_type, a_dense, a, threshold, m, distrib_mode)
    character(len=*), intent(in) :: matrix_type
    real(bml_dp), target, intent(in) :: a_dense(:, :)
    type(bml_matrix_t), intent(inout) :: a
    real(bml_dp), intent(in) :: threshold
    integer, intent(in) :: m
    character(len=*), optional, intent(in) :: distrib_mode
    if (c_associated(a%ptr)) call c_bml_deallocate(a%ptr)
    a%ptr = c_bml_import(type_code(matrix_type), 2, 1, size(a_dense, 1), m, c_loc(a_dense), &
         threshold, 0)
  end subroutine bml_import_from_dense

  subroutine bml_export_to_dense(a, a_dense)
    type(bml_matrix_t), intent(in) :: a
    real(bml_dp), intent(inout) :: a_dense(:, :)
    type(c_ptr) :: p
    real(bml_dp), pointer :: buf(:, :)
    integer :: n
    n = c_bml_get_n(a%ptr)
    p = c_bml_export(a%ptr, 1)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, buf, [n, n])
    a_dense(1:n, 1:n) = buf
    call c_free(p)
  end subroutine bml_export_to_dense

  subroutine bml_scale(alpha, a)
    real(bml_dp), target, intent(in) :: alpha
    type(bml_matrix_t), intent(inout) :: a
    real(bml_dp), target :: s
    s = alpha
    call c_bml_scale(c_loc(s), a%ptr, a%ptr)
  end subroutine bml_scale

  real(bml_dp) function bml_trace(a)
    type(bml_matrix_t), intent(in) :: a
    bml_trace = c_bml_trace(a%ptr)
  end function bml_trace

  subroutine bml_copy(a, b)
    type(bml_matrix_t), intent(in) :: a
    type(bml_matrix_t), intent(inout) :: b
    call c_bml_copy(a%ptr, b%ptr)
  end subroutine bml_copy

  subroutine bml_copy_new(a, b)
    type(bml_matrix_t), intent(in) :: a
    type(bml_matrix_t), intent(inout) :: b
    if (c_associated(b%ptr)) call c_bml_deallocate(b%ptr)
    b%ptr = c_bml_copy_new(a%ptr)
  end subroutine bml_copy_new

  integer function bml_get_m(a)
    type(bml_matrix_t), intent(in) :: a
    bml_get_m = c_bml_get_m(a%ptr)
  end function bml_get_m

  !> entries with |a| <= threshold become zero
  subroutine bml_threshold(a, threshold)
    type(bml_matrix_t), intent(inout) :: a
    real(bml_dp), intent(in) :: threshold
    call c_bml_threshold(a%ptr, threshold)
  end subroutine bml_threshold

  !> Tr(A B)
  real(bml_dp) function bml_trace_mult(a, b)
    type(bml_matrix_t), intent(in) :: a, b
    bml_trace_mult = c_bml_trace_mult(a%ptr, b%ptr)
  end function bml_trace_mult

  !> ||A - B||_F
  real(bml_dp) function bml_fnorm2(a, b)
    type(bml_matrix_t), intent(in) :: a, b
    bml_fnorm2 = c_bml_fnorm2(a%ptr, b%ptr)
  end function bml_fnorm2

  !> Gershgorin interval: gbnd(1) = min, gbnd(2) = max
  subroutine bml_gershgorin(a, gbnd)
    type(bml_matrix_t), intent(in) :: a
    real(bml_dp), allocatable, intent(inout) :: gbnd(:)
    type(c_ptr) :: p
    real(bml_dp), pointer :: g(:)
    if (.not. allocated(gbnd)) allocate(gbnd(2))
    p = c_bml_gershgorin(a%ptr)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, g, [2])
    gbnd(1:2) = g(1:2)
    call c_free(p)
  end subroutine bml_gershgorin

end module bml

! src/prg_progress_mod.F90 (library start / stop / version) over
! src/prg_c_interface.F90:51-61's C symbols.
module prg_progress_mod
  implicit none
  private
  public :: prg_progress_init, prg_progress_shutdown, prg_version
  interface
    subroutine prg_progress_init() bind(C, name="prg_progress_init")
    end subroutine
    subroutine prg_progress_shutdown() bind(C, name="prg_progress_shutdown")
    end subroutine
    subroutine prg_version() bind(C, name="prg_version")
    end subroutine
  end interface
end module prg_progress_mod

module prg_gpu_cstr
  use iso_c_binding
  implicit none
contains
  function cstr(s) result(r)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: r
    r = trim(s)//c_null_char
  end function cstr
end module prg_gpu_cstr

module prg_sp2_mod
  use iso_c_binding
  use bml
  use prg_gpu_cstr
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)
  public :: prg_sp2_basic, prg_sp2_basic_tcore, prg_sp2_alg1, prg_sp2_alg2
  public :: prg_sp2_alg2_genseq, prg_sp2_alg1_genseq
  public :: prg_sp2_alg2_seq, prg_sp2_alg1_seq, prg_prg_sp2_alg2_seq_inplace
  public :: prg_prg_sp2_alg1_seq_inplace, prg_sp2_submatrix, prg_sp2_submatrix_inplace

  interface
    subroutine c_sp2_3(h, r, thr, bndfil, mn, mx, conv, tol, v) bind(C, name="prg_sp2_alg2")
      import :: c_ptr, c_double, c_int, c_char
      type(c_ptr), value :: h, r
      real(c_double), value :: thr, bndfil, tol
      integer(c_int), value :: mn, mx, v
      character(kind=c_char) :: conv(*)
    end subroutine
    subroutine c_sp2_alg1(h, r, thr, bndfil, mn, mx, conv, tol, v) bind(C, name="prg_sp2_alg1")
      import :: c_ptr, c_double, c_int, c_char
      type(c_ptr), value :: h, r
      real(c_double), value :: thr, bndfil, tol
      integer(c_int), value :: mn, mx, v
      character(kind=c_char) :: conv(*)
    end subroutine
    subroutine c_sp2_basic(h, r, thr, bndfil, mn, mx, conv, tol, v) bind(C, name="prg_sp2_basic")
      import :: c_ptr, c_double, c_int, c_char
      type(c_ptr), value :: h, r
      real(c_double), value :: thr, bndfil, tol
      integer(c_int), value :: mn, mx, v
      character(kind=c_char) :: conv(*)
    end subroutine
    subroutine c_sp2_basic_tcore(h, r, rf, thr, bndfil, mn, mx, conv, tol, v) &
         bind(C, name="prg_sp2_basic_tcore")
      import :: c_ptr, c_double, c_int, c_char
      type(c_ptr), value :: h, r, rf
      real(c_double), value :: thr, bndfil, tol
      integer(c_int), value :: mn, mx, v
      character(kind=c_char) :: conv(*)
    end subroutine
    subroutine c_alg2_genseq(h, r, thr, bndfil, mn, mx, conv, tol, pp, ic, vv, v) &
         bind(C, name="prg_sp2_alg2_genseq")
      import :: c_ptr, c_double, c_int, c_char
      type(c_ptr), value :: h, r
      real(c_double), value :: thr, bndfil, tol
      integer(c_int), value :: mn, mx, v
      character(kind=c_char) :: conv(*)
      integer(c_int) :: pp(*), ic
      real(c_double) :: vv(*)
    end subroutine
    subroutine c_alg1_genseq(h, r, thr, bndfil, mn, mx, conv, tol, pp, ic, vv) &
         bind(C, name="prg_sp2_alg1_genseq")
      import :: c_ptr, c_double, c_int, c_char
      type(c_ptr), value :: h, r
      real(c_double), value :: thr, bndfil, tol
      integer(c_int), value :: mn, mx
      character(kind=c_char) :: conv(*)
      integer(c_int) :: pp(*), ic
      real(c_double) :: vv(*)
    end subroutine
    subroutine c_alg2_seq(h, r, thr, pp, ic, vv, v) bind(C, name="prg_sp2_alg2_seq")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, r
      real(c_double), value :: thr
      integer(c_int) :: pp(*), ic
      real(c_double) :: vv(*)
      integer(c_int), value :: v
    end subroutine
    subroutine c_alg1_seq(h, r, thr, pp, ic, vv) bind(C, name="prg_sp2_alg1_seq")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, r
      real(c_double), value :: thr
      integer(c_int) :: pp(*), ic
      real(c_double) :: vv(*)
    end subroutine
    subroutine c_alg2_inpl(r, thr, n, pp, ic, vv, mn, mx) bind(C, name="prg_prg_sp2_alg2_seq_inplace")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: r
      real(c_double), value :: thr, mn, mx
      integer(c_int), value :: n
      integer(c_int) :: pp(*), ic
      real(c_double) :: vv(*)
    end subroutine
    subroutine c_alg1_inpl(r, thr, n, pp, ic, vv, mn, mx) bind(C, name="prg_prg_sp2_alg1_seq_inplace")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: r
      real(c_double), value :: thr, mn, mx
      integer(c_int), value :: n
      integer(c_int) :: pp(*), ic
      real(c_double) :: vv(*)
    end subroutine
    subroutine c_submatrix(h, r, thr, pp, ic, vv, mn, mx, core) bind(C, name="prg_sp2_submatrix")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, r
      real(c_double), value :: thr, mn, mx
      integer(c_int) :: pp(*), ic
      real(c_double) :: vv(*)
      integer(c_int), value :: core
    end subroutine
    subroutine c_submatrix_inpl(r, thr, pp, ic, vv, mn, mx, core) &
         bind(C, name="prg_sp2_submatrix_inplace")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: r
      real(c_double), value :: thr, mn, mx
      integer(c_int) :: pp(*), ic
      real(c_double) :: vv(*)
      integer(c_int), value :: core
    end subroutine
  end interface

contains

  integer function vb(verbose)
    integer, optional, intent(in) :: verbose
    vb = 0
    if (present(verbose)) vb = verbose
  end function vb

  !> src/prg_sp2_mod.F90:49-135
  subroutine prg_sp2_basic(h_bml, rho_bml, threshold, bndfil, minsp2iter, maxsp2iter, sp2conv, &
       idemtol, verbose)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold, bndfil, idemtol
    integer, intent(in) :: minsp2iter, maxsp2iter
    character(len=*), intent(in) :: sp2conv
    integer, optional, intent(in) :: verbose
    call c_sp2_basic(h_bml%ptr, rho_bml%ptr, threshold, bndfil, minsp2iter, maxsp2iter, &
         cstr(sp2conv), idemtol, vb(verbose))
  end subroutine prg_sp2_basic

  !> src/prg_sp2_mod.F90:138-246
  subroutine prg_sp2_basic_tcore(h_bml, rho_bml, rhofull_bml, threshold, bndfil, minsp2iter, &
       maxsp2iter, sp2conv, idemtol, verbose)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: rho_bml, rhofull_bml
    real(dp), intent(in) :: threshold, bndfil, idemtol
    integer, intent(in) :: minsp2iter, maxsp2iter
    character(len=*), intent(in) :: sp2conv
    integer, optional, intent(in) :: verbose
    call c_sp2_basic_tcore(h_bml%ptr, rho_bml%ptr, rhofull_bml%ptr, threshold, bndfil, &
         minsp2iter, maxsp2iter, cstr(sp2conv), idemtol, vb(verbose))
  end subroutine prg_sp2_basic_tcore

  !> src/prg_sp2_mod.F90:263-380
  subroutine prg_sp2_alg2(h_bml, rho_bml, threshold, bndfil, minsp2iter, maxsp2iter, sp2conv, &
       idemtol, verbose)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold, bndfil, idemtol
    integer, intent(in) :: minsp2iter, maxsp2iter
    character(len=*), intent(in) :: sp2conv
    integer, optional, intent(in) :: verbose
    call c_sp2_3(h_bml%ptr, rho_bml%ptr, threshold, bndfil, minsp2iter, maxsp2iter, &
         cstr(sp2conv), idemtol, vb(verbose))
  end subroutine prg_sp2_alg2

  !> src/prg_sp2_mod.F90:734-864
  subroutine prg_sp2_alg1(h_bml, rho_bml, threshold, bndfil, minsp2iter, maxsp2iter, sp2conv, &
       idemtol, verbose)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold, bndfil, idemtol
    integer, intent(in) :: minsp2iter, maxsp2iter
    character(len=*), intent(in) :: sp2conv
    integer, optional, intent(in) :: verbose
    call c_sp2_alg1(h_bml%ptr, rho_bml%ptr, threshold, bndfil, minsp2iter, maxsp2iter, &
         cstr(sp2conv), idemtol, vb(verbose))
  end subroutine prg_sp2_alg1

  !> src/prg_sp2_mod.F90:395-533
  subroutine prg_sp2_alg2_genseq(h_bml, rho_bml, threshold, bndfil, minsp2iter, maxsp2iter, &
       sp2conv, idemtol, pp, icount, vv, verbose)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold, bndfil, idemtol
    integer, intent(in) :: minsp2iter, maxsp2iter
    character(len=*), intent(in) :: sp2conv
    integer, intent(inout) :: pp(:), icount
    real(dp), intent(inout) :: vv(:)
    integer, optional, intent(in) :: verbose
    call c_alg2_genseq(h_bml%ptr, rho_bml%ptr, threshold, bndfil, minsp2iter, &
         min(maxsp2iter, size(pp)), cstr(sp2conv), idemtol, pp, icount, vv, vb(verbose))
  end subroutine prg_sp2_alg2_genseq

  !> src/prg_sp2_mod.F90:879-1000
  subroutine prg_sp2_alg1_genseq(h_bml, rho_bml, threshold, bndfil, minsp2iter, maxsp2iter, &
       sp2conv, idemtol, pp, icount, vv)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold, bndfil, idemtol
    integer, intent(in) :: minsp2iter, maxsp2iter
    character(len=*), intent(in) :: sp2conv
    integer, intent(inout) :: pp(:), icount
    real(dp), intent(inout) :: vv(:)
    call c_alg1_genseq(h_bml%ptr, rho_bml%ptr, threshold, bndfil, minsp2iter, &
         min(maxsp2iter, size(pp)), cstr(sp2conv), idemtol, pp, icount, vv)
  end subroutine prg_sp2_alg1_genseq

  !> src/prg_sp2_mod.F90:543-627
  subroutine prg_sp2_alg2_seq(h_bml, rho_bml, threshold, pp, icount, vv, verbose)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold
    integer, intent(inout) :: pp(:), icount
    real(dp), intent(inout) :: vv(:)
    integer, optional, intent(in) :: verbose
    call c_alg2_seq(h_bml%ptr, rho_bml%ptr, threshold, pp, icount, vv, vb(verbose))
  end subroutine prg_sp2_alg2_seq

  !> src/prg_sp2_mod.F90:1010-1089
  subroutine prg_sp2_alg1_seq(h_bml, rho_bml, threshold, pp, icount, vv)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold
    integer, intent(inout) :: pp(:), icount
    real(dp), intent(inout) :: vv(:)
    call c_alg1_seq(h_bml%ptr, rho_bml%ptr, threshold, pp, icount, vv)
  end subroutine prg_sp2_alg1_seq

  !> src/prg_sp2_mod.F90:638-722
  subroutine prg_prg_sp2_alg2_seq_inplace(rho_bml, threshold, pp, icount, vv, mineval, maxeval, &
       verbose)
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold
    integer, intent(inout) :: pp(:), icount
    real(dp), intent(inout) :: vv(:)
    real(dp), intent(in) :: mineval, maxeval
    integer, optional, intent(in) :: verbose
    call c_alg2_inpl(rho_bml%ptr, threshold, size(pp), pp, icount, vv, mineval, maxeval)
  end subroutine prg_prg_sp2_alg2_seq_inplace

  !> src/prg_sp2_mod.F90:1100-1177
  subroutine prg_prg_sp2_alg1_seq_inplace(rho_bml, threshold, pp, icount, vv, mineval, maxeval)
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold
    integer, intent(inout) :: pp(:), icount
    real(dp), intent(inout) :: vv(:)
    real(dp), intent(in) :: mineval, maxeval
    call c_alg1_inpl(rho_bml%ptr, threshold, size(pp), pp, icount, vv, mineval, maxeval)
  end subroutine prg_prg_sp2_alg1_seq_inplace

  !> src/prg_sp2_mod.F90:1190-1246
  subroutine prg_sp2_submatrix(ham_bml, rho_bml, threshold, pp, icount, vv, mineval, maxeval, &
       core_size)
    type(bml_matrix_t), intent(in) :: ham_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold, mineval, maxeval
    integer, intent(inout) :: pp(:), icount
    real(dp), intent(inout) :: vv(:)
    integer, intent(in) :: core_size
    call c_submatrix(ham_bml%ptr, rho_bml%ptr, threshold, pp, icount, vv, mineval, maxeval, &
         core_size)
  end subroutine prg_sp2_submatrix

  !> src/prg_sp2_mod.F90:1259-1310
  subroutine prg_sp2_submatrix_inplace(rho_bml, threshold, pp, icount, vv, mineval, maxeval, &
       core_size)
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold, mineval, maxeval
    integer, intent(inout) :: pp(:), icount
    real(dp), intent(inout) :: vv(:)
    integer, intent(in) :: core_size
    call c_submatrix_inpl(rho_bml%ptr, threshold, pp, icount, vv, mineval, maxeval, core_size)
  end subroutine prg_sp2_submatrix_inplace

end module prg_sp2_mod

module prg_normalize_mod
  use iso_c_binding
  use bml
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)
  public :: prg_normalize, prg_normalize_fermi, prg_normalize_implicit_fermi
  interface
    subroutine c_normalize(h) bind(C, name="prg_normalize")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine
    subroutine c_normalize_fermi(h, h1, hn, mu) bind(C, name="prg_normalize_fermi")
      import :: c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: h1, hn, mu
    end subroutine
    subroutine c_normalize_if(h, c, mu) bind(C, name="prg_normalize_implicit_fermi")
      import :: c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: c, mu
    end subroutine
  end interface
contains
  !> src/prg_normalize_mod.F90:32-46
  subroutine prg_normalize(h_bml)
    type(bml_matrix_t), intent(inout) :: h_bml
    call c_normalize(h_bml%ptr)
  end subroutine prg_normalize
  !> src/prg_normalize_mod.F90:59-72
  subroutine prg_normalize_fermi(h_bml, h1, hN, mu)
    type(bml_matrix_t), intent(inout) :: h_bml
    real(dp), intent(in) :: h1, hN, mu
    call c_normalize_fermi(h_bml%ptr, h1, hN, mu)
  end subroutine prg_normalize_fermi
  !> src/prg_normalize_mod.F90:83-94
  subroutine prg_normalize_implicit_fermi(h_bml, cnst, mu)
    type(bml_matrix_t), intent(inout) :: h_bml
    real(dp), intent(in) :: cnst, mu
    call c_normalize_if(h_bml%ptr, cnst, mu)
  end subroutine prg_normalize_implicit_fermi
end module prg_normalize_mod

module prg_homolumo_mod
  use iso_c_binding
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)
  public :: prg_homolumogap, prg_sp2sequence
  interface
    subroutine c_homolumogap(vv, imax, pp, mn, mx, eh, el, eg, v) bind(C, name="prg_homolumogap")
      import :: c_double, c_int
      real(c_double) :: vv(*), eh, el, eg
      integer(c_int), value :: imax, v
      integer(c_int) :: pp(*)
      real(c_double), value :: mn, mx
    end subroutine
    subroutine c_sp2sequence(pp, imax, mn, mx, eh, el, err, v) bind(C, name="prg_sp2sequence")
      import :: c_double, c_int
      integer(c_int) :: pp(*), imax
      real(c_double), value :: mn, mx, eh, el, err
      integer(c_int), value :: v
    end subroutine
  end interface
contains
  !> src/prg_homolumo_mod.F90:22-89
  subroutine prg_homolumogap(vv, imax, pp, mineval, maxeval, ehomo, elumo, egap, verbose)
    integer, intent(in) :: imax
    integer, intent(in) :: pp(:)
    real(dp), intent(in) :: vv(:)
    real(dp), intent(in) :: mineval, maxeval
    real(dp), intent(inout) :: ehomo, elumo, egap
    integer, optional, intent(in) :: verbose
    integer :: v
    integer, allocatable :: ppc(:)
    real(dp), allocatable :: vvc(:)
    v = 0
    if (present(verbose)) v = verbose
    ppc = pp
    vvc = vv
    call c_homolumogap(vvc, imax, ppc, mineval, maxeval, ehomo, elumo, egap, v)
  end subroutine prg_homolumogap
  !> src/prg_homolumo_mod.F90:97-147
  subroutine prg_sp2sequence(pp, imax, mineval, maxeval, ehomo, elumo, errlimit, verbose)
    integer, intent(inout) :: imax
    integer, intent(inout) :: pp(:)
    real(dp), intent(in) :: mineval, maxeval, ehomo, elumo, errlimit
    integer, optional, intent(in) :: verbose
    integer :: v
    integer :: tmp(100)
    v = 0
    if (present(verbose)) v = verbose
    call c_sp2sequence(tmp, imax, mineval, maxeval, ehomo, elumo, errlimit, v)
    pp(1:min(imax, size(pp))) = tmp(1:min(imax, size(pp)))
  end subroutine prg_sp2sequence
end module prg_homolumo_mod

module prg_densitymatrix_mod
  use iso_c_binding
  use bml
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)
  public :: prg_build_density_T0, prg_check_idempotency, prg_build_density_T_Fermi
  interface
    subroutine c_t0(n, h, r, thr, bndfil, ev) bind(C, name="prg_build_density_T0")
      import :: c_ptr, c_double, c_int
      integer(c_int), value :: n
      type(c_ptr), value :: h, r
      real(c_double), value :: thr, bndfil
      real(c_double) :: ev(*)
    end subroutine
    subroutine c_idem(a, thr, v) bind(C, name="prg_check_idempotency")
      import :: c_ptr, c_double
      type(c_ptr), value :: a
      real(c_double), value :: thr
      real(c_double) :: v
    end subroutine
    subroutine c_tfermi(h, r, thr, kbt, ef, v, drho) bind(C, name="prg_build_density_T_fermi")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, r
      real(c_double), value :: thr, kbt, ef
      integer(c_int), value :: v
      real(c_double) :: drho
    end subroutine
  end interface
contains
  !> src/prg_densitymatrix_mod.F90:36-101
  subroutine prg_build_density_T0(ham_bml, rho_bml, threshold, bndfil, eigenvalues_out)
    type(bml_matrix_t), intent(in) :: ham_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold, bndfil
    real(dp), allocatable, optional, intent(out) :: eigenvalues_out(:)
    real(dp), allocatable :: ev(:)
    integer :: n
    n = bml_get_n(ham_bml)
    allocate(ev(n))
    call c_t0(n, ham_bml%ptr, rho_bml%ptr, threshold, bndfil, ev)
    if (present(eigenvalues_out)) eigenvalues_out = ev
  end subroutine prg_build_density_T0
  !> src/prg_densitymatrix_mod.F90:862-880
  subroutine prg_check_idempotency(mat_bml, threshold, idempotency)
    type(bml_matrix_t), intent(in) :: mat_bml
    real(dp), intent(in) :: threshold
    real(dp), intent(out) :: idempotency
    call c_idem(mat_bml%ptr, threshold, idempotency)
  end subroutine prg_check_idempotency
  !> src/prg_densitymatrix_mod.F90:509-582
  subroutine prg_build_density_T_Fermi(ham_bml, rho_bml, threshold, kbt, ef, verbose, drho)
    type(bml_matrix_t), intent(in) :: ham_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp), intent(in) :: threshold, kbt, ef
    integer, optional, intent(in) :: verbose
    real(dp), optional, intent(inout) :: drho
    real(dp) :: d
    integer :: v
    v = 0
    if (present(verbose)) v = verbose
    call c_tfermi(ham_bml%ptr, rho_bml%ptr, threshold, kbt, ef, v, d)
    if (present(drho)) drho = d
  end subroutine prg_build_density_T_Fermi
end module prg_densitymatrix_mod

module prg_implicit_fermi_mod
  use iso_c_binding
  use bml
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)
  public :: prg_implicit_fermi, prg_implicit_fermi_save_inverse, prg_test_density_matrix
  public :: prg_implicit_fermi_first_order_response, prg_implicit_fermi_zero
  public :: prg_implicit_fermi_response
  public :: prg_conjgrad, prg_newtonschulz, prg_pcg, prg_finite_diff
  interface
    subroutine c_cg(a, p, p2, t, d, w, tol, thr) bind(C, name="prg_conjgrad")
      import :: c_ptr, c_double
      type(c_ptr), value :: a, p, p2, t, d, w
      real(c_double), value :: tol, thr
    end subroutine
    subroutine c_pcg(a, p, p2, d, w, tol, thr) bind(C, name="prg_pcg")
      import :: c_ptr, c_double
      type(c_ptr), value :: a, p, p2, d, w
      real(c_double), value :: tol, thr
    end subroutine
    subroutine c_ns(a, ai, r, t, d, ii, tol, thr, it) bind(C, name="prg_newtonschulz")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: a, ai, r, t, d, ii
      real(c_double), value :: tol, thr
      integer(c_int) :: it
    end subroutine
    subroutine c_fd(h0, hl, mu0, mul, beta, order, lam, h, thr, outl) &
         bind(C, name="prg_gpu_finite_diff")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h0, outl
      type(c_ptr) :: hl(*)
      real(c_double), value :: mu0, beta, lam, h, thr
      real(c_double) :: mul(*)
      integer(c_int), value :: order
    end subroutine
    subroutine c_if(h, p, ns, k, nocc, mu, beta, method, osteps, oel, thr, tol) &
         bind(C, name="prg_implicit_fermi")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, p
      integer(c_int), value :: ns, k, method, osteps
      real(c_double), value :: nocc, mu, beta, oel, thr, tol
    end subroutine
    function c_if_mu() bind(C, name="prg_implicit_fermi_last_mu") result(r)
      import :: c_double
      real(c_double) :: r
    end function
    subroutine c_ifs(inv, h, p, ns, nocc, mu, beta, oel, thr, tol, scf, oi, tn) &
         bind(C, name="prg_implicit_fermi_save_inverse")
      import :: c_ptr, c_double, c_int
      type(c_ptr) :: inv(*)
      type(c_ptr), value :: h, p
      integer(c_int), value :: ns, scf, oi, tn
      real(c_double), value :: nocc, beta, oel, thr, tol
      real(c_double) :: mu
    end subroutine
    subroutine c_counters(oi, tn) bind(C, name="prg_implicit_fermi_counters")
      import :: c_int
      integer(c_int) :: oi, tn
    end subroutine
    subroutine c_tdm(h, p, beta, mu, nocc, osteps, oel, thr) bind(C, name="prg_test_density_matrix")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, p
      real(c_double), value :: beta, mu, nocc, oel, thr
      integer(c_int), value :: osteps
    end subroutine
    function c_tdm_mu() bind(C, name="prg_test_density_matrix_mu") result(r)
      import :: c_double
      real(c_double) :: r
    end function
    subroutine c_ifz(h, p, ns, mu, method, thr, tol) bind(C, name="prg_implicit_fermi_zero")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, p
      integer(c_int), value :: ns, method
      real(c_double), value :: mu, thr, tol
    end subroutine
    subroutine c_ifr(h0, hl, p0, pl, ns, mu0, mu, beta, nocc, otol, ltol, order, thr) &
         bind(C, name="prg_gpu_implicit_fermi_response")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h0, p0
      type(c_ptr) :: hl(*), pl(*)
      integer(c_int), value :: ns, order
      real(c_double) :: mu0, mu(*)
      real(c_double), value :: beta, nocc, otol, ltol, thr
    end subroutine
    subroutine c_for(h0, h1, p0, p1, inv, ns, mu0, beta, nocc, thr) &
         bind(C, name="prg_implicit_fermi_first_order_response")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h0, h1, p0, p1
      type(c_ptr) :: inv(*)
      integer(c_int), value :: ns
      real(c_double), value :: mu0, beta, nocc, thr
    end subroutine
  end interface
contains
  !> src/prg_implicit_fermi_mod.F90:209-336 (mu updated, as in Fortran)
  subroutine prg_implicit_fermi(h_bml, p_bml, nsteps, k, nocc, mu, beta, method, osteps, &
       occErrLimit, threshold, tol)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: p_bml
    integer, intent(in) :: osteps, nsteps, method, k
    real(dp), intent(in) :: nocc, threshold, tol, occErrLimit, beta
    real(dp), intent(inout) :: mu
    call c_if(h_bml%ptr, p_bml%ptr, nsteps, k, nocc, mu, beta, method, osteps, occErrLimit, &
         threshold, tol)
    mu = c_if_mu()
  end subroutine prg_implicit_fermi
  !> src/prg_implicit_fermi_mod.F90:46-193
  subroutine prg_implicit_fermi_save_inverse(Inv_bml, h_bml, p_bml, nsteps, nocc, mu, beta, &
       occErrLimit, threshold, tol, SCF_IT, occiter, totns)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: p_bml, Inv_bml(nsteps)
    integer, intent(in) :: nsteps, SCF_IT
    real(dp), intent(in) :: nocc, threshold, tol, occErrLimit, beta
    real(dp), intent(inout) :: mu
    integer, intent(inout) :: occiter, totns
    type(c_ptr) :: h(nsteps)
    integer :: i
    do i = 1, nsteps
      h(i) = Inv_bml(i)%ptr
    enddo
    call c_ifs(h, h_bml%ptr, p_bml%ptr, nsteps, nocc, mu, beta, occErrLimit, threshold, tol, &
         SCF_IT, occiter, totns)
    call c_counters(occiter, totns)
  end subroutine prg_implicit_fermi_save_inverse
  !> src/prg_implicit_fermi_mod.F90:1112-1186 (mu updated, as in Fortran)
  subroutine prg_test_density_matrix(ham_bml, p_bml, beta, mu, nocc, osteps, occErrLimit, &
       threshold)
    type(bml_matrix_t), intent(in) :: ham_bml
    type(bml_matrix_t), intent(inout) :: p_bml
    real(dp), intent(in) :: beta, nocc, occErrLimit, threshold
    real(dp), intent(inout) :: mu
    integer, intent(in) :: osteps
    call c_tdm(ham_bml%ptr, p_bml%ptr, beta, mu, nocc, osteps, occErrLimit, threshold)
    mu = c_tdm_mu()
  end subroutine prg_test_density_matrix
  !> src/prg_implicit_fermi_mod.F90:439-529
  subroutine prg_implicit_fermi_first_order_response(H0_bml, H1_bml, P0_bml, P1_bml, Inv_bml, &
       nsteps, mu0, beta, nocc, threshold)
    type(bml_matrix_t), intent(in) :: H0_bml, H1_bml, Inv_bml(nsteps)
    type(bml_matrix_t), intent(inout) :: P0_bml, P1_bml
    real(dp), intent(in) :: mu0, threshold, beta, nocc
    integer, intent(in) :: nsteps
    type(c_ptr) :: h(nsteps)
    integer :: i
    do i = 1, nsteps
      h(i) = Inv_bml(i)%ptr
    enddo
    call c_for(H0_bml%ptr, H1_bml%ptr, P0_bml%ptr, P1_bml%ptr, h, nsteps, mu0, beta, nocc, &
         threshold)
  end subroutine prg_implicit_fermi_first_order_response
  !> src/prg_implicit_fermi_mod.F90:347-425 (tol optional, as there)
  subroutine prg_implicit_fermi_zero(h_bml, p_bml, nsteps, mu, method, threshold, tol)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: p_bml
    integer, intent(in) :: nsteps, method
    real(dp), intent(in) :: mu, threshold
    real(dp), intent(inout), optional :: tol
    real(dp) :: t
    t = 1.0e-10_dp
    if (present(tol)) t = tol
    call c_ifz(h_bml%ptr, p_bml%ptr, nsteps, mu, method, threshold, t)
  end subroutine prg_implicit_fermi_zero
  !> src/prg_implicit_fermi_mod.F90:547-712: H2/H3, P2/P3 are read only up to
  !> `order`; mu0 and mu(1:order) are updated
  subroutine prg_implicit_fermi_response(H0_bml, H1_bml, H2_bml, H3_bml, P0_bml, P1_bml, &
       P2_bml, P3_bml, nsteps, mu0, mu, beta, nocc, occ_tol, lin_tol, order, threshold)
    type(bml_matrix_t), intent(in) :: H0_bml, H1_bml, H2_bml, H3_bml
    type(bml_matrix_t), intent(inout) :: P0_bml, P1_bml, P2_bml, P3_bml
    real(dp), intent(inout) :: mu0
    real(dp), allocatable, intent(inout) :: mu(:)
    real(dp), intent(in) :: beta, occ_tol, lin_tol, nocc
    integer, intent(in) :: nsteps
    integer :: order
    real(dp) :: threshold
    type(c_ptr) :: hl(3), pl(3)
    real(c_double) :: m0, mr(3)
    hl = (/ H1_bml%ptr, H2_bml%ptr, H3_bml%ptr /)
    pl = (/ P1_bml%ptr, P2_bml%ptr, P3_bml%ptr /)
    m0 = mu0
    mr = 0.0_dp
    call c_ifr(H0_bml%ptr, hl, P0_bml%ptr, pl, nsteps, m0, mr, beta, nocc, occ_tol, lin_tol, &
         order, threshold)
    mu0 = m0
    if (.not. allocated(mu)) allocate(mu(order))
    mu(1:order) = mr(1:order)
  end subroutine prg_implicit_fermi_response
  !> src/prg_implicit_fermi_mod.F90:1006-1048
  subroutine prg_conjgrad(A_bml, p_bml, p2_bml, tmp_bml, d_bml, w_bml, cg_tol, threshold)
    type(bml_matrix_t), intent(in) :: A_bml, p2_bml
    type(bml_matrix_t), intent(inout) :: p_bml, tmp_bml, d_bml, w_bml
    real(dp), intent(in) :: cg_tol, threshold
    call c_cg(A_bml%ptr, p_bml%ptr, p2_bml%ptr, tmp_bml%ptr, d_bml%ptr, w_bml%ptr, cg_tol, &
         threshold)
  end subroutine prg_conjgrad
  !> src/prg_implicit_fermi_mod.F90:934-996 (p2_bml returns the residual)
  subroutine prg_pcg(A_bml, p_bml, p2_bml, d_bml, wtmp_bml, cg_tol, threshold)
    type(bml_matrix_t), intent(in) :: A_bml
    type(bml_matrix_t), intent(inout) :: p_bml, p2_bml, d_bml, wtmp_bml
    real(dp), intent(in) :: cg_tol, threshold
    call c_pcg(A_bml%ptr, p_bml%ptr, p2_bml%ptr, d_bml%ptr, wtmp_bml%ptr, cg_tol, threshold)
  end subroutine prg_pcg
  !> src/prg_implicit_fermi_mod.F90:888-922
  subroutine prg_newtonschulz(a_bml, ai_bml, r_bml, tmp_bml, d_bml, I_bml, tol, threshold, &
       num_iter)
    type(bml_matrix_t), intent(inout) :: ai_bml, r_bml, tmp_bml, d_bml
    type(bml_matrix_t), intent(in) :: a_bml, I_bml
    real(dp), intent(in) :: threshold, tol
    integer, intent(out) :: num_iter
    integer(c_int) :: it
    call c_ns(a_bml%ptr, ai_bml%ptr, r_bml%ptr, tmp_bml%ptr, d_bml%ptr, I_bml%ptr, tol, &
         threshold, it)
    num_iter = it
  end subroutine prg_newtonschulz
  !> src/prg_implicit_fermi_mod.F90:726-831 (prints the differences, as the reference)
  subroutine prg_finite_diff(H0_bml, H_list, mu0, mu_list, beta, order, lambda, h, threshold)
    type(bml_matrix_t), intent(in) :: H0_bml
    type(bml_matrix_t), allocatable, intent(in) :: H_list(:)
    real(dp), allocatable, intent(in) :: mu_list(:)
    real(dp), intent(in) :: mu0, lambda, beta, threshold, h
    integer, intent(in) :: order
    type(c_ptr) :: hl(order)
    real(c_double) :: ml(order)
    integer :: i
    do i = 1, order
      hl(i) = H_list(i)%ptr
      ml(i) = mu_list(i)
    enddo
    call c_fd(H0_bml%ptr, hl, mu0, ml, beta, order, lambda, h, threshold, c_null_ptr)
  end subroutine prg_finite_diff
end module prg_implicit_fermi_mod

module prg_response_mod
  use iso_c_binding
  use bml
  use prg_gpu_cstr
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)
  public :: prg_compute_response_SP2, prg_compute_response_RS, prg_compute_response_FD
  public :: prg_canon_response, prg_canon_response_orig, prg_canon_response_vector
  public :: prg_canon_response_p1_dpdmu
  interface
    subroutine c_rs(h, h1, rsp, lam, bndfil, thr, v) bind(C, name="prg_compute_response_RS")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, h1, rsp
      real(c_double), value :: lam, bndfil, thr
      integer(c_int), value :: v
    end subroutine
    subroutine c_fdr(h, h1, rsp, delta, bndfil, thr, v) bind(C, name="prg_compute_response_FD")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, h1, rsp
      real(c_double), value :: delta, bndfil, thr
      integer(c_int), value :: v
    end subroutine
    subroutine c_canon(p1, h1, nocc, beta, e, mu0, m, hdim) bind(C, name="prg_canon_response")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: p1, h1
      real(c_double), value :: nocc, beta, mu0
      real(c_double) :: e(*)
      integer(c_int), value :: m, hdim
    end subroutine
    subroutine c_canon_o(p1, h1, nocc, beta, e, mu0, m, thr, hdim) &
         bind(C, name="prg_canon_response_orig")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: p1, h1
      real(c_double), value :: nocc, beta, mu0, thr
      real(c_double) :: e(*)
      integer(c_int), value :: m, hdim
    end subroutine
    subroutine c_canon_v(p1, h1, nocc, beta, e, mu0, m, thr, hdim) &
         bind(C, name="prg_canon_response_vector")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: p1, h1
      real(c_double), value :: nocc, beta, mu0, thr
      real(c_double) :: e(*)
      integer(c_int), value :: m, hdim
    end subroutine
    subroutine c_canon_p(p1, dp1, h1, norbs, beta, q, e, mu0, m, hdim) &
         bind(C, name="prg_canon_response_p1_dpdmu")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: p1, h1, q
      real(c_double) :: dp1(*), e(*)
      real(c_double), value :: beta, mu0
      integer(c_int), value :: norbs, m, hdim
    end subroutine
    subroutine c_rsp(h, h1, rsp, rho, lam, bndfil, mn, mx, conv, tol, thr, v) &
         bind(C, name="prg_compute_response_SP2")
      import :: c_ptr, c_double, c_int, c_char
      type(c_ptr), value :: h, h1, rsp, rho
      real(c_double), value :: lam, bndfil, tol, thr
      integer(c_int), value :: mn, mx, v
      character(kind=c_char) :: conv(*)
    end subroutine
  end interface
contains
  !> src/prg_response_mod.F90:653-782
  subroutine prg_compute_response_SP2(ham_bml, prt_bml, rsp_bml, rho_bml, lambda, bndfil, &
       minsp2iter, maxsp2iter, sp2conv, idemtol, threshold, verbose)
    type(bml_matrix_t), intent(in) :: ham_bml, prt_bml
    type(bml_matrix_t), intent(inout) :: rsp_bml, rho_bml
    real(dp) :: lambda
    real(dp), intent(in) :: bndfil, threshold, idemtol
    integer, intent(in) :: minsp2iter, maxsp2iter
    character(len=*), intent(in) :: sp2conv
    integer :: verbose
    call c_rsp(ham_bml%ptr, prt_bml%ptr, rsp_bml%ptr, rho_bml%ptr, lambda, bndfil, minsp2iter, &
         maxsp2iter, cstr(sp2conv), idemtol, threshold, verbose)
  end subroutine prg_compute_response_SP2
  !> src/prg_response_mod.F90:250-361
  subroutine prg_compute_response_RS(ham_bml, prt_bml, rsp_bml, lambda, bndfil, threshold, verbose)
    type(bml_matrix_t), intent(in) :: ham_bml, prt_bml
    type(bml_matrix_t), intent(inout) :: rsp_bml
    real(dp) :: lambda
    real(dp), intent(in) :: bndfil, threshold
    integer :: verbose
    call c_rs(ham_bml%ptr, prt_bml%ptr, rsp_bml%ptr, lambda, bndfil, threshold, verbose)
  end subroutine prg_compute_response_RS
  !> src/prg_response_mod.F90:380-424
  subroutine prg_compute_response_FD(ham_bml, prt_bml, rsp_bml, prg_delta, bndfil, threshold, &
       verbose)
    type(bml_matrix_t), intent(in) :: ham_bml, prt_bml
    type(bml_matrix_t), intent(inout) :: rsp_bml
    real(dp) :: prg_delta
    real(dp), intent(in) :: bndfil, threshold
    integer :: verbose
    call c_fdr(ham_bml%ptr, prt_bml%ptr, rsp_bml%ptr, prg_delta, bndfil, threshold, verbose)
  end subroutine prg_compute_response_FD
  !> src/prg_response_mod.F90:947-1022
  subroutine prg_canon_response(P1_bml, H1_bml, Nocc, beta, evals, mu0, m, HDIM)
    type(bml_matrix_t), intent(inout) :: P1_bml, H1_bml
    integer, intent(in) :: HDIM, m
    real(dp), intent(in) :: evals(HDIM), beta, mu0, nocc
    real(c_double) :: e(HDIM)
    e = evals
    call c_canon(P1_bml%ptr, H1_bml%ptr, nocc, beta, e, mu0, m, HDIM)
  end subroutine prg_canon_response
  !> src/prg_response_mod.F90:1038-1105
  subroutine prg_canon_response_orig(P1_bml, H1_bml, Nocc, beta, evals, mu0, m, thresh, HDIM)
    type(bml_matrix_t), intent(inout) :: P1_bml, H1_bml
    integer, intent(in) :: HDIM, m
    real(dp), intent(in) :: evals(HDIM), thresh, beta, mu0, nocc
    real(c_double) :: e(HDIM)
    e = evals
    call c_canon_o(P1_bml%ptr, H1_bml%ptr, nocc, beta, e, mu0, m, thresh, HDIM)
  end subroutine prg_canon_response_orig
  !> src/prg_response_mod.F90:849-931
  subroutine prg_canon_response_vector(P1_bml, H1_bml, Nocc, beta, evals, mu0, m, thresh, HDIM)
    type(bml_matrix_t), intent(inout) :: P1_bml, H1_bml
    integer, intent(in) :: HDIM, m
    real(dp), intent(in) :: evals(HDIM), thresh, beta, mu0, nocc
    real(c_double) :: e(HDIM)
    e = evals
    call c_canon_v(P1_bml%ptr, H1_bml%ptr, nocc, beta, e, mu0, m, thresh, HDIM)
  end subroutine prg_canon_response_vector
  !> src/prg_response_mod.F90:1112-1183
  subroutine prg_canon_response_p1_dpdmu(P1_bml, dPdMu, H1_bml, Norbs, beta, Q_bml, evals, mu0, &
       m, HDIM)
    type(bml_matrix_t), intent(inout) :: P1_bml, H1_bml
    type(bml_matrix_t) :: Q_bml
    integer, intent(in) :: HDIM, m, Norbs
    real(dp), intent(in) :: evals(HDIM), beta, mu0
    real(dp), intent(inout) :: dPdMu(HDIM)
    real(c_double) :: e(HDIM), d(HDIM)
    e = evals
    call c_canon_p(P1_bml%ptr, d, H1_bml%ptr, Norbs, beta, Q_bml%ptr, e, mu0, m, HDIM)
    dPdMu = d
  end subroutine prg_canon_response_p1_dpdmu
end module prg_response_mod

module prg_sp2_fermi_mod
  use iso_c_binding
  use bml
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)
  public :: prg_sp2_fermi_init, prg_sp2_fermi_init_norecs, prg_sp2_fermi, sp2_inverse
  interface
    subroutine c_init(h, ns, nocc, ts, thr, oel, tl, x, mu, beta, h1, hn, sg, v) &
         bind(C, name="prg_sp2_fermi_init")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, x
      integer(c_int), value :: ns, v
      real(c_double), value :: nocc, ts, thr, oel, tl
      real(c_double) :: mu, beta, h1, hn
      integer(c_int) :: sg(*)
    end subroutine
    subroutine c_init_nr(h, ns, nocc, ts, thr, oel, tl, x, mu, beta, h1, hn, sg, v) &
         bind(C, name="prg_sp2_fermi_init_norecs")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, x
      integer(c_int) :: ns
      integer(c_int), value :: v
      real(c_double), value :: nocc, ts, thr, oel, tl
      real(c_double) :: mu, beta, h1, hn
      integer(c_int) :: sg(*)
    end subroutine
    subroutine c_fermi(h, os, ns, nocc, mu, beta, h1, hn, sg, thr, eps, tl, x) &
         bind(C, name="prg_sp2_fermi")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, x
      integer(c_int), value :: os, ns
      real(c_double), value :: nocc, thr, eps, tl
      real(c_double) :: mu, beta, h1, hn
      integer(c_int) :: sg(*)
    end subroutine
    function c_inverse(f, mu, h1, hn, ns, sg) bind(C, name="prg_sp2_inverse") result(r)
      import :: c_double, c_int
      real(c_double), value :: f, mu, h1, hn
      integer(c_int), value :: ns
      integer(c_int) :: sg(*)
      real(c_double) :: r
    end function
  end interface
contains
  !> src/prg_sp2_fermi_mod.F90:43-185
  subroutine prg_sp2_fermi_init(h_bml, nsteps, nocc, tscale, threshold, occErrLimit, &
       traceLimit, x_bml, mu, beta, h1, hN, sgnlist)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: x_bml
    integer, intent(in) :: nsteps
    integer, intent(inout) :: sgnlist(:)
    real(dp), intent(in) :: nocc, tscale, threshold, occErrLimit, traceLimit
    real(dp), intent(inout) :: mu, beta, h1, hN
    call c_init(h_bml%ptr, nsteps, nocc, tscale, threshold, occErrLimit, traceLimit, &
         x_bml%ptr, mu, beta, h1, hN, sgnlist, 0)
  end subroutine prg_sp2_fermi_init
  !> src/prg_sp2_fermi_mod.F90:203-387
  subroutine prg_sp2_fermi_init_norecs(h_bml, nsteps, nocc, tscale, threshold, occErrLimit, &
       traceLimit, x_bml, mu, beta, h1, hN, sgnlist, verbose)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: x_bml
    integer, intent(inout) :: nsteps
    integer, intent(inout) :: sgnlist(:)
    real(dp), intent(in) :: nocc, tscale, threshold, occErrLimit, traceLimit
    real(dp), intent(inout) :: mu, beta, h1, hN
    integer, optional :: verbose
    integer :: v
    v = 0
    if (present(verbose)) v = verbose
    call c_init_nr(h_bml%ptr, nsteps, nocc, tscale, threshold, occErrLimit, traceLimit, &
         x_bml%ptr, mu, beta, h1, hN, sgnlist, v)
  end subroutine prg_sp2_fermi_init_norecs
  !> src/prg_sp2_fermi_mod.F90:405-470
  subroutine prg_sp2_fermi(h_bml, osteps, nsteps, nocc, mu, beta, h1, hN, sgnlist, &
       threshold, eps, traceLimit, x_bml)
    type(bml_matrix_t), intent(in) :: h_bml
    type(bml_matrix_t), intent(inout) :: x_bml
    integer, intent(in) :: osteps, nsteps
    integer, intent(in) :: sgnlist(:)
    real(dp), intent(in) :: nocc, threshold, eps, traceLimit
    real(dp), intent(inout) :: beta, h1, hN, mu
    integer, allocatable :: sg(:)
    sg = sgnlist
    call c_fermi(h_bml%ptr, osteps, nsteps, nocc, mu, beta, h1, hN, sg, threshold, eps, &
         traceLimit, x_bml%ptr)
  end subroutine prg_sp2_fermi
  !> src/prg_sp2_fermi_mod.F90:720-745
  function sp2_inverse(f, mu, h1, hN, nsteps, sgnlist) result(ee)
    real(dp), intent(in) :: f, mu, h1, hN
    integer, intent(in) :: nsteps, sgnlist(:)
    real(dp) :: ee
    integer, allocatable :: sg(:)
    sg = sgnlist
    ee = c_inverse(f, mu, h1, hN, nsteps, sg)
  end function sp2_inverse
end module prg_sp2_fermi_mod

module prg_nonortho_mod
  use iso_c_binding
  use bml
  use prg_gpu_cstr
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)
  public :: prg_orthogonalize, prg_deorthogonalize
  interface
    subroutine c_ortho(a, z, o, thr, t, v) bind(C, name="prg_orthogonalize")
      import :: c_ptr, c_double, c_int, c_char
      type(c_ptr), value :: a, z, o
      real(c_double), value :: thr
      character(kind=c_char) :: t(*)
      integer(c_int), value :: v
    end subroutine
    subroutine c_deortho(o, z, a, thr, t, v) bind(C, name="prg_deorthogonalize")
      import :: c_ptr, c_double, c_int, c_char
      type(c_ptr), value :: o, z, a
      real(c_double), value :: thr
      character(kind=c_char) :: t(*)
      integer(c_int), value :: v
    end subroutine
  end interface
contains
  !> src/prg_nonortho_mod.F90:35-68
  subroutine prg_orthogonalize(A_bml, zmat_bml, orthoA_bml, threshold, bml_type, verbose)
    type(bml_matrix_t), intent(inout) :: A_bml, zmat_bml, orthoA_bml
    real(dp), intent(in) :: threshold
    character(len=*), intent(in) :: bml_type
    integer, intent(in) :: verbose
    if (.not. c_associated(orthoA_bml%ptr)) &
         call bml_zero_matrix(bml_type, bml_element_real, dp, bml_get_n(A_bml), &
         bml_get_n(A_bml), orthoA_bml)
    call c_ortho(A_bml%ptr, zmat_bml%ptr, orthoA_bml%ptr, threshold, cstr(bml_type), verbose)
  end subroutine prg_orthogonalize
  !> src/prg_nonortho_mod.F90:81-120
  subroutine prg_deorthogonalize(orthoA_bml, zmat_bml, a_bml, threshold, bml_type, verbose)
    type(bml_matrix_t), intent(inout) :: a_bml
    type(bml_matrix_t), intent(in) :: zmat_bml, orthoA_bml
    real(dp) :: threshold
    character(len=*) :: bml_type
    integer :: verbose
    if (.not. c_associated(a_bml%ptr)) &
         call bml_zero_matrix(bml_type, bml_element_real, dp, bml_get_n(orthoA_bml), &
         bml_get_n(orthoA_bml), a_bml)
    call c_deortho(orthoA_bml%ptr, zmat_bml%ptr, a_bml%ptr, threshold, cstr(bml_type), verbose)
  end subroutine prg_deorthogonalize
end module prg_nonortho_mod

! Graph partitioning (src/prg_graph_mod.F90) on one rank.  The Fortran type
! keeps the reference's fields; node ids are 0-based throughout (the C
! routines and the core+halo lists are 0-based; the reference's own
! prg_readPart adds 1 to the ids it reads, which its 0-based equal partitions
! do not -- here both are 0-based).  The library routines work on the flat C
! struct prg_graph_partitioning_t, filled from and copied back into this type.
module prg_graph_mod
  use iso_c_binding
  use prg_gpu_cstr
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)

  !> src/prg_graph_mod.F90:32-56
  type, public :: subgraph_t
    integer :: part = 0
    integer :: hsize = 0
    integer :: lsize = 0
    integer :: llsize = 0
    integer, allocatable :: core_halo_index(:)
    integer, allocatable :: nodeInPart(:)
  end type subgraph_t

  !> src/prg_graph_mod.F90:58-140
  type, public :: graph_partitioning_t
    character(len=100) :: pname = ""
    integer :: myRank = 0
    integer :: totalProcs = 1
    integer :: totalParts = 0
    integer :: totalNodes = 0
    integer :: totalNodes2 = 0
    integer :: globalPartMin = 0
    integer :: globalPartMax = 0
    integer :: globalPartExtent = 0
    integer, allocatable :: localPartMin(:)
    integer, allocatable :: localPartMax(:)
    integer, allocatable :: localPartExtent(:)
    integer, allocatable :: order(:)
    integer, allocatable :: reorder(:)
    integer :: nparts = 0
    integer :: nx = 0, ny = 0, nz = 0
    integer, allocatable :: nnodesInPart(:)
    integer, allocatable :: nnodesInPartAll(:)
    integer :: pp(100) = 0
    integer :: maxIter = 0
    real(dp) :: ehomo = 0.0_dp
    real(dp) :: elumo = 0.0_dp
    real(dp) :: mineval = 0.0_dp
    real(dp) :: maxeval = 0.0_dp
    real(dp) :: vv(100) = 0.0_dp
    type(subgraph_t), allocatable :: sgraph(:)
  end type graph_partitioning_t

  !> the library's flat prg_graph_partitioning_t (include/progress_gpu.h)
  type, bind(C), public :: prg_c_partitioning_t
    integer(c_int) :: totalParts
    integer(c_int) :: totalNodes
    type(c_ptr) :: nnodesInPart
    type(c_ptr) :: nodeInPart
    type(c_ptr) :: lsize
    type(c_ptr) :: llsize
    type(c_ptr) :: core_halo_index
    integer(c_int) :: pp(100)
    integer(c_int) :: maxIter
    real(c_double) :: ehomo, elumo, mineval, maxeval
    real(c_double) :: vv(100)
  end type prg_c_partitioning_t

  public :: prg_initSubgraph, prg_destroySubgraph, prg_initGraphPartitioning
  public :: prg_destroyGraphPartitioning, prg_equalPartition, prg_filePartition
  public :: prg_fnormGraph
  public :: prg_gp_to_c, prg_gp_from_c, prg_gp_c_free

  interface
    subroutine c_set_partition(gp, np, nn, sizes, nodes) bind(C, name="prg_setPartition")
      import :: prg_c_partitioning_t, c_int
      type(prg_c_partitioning_t) :: gp
      integer(c_int), value :: np, nn
      integer(c_int) :: sizes(*), nodes(*)
    end subroutine
    subroutine c_file_partition(gp, f) bind(C, name="prg_filePartition")
      import :: prg_c_partitioning_t, c_char
      type(prg_c_partitioning_t) :: gp
      character(kind=c_char) :: f(*)
    end subroutine
    subroutine c_destroy(gp) bind(C, name="prg_destroyGraphPartitioning")
      import :: prg_c_partitioning_t
      type(prg_c_partitioning_t) :: gp
    end subroutine
    function c_err() bind(C, name="prg_last_error_code") result(r)
      import :: c_int
      integer(c_int) :: r
    end function
  end interface

contains

  subroutine c_clear(cg)
    type(prg_c_partitioning_t), intent(out) :: cg
    cg%totalParts = 0
    cg%totalNodes = 0
    cg%nnodesInPart = c_null_ptr
    cg%nodeInPart = c_null_ptr
    cg%lsize = c_null_ptr
    cg%llsize = c_null_ptr
    cg%core_halo_index = c_null_ptr
    cg%pp = 0
    cg%maxIter = 0
    cg%ehomo = 0.0_dp
    cg%elumo = 0.0_dp
    cg%mineval = 0.0_dp
    cg%maxeval = 0.0_dp
    cg%vv = 0.0_dp
  end subroutine c_clear

  !> src/prg_graph_mod.F90:148-159
  subroutine prg_initSubgraph(sg, pnum, hsize)
    type(subgraph_t), intent(inout) :: sg
    integer, intent(in) :: pnum, hsize
    sg%part = pnum
    sg%hsize = hsize
    sg%lsize = 0
    sg%llsize = 0
    if (allocated(sg%core_halo_index)) deallocate(sg%core_halo_index)
    allocate(sg%core_halo_index(hsize))
  end subroutine prg_initSubgraph

  !> src/prg_graph_mod.F90:164-172
  subroutine prg_destroySubgraph(sg)
    type(subgraph_t), intent(inout) :: sg
    if (allocated(sg%core_halo_index)) deallocate(sg%core_halo_index)
    if (allocated(sg%nodeInPart)) deallocate(sg%nodeInPart)
  end subroutine prg_destroySubgraph

  !> src/prg_graph_mod.F90:179-260, one rank: every part is local
  subroutine prg_initGraphPartitioning(gp, pname, np, nnodes, nnodes2)
    type(graph_partitioning_t), intent(inout) :: gp
    character(len=*), intent(in) :: pname
    integer, intent(in) :: np, nnodes, nnodes2
    gp%myRank = 0
    gp%pname = pname
    gp%totalProcs = 1
    gp%totalParts = np
    gp%totalNodes = nnodes
    gp%totalNodes2 = nnodes2
    gp%globalPartMin = 1
    gp%globalPartMax = np
    gp%globalPartExtent = np
    allocate(gp%localPartMin(1), gp%localPartMax(1), gp%localPartExtent(1))
    gp%localPartMin(1) = 1
    gp%localPartMax(1) = np
    gp%localPartExtent(1) = np
    gp%nparts = np
    allocate(gp%nnodesInPart(np), gp%nnodesInPartAll(np), gp%sgraph(np))
    allocate(gp%order(nnodes), gp%reorder(nnodes))
    gp%maxIter = 0
    gp%mineval = 0.0_dp
    gp%maxeval = 0.0_dp
  end subroutine prg_initGraphPartitioning

  !> src/prg_graph_mod.F90:268-290
  subroutine prg_destroyGraphPartitioning(gp)
    type(graph_partitioning_t), intent(inout) :: gp
    integer :: i
    if (allocated(gp%localPartMin)) deallocate(gp%localPartMin)
    if (allocated(gp%localPartMax)) deallocate(gp%localPartMax)
    if (allocated(gp%localPartExtent)) deallocate(gp%localPartExtent)
    if (allocated(gp%order)) deallocate(gp%order)
    if (allocated(gp%reorder)) deallocate(gp%reorder)
    if (allocated(gp%nnodesInPart)) deallocate(gp%nnodesInPart)
    if (allocated(gp%nnodesInPartAll)) deallocate(gp%nnodesInPartAll)
    if (allocated(gp%sgraph)) then
      do i = 1, size(gp%sgraph)
        call prg_destroySubgraph(gp%sgraph(i))
      enddo
      deallocate(gp%sgraph)
    endif
  end subroutine prg_destroyGraphPartitioning

  !> src/prg_graph_mod.F90:359-399: ceiling(nnodes / nodesPerPart) parts of
  !> consecutive 0-based node ids
  subroutine prg_equalPartition(gp, nodesPerPart, nnodes)
    type(graph_partitioning_t), intent(inout) :: gp
    integer, intent(in) :: nodesPerPart, nnodes
    integer :: i, j, np, psize
    if (nnodes <= nodesPerPart) then
      np = 1
    else
      np = ceiling(real(nnodes) / real(nodesPerPart))
    endif
    call prg_destroyGraphPartitioning(gp)
    call prg_initGraphPartitioning(gp, "equalParts", np, nnodes, nnodes)
    do i = 1, np
      call prg_initSubgraph(gp%sgraph(i), i, nnodes)
      if (i * nodesPerPart <= nnodes) then
        psize = nodesPerPart
      else
        psize = nnodes - nodesPerPart * (i - 1)
      endif
      allocate(gp%sgraph(i)%nodeInPart(psize))
      do j = 1, psize
        gp%sgraph(i)%nodeInPart(j) = (i - 1) * nodesPerPart + j - 1
      enddo
      gp%nnodesInPart(i) = psize
      gp%nnodesInPartAll(i) = psize
    enddo
  end subroutine prg_equalPartition

  !> src/prg_graph_mod.F90:880-929 (the library's reader; ids kept 0-based)
  subroutine prg_filePartition(gp, partFile)
    type(graph_partitioning_t), intent(inout) :: gp
    character(len=*), intent(in) :: partFile
    type(prg_c_partitioning_t) :: cg
    integer, pointer :: sizes(:), nodes(:)
    type(c_ptr), pointer :: lists(:)
    character(len=100) :: pname
    integer :: i, np, u
    call c_clear(cg)
    call c_file_partition(cg, cstr(partFile))
    if (c_err() /= 0) return
    pname = "filePart"
    open(newunit=u, file=partFile, status='old')
    read(u, *) pname
    close(u)
    np = cg%totalParts
    call prg_destroyGraphPartitioning(gp)
    call prg_initGraphPartitioning(gp, trim(pname), np, cg%totalNodes, cg%totalNodes)
    call c_f_pointer(cg%nnodesInPart, sizes, [np])
    call c_f_pointer(cg%nodeInPart, lists, [np])
    do i = 1, np
      call prg_initSubgraph(gp%sgraph(i), i, cg%totalNodes)
      allocate(gp%sgraph(i)%nodeInPart(sizes(i)))
      if (sizes(i) > 0) then
        call c_f_pointer(lists(i), nodes, [sizes(i)])
        gp%sgraph(i)%nodeInPart(:) = nodes(:)
      endif
      gp%nnodesInPart(i) = sizes(i)
      gp%nnodesInPartAll(i) = sizes(i)
    enddo
    call c_destroy(cg)
  end subroutine prg_filePartition

  !> src/prg_graph_mod.F90:933-971 (one rank): vv <- sqrt(vv)
  subroutine prg_fnormGraph(gp)
    type(graph_partitioning_t), intent(inout) :: gp
    integer :: i
    do i = 1, gp%maxIter
      gp%vv(i) = sqrt(gp%vv(i))
    enddo
  end subroutine prg_fnormGraph

  !> fills the library's flat struct from gp (cores, SP2 sequence, bounds, vv)
  subroutine prg_gp_to_c(gp, cg)
    type(graph_partitioning_t), intent(in) :: gp
    type(prg_c_partitioning_t), intent(inout) :: cg
    integer(c_int), allocatable :: sizes(:), nodes(:)
    integer :: i, k
    call c_clear(cg)
    allocate(sizes(gp%totalParts))
    do i = 1, gp%totalParts
      sizes(i) = gp%nnodesInPart(i)
    enddo
    allocate(nodes(max(1, sum(sizes))))
    k = 0
    do i = 1, gp%totalParts
      nodes(k + 1:k + sizes(i)) = gp%sgraph(i)%nodeInPart(1:sizes(i))
      k = k + sizes(i)
    enddo
    call c_set_partition(cg, gp%totalParts, gp%totalNodes, sizes, nodes)
    cg%pp = gp%pp
    cg%maxIter = gp%maxIter
    cg%ehomo = gp%ehomo
    cg%elumo = gp%elumo
    cg%mineval = gp%mineval
    cg%maxeval = gp%maxeval
    cg%vv = gp%vv
  end subroutine prg_gp_to_c

  !> copies the core+halo lists (lsize, llsize) and vv back into gp
  subroutine prg_gp_from_c(cg, gp)
    type(prg_c_partitioning_t), intent(in) :: cg
    type(graph_partitioning_t), intent(inout) :: gp
    integer, pointer :: ls(:), lls(:), idx(:)
    type(c_ptr), pointer :: lists(:)
    integer :: i, np
    np = cg%totalParts
    gp%vv = cg%vv
    if (.not. c_associated(cg%lsize)) return
    call c_f_pointer(cg%lsize, ls, [np])
    call c_f_pointer(cg%llsize, lls, [np])
    call c_f_pointer(cg%core_halo_index, lists, [np])
    do i = 1, np
      gp%sgraph(i)%lsize = ls(i)
      gp%sgraph(i)%llsize = lls(i)
      if (allocated(gp%sgraph(i)%core_halo_index)) deallocate(gp%sgraph(i)%core_halo_index)
      allocate(gp%sgraph(i)%core_halo_index(max(1, ls(i))))
      if (ls(i) > 0 .and. c_associated(lists(i))) then
        call c_f_pointer(lists(i), idx, [ls(i)])
        gp%sgraph(i)%core_halo_index(1:ls(i)) = idx(:)
      endif
    enddo
  end subroutine prg_gp_from_c

  subroutine prg_gp_c_free(cg)
    type(prg_c_partitioning_t), intent(inout) :: cg
    call c_destroy(cg)
  end subroutine prg_gp_c_free

end module prg_graph_mod

! Subgraph SP2 loop (src/prg_subgraphloop_mod.F90) over the library's batched
! GPU loop.
module prg_subgraphloop_mod
  use iso_c_binding
  use bml
  use prg_graph_mod
  implicit none
  private
  public :: prg_subgraphSP2Loop, prg_getPartitionHalosFromGraph
  interface
    subroutine c_loop(h, g, r, gp, thr) bind(C, name="prg_subgraphSP2Loop")
      import :: c_ptr, c_double, prg_c_partitioning_t
      type(c_ptr), value :: h, g, r
      type(prg_c_partitioning_t) :: gp
      real(c_double), value :: thr
    end subroutine
    subroutine c_halos(gp, g, dj) bind(C, name="prg_getPartitionHalosFromGraph")
      import :: c_ptr, c_int, prg_c_partitioning_t
      type(prg_c_partitioning_t) :: gp
      type(c_ptr), value :: g
      integer(c_int), value :: dj
    end subroutine
  end interface
contains
  !> src/prg_subgraphloop_mod.F90:36-126 (one rank): core+halo lists from G
  !> and H, every part purified with gp%pp / gp%maxIter on the GPU, core rows
  !> into rho_bml, gp%vv accumulated then square-rooted (prg_fnormGraph)
  subroutine prg_subgraphSP2Loop(h_bml, g_bml, rho_bml, gp, threshold)
    type(bml_matrix_t), intent(in) :: h_bml, g_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    type(graph_partitioning_t), intent(inout) :: gp
    real(kind(1.0d0)), intent(in) :: threshold
    type(prg_c_partitioning_t) :: cg
    call prg_gp_to_c(gp, cg)
    call c_loop(h_bml%ptr, g_bml%ptr, rho_bml%ptr, cg, threshold)
    call prg_gp_from_c(cg, gp)
    call prg_gp_c_free(cg)
  end subroutine prg_subgraphSP2Loop

  !> src/prg_subgraphloop_mod.F90:336-368
  subroutine prg_getPartitionHalosFromGraph(gp, g_bml, djflag)
    type(graph_partitioning_t), intent(inout) :: gp
    type(bml_matrix_t), intent(in) :: g_bml
    logical, intent(in) :: djflag
    type(prg_c_partitioning_t) :: cg
    integer(c_int) :: dj
    dj = 0
    if (djflag) dj = 1
    call prg_gp_to_c(gp, cg)
    call c_halos(cg, g_bml%ptr, dj)
    call prg_gp_from_c(cg, gp)
    call prg_gp_c_free(cg)
  end subroutine prg_getPartitionHalosFromGraph
end module prg_subgraphloop_mod

module prg_graphsolver_mod
  use iso_c_binding
  use bml
  implicit none
  private
  integer, parameter :: dp = kind(1.0d0)
  public :: prg_build_densityGP_T0, prg_build_zmatGP
  interface
    subroutine c_zmat_gp(s, g, z, thr, nparts, v) bind(C, name="prg_build_zmatGP")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: s, g, z
      real(c_double), value :: thr
      integer(c_int), value :: nparts, v
    end subroutine
    subroutine c_gp_t0(h, g, r, thr, bndfil, ef, nparts, v) bind(C, name="prg_build_densityGP_T0")
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: h, g, r
      real(c_double), value :: thr, bndfil, ef
      integer(c_int), value :: nparts, v
    end subroutine
    function c_gp_t0_ef() bind(C, name="prg_build_densityGP_T0_ef") result(r)
      import :: c_double
      real(c_double) :: r
    end function
  end interface
contains
  !> src/prg_graphsolver_mod.F90:39-191 (partition by the library's BFS
  !> partitioner in place of METIS; Ef updated as in Fortran)
  subroutine prg_build_densityGP_T0(ham_bml, g_bml, rho_bml, threshold, bndfil, Ef, nparts, &
       verbose)
    type(bml_matrix_t), intent(in) :: ham_bml, g_bml
    type(bml_matrix_t), intent(inout) :: rho_bml
    real(dp) :: bndfil, threshold
    integer, intent(inout) :: nparts
    real(dp), intent(inout) :: Ef
    integer, optional, intent(in) :: verbose
    integer :: v
    v = 0
    if (present(verbose)) v = verbose
    if (.not. c_associated(rho_bml%ptr)) &
         call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, bml_get_n(ham_bml), &
         bml_get_n(ham_bml), rho_bml)
    call c_gp_t0(ham_bml%ptr, g_bml%ptr, rho_bml%ptr, threshold, bndfil, Ef, nparts, v)
    Ef = c_gp_t0_ef()
  end subroutine prg_build_densityGP_T0

  !> src/prg_graphsolver_mod.F90:202-331: Z ~ S^-1/2 from the parts' Loewdin
  !> factors (partition by the library's BFS partitioner in place of METIS)
  subroutine prg_build_zmatGP(over_bml, g_bml, zmat_bml, threshold, nparts, verbose)
    type(bml_matrix_t), intent(in) :: over_bml, g_bml
    type(bml_matrix_t), intent(inout) :: zmat_bml
    real(dp) :: threshold
    integer, intent(inout) :: nparts
    integer, optional, intent(in) :: verbose
    integer :: v
    v = 0
    if (present(verbose)) v = verbose
    if (.not. c_associated(zmat_bml%ptr)) &
         call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, bml_get_n(over_bml), &
         bml_get_n(over_bml), zmat_bml)
    call c_zmat_gp(over_bml%ptr, g_bml%ptr, zmat_bml%ptr, threshold, nparts, v)
  end subroutine prg_build_zmatGP
end module prg_graphsolver_mod

! src/sp2_tensorcore/src/prg_sp2_tensorcore_mod.F90: the mixed-precision SP2
! entry with host arrays (H single precision in, D = X double precision out).
! The C symbol takes sp2conv as one character by value (the C prototype's
! form); the library does not use it.
module prg_sp2_tensorcore_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: prg_sp2_tensorcore_f

  interface
    subroutine c_sp2_tensorcore(n, h, d, eps, bndfil, mn, mx, conv, tol, v) &
         bind(C, name="prg_sp2_tensorcore")
      import :: c_ptr, c_int, c_float, c_char
      integer(c_int), value :: n, mn, mx, v
      type(c_ptr), value :: h, d
      real(c_float), value :: eps, bndfil, tol
      character(kind=c_char), value :: conv
    end subroutine
  end interface

contains

  subroutine prg_sp2_tensorcore_f(N, H, D, eps, bndfil, minsp2iter, maxsp2iter, sp2conv, &
       idemtol, verbose)
    integer(c_int), intent(in) :: N, minsp2iter, maxsp2iter, verbose
    real(c_double), target, intent(inout) :: D(*)
    real(c_float), target, intent(in) :: H(*)
    real(c_float), intent(in) :: eps, bndfil, idemtol
    character(c_char), intent(in) :: sp2conv
    call c_sp2_tensorcore(N, c_loc(H), c_loc(D), eps, bndfil, minsp2iter, maxsp2iter, &
         sp2conv, idemtol, verbose)
  end subroutine prg_sp2_tensorcore_f

end module prg_sp2_tensorcore_mod
