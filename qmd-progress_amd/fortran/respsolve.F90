! Driver for the response and linear-solver routines through the Fortran
! modules: reads H, H1 and an SPD matrix A (Matrix Market), writes
!   <prefix>_rs.mtx   prg_compute_response_RS(H, H1)        (bndfil 0.5)
!   <prefix>_fd.mtx   prg_compute_response_FD(H, H1, 1e-4)  (bndfil 0.5)
!   <prefix>_pcg.mtx  prg_pcg(A, X0 = 0, B = I)             (tolerance 1e-16)
!   <prefix>_ns.mtx   prg_newtonschulz(A, Ai0 = I / 4)      (tolerance 1e-10)
!   <prefix>_ifz.mtx  prg_implicit_fermi_zero(H, 6 steps, mu 0, CG, tolerance 1e-18)
!   <prefix>_ifr0.mtx, _ifr1.mtx, _ifr2.mtx
!                     prg_implicit_fermi_response(H, H1, H2 = H1 / 2, order 2,
!                     6 steps, mu0 0, beta 10, nocc n / 2, occ_tol 1e-8)
!
! usage: respsolve <h.mtx> <h1.mtx> <a.mtx> <prefix>
program respsolve
  use bml
  use prg_response_mod
  use prg_implicit_fermi_mod
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  type(bml_matrix_t) :: h_bml, h1_bml, a_bml, rsp_bml, x_bml, b_bml, ai_bml
  type(bml_matrix_t) :: aux1, aux2, aux3, aux4, h2_bml, p0_bml, p1_bml, p2_bml, p3_bml
  real(dp), allocatable :: mu(:)
  real(dp) :: mu0, ifz_tol
  character(len=256) :: hfile, h1file, afile, prefix
  integer :: n, iters, ierr

  if (command_argument_count() < 4) then
    write(*, '(a)') "usage: respsolve <h.mtx> <h1.mtx> <a.mtx> <prefix>"
    stop 2
  end if
  call get_command_argument(1, hfile)
  call get_command_argument(2, h1file)
  call get_command_argument(3, afile)
  call get_command_argument(4, prefix)
  n = mtx_size(hfile)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, h_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, h1_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, a_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, rsp_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, x_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, aux1)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, aux2)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, aux3)
  call bml_identity_matrix(bml_matrix_dense, bml_element_real, dp, n, n, aux4)
  call bml_identity_matrix(bml_matrix_dense, bml_element_real, dp, n, n, b_bml)
  call bml_identity_matrix(bml_matrix_dense, bml_element_real, dp, n, n, ai_bml)
  call bml_read_matrix(h_bml, hfile)
  call bml_read_matrix(h1_bml, h1file)
  call bml_read_matrix(a_bml, afile)

  call prg_compute_response_RS(h_bml, h1_bml, rsp_bml, 0.0_dp, 0.5_dp, 0.0_dp, 0)
  call check("prg_compute_response_RS")
  call bml_write_matrix(rsp_bml, trim(prefix) // "_rs.mtx")
  call prg_compute_response_FD(h_bml, h1_bml, rsp_bml, 1.0e-4_dp, 0.5_dp, 0.0_dp, 0)
  call check("prg_compute_response_FD")
  call bml_write_matrix(rsp_bml, trim(prefix) // "_fd.mtx")
  call prg_pcg(a_bml, x_bml, b_bml, aux1, aux2, 1.0e-16_dp, 0.0_dp)
  call check("prg_pcg")
  call bml_write_matrix(x_bml, trim(prefix) // "_pcg.mtx")
  call bml_scale(0.25_dp, ai_bml)
  call prg_newtonschulz(a_bml, ai_bml, aux1, aux2, aux3, aux4, 1.0e-10_dp, 0.0_dp, iters)
  call check("prg_newtonschulz")
  call bml_write_matrix(ai_bml, trim(prefix) // "_ns.mtx")
  write(*, '(a,i0,a,i0)') "respsolve: n=", n, " newton-schulz iterations=", iters

  ifz_tol = 1.0e-18_dp
  call prg_implicit_fermi_zero(h_bml, rsp_bml, 6, 0.0_dp, 0, 0.0_dp, ifz_tol)
  call check("prg_implicit_fermi_zero")
  call bml_write_matrix(rsp_bml, trim(prefix) // "_ifz.mtx")

  call bml_copy_new(h1_bml, h2_bml)
  call bml_scale(0.5_dp, h2_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, p0_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, p1_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, p2_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, n, n, p3_bml)
  mu0 = 0.0_dp
  call prg_implicit_fermi_response(h_bml, h1_bml, h2_bml, h2_bml, p0_bml, p1_bml, p2_bml, &
       p3_bml, 6, mu0, mu, 10.0_dp, 0.5_dp * n, 1.0e-8_dp, 1.0e-10_dp, 2, 0.0_dp)
  call check("prg_implicit_fermi_response")
  call bml_write_matrix(p0_bml, trim(prefix) // "_ifr0.mtx")
  call bml_write_matrix(p1_bml, trim(prefix) // "_ifr1.mtx")
  call bml_write_matrix(p2_bml, trim(prefix) // "_ifr2.mtx")
  write(*, '(a,es24.16,a,es24.16,a,es24.16)') "respsolve: mu0=", mu0, " mu1=", mu(1), &
       " mu2=", mu(2)

contains

  subroutine check(who)
    character(len=*), intent(in) :: who
    ierr = prg_gpu_last_error()
    if (ierr /= 0) then
      write(*, '(a,a,a,i0)') "respsolve: ", who, " library error ", ierr
      stop 1
    end if
  end subroutine check

  integer function mtx_size(fname)
    character(len=*), intent(in) :: fname
    character(len=512) :: line
    integer :: u, nr
    open(newunit=u, file=fname, status="old", action="read")
    do
      read(u, '(a)') line
      if (line(1:1) /= "%") exit
    end do
    read(line, *) nr
    close(u)
    mtx_size = nr
  end function mtx_size
end program respsolve
