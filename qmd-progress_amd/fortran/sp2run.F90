! SP2 driver in the shape of examples/sp2run/src/main.F90: reads a Matrix
! Market Hamiltonian, computes the density matrix with one of the SP2
! variants through the GPU library and writes it out.
!
! usage: sp2run <hamiltonian.mtx> <rho_out.mtx> [alg2|alg1|basic] [bndfil] [Rel|Abs] [idemtol]
program sp2run
  use bml
  use prg_sp2_mod
  use prg_densitymatrix_mod
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  type(bml_matrix_t) :: ham_bml, rho_bml
  character(len=256) :: hfile, rfile, arg
  character(len=8) :: method, sp2conv
  real(dp) :: bndfil, idemtol, threshold, idem
  integer :: norb, minsp2iter, maxsp2iter, ierr

  if (command_argument_count() < 2) then
    write(*, '(a)') "usage: sp2run <hamiltonian.mtx> <rho_out.mtx> [alg2|alg1|basic] [bndfil]" // &
         " [Rel|Abs] [idemtol]"
    stop 2
  end if
  call get_command_argument(1, hfile)
  call get_command_argument(2, rfile)
  method = "alg2"
  bndfil = 0.5_dp
  sp2conv = "Rel"
  idemtol = 1.0e-10_dp
  if (command_argument_count() >= 3) call get_command_argument(3, method)
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    read(arg, *) bndfil
  end if
  if (command_argument_count() >= 5) call get_command_argument(5, sp2conv)
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg)
    read(arg, *) idemtol
  end if
  threshold = 0.0_dp
  minsp2iter = 10
  maxsp2iter = 200

  norb = mtx_size(hfile)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, norb, norb, ham_bml)
  call bml_read_matrix(ham_bml, hfile)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, norb, norb, rho_bml)

  select case (trim(method))
  case ("alg1")
    call prg_sp2_alg1(ham_bml, rho_bml, threshold, bndfil, minsp2iter, maxsp2iter, sp2conv, &
         idemtol, 1)
  case ("basic")
    call prg_sp2_basic(ham_bml, rho_bml, threshold, bndfil, minsp2iter, maxsp2iter, sp2conv, &
         idemtol, 1)
  case default
    call prg_sp2_alg2(ham_bml, rho_bml, threshold, bndfil, minsp2iter, maxsp2iter, sp2conv, &
         idemtol, 1)
  end select
  ierr = prg_gpu_last_error()
  if (ierr /= 0) then
    write(*, '(a,i0)') "sp2run: library error ", ierr
    stop 1
  end if

  call bml_scale(0.5_dp, rho_bml)
  call prg_check_idempotency(rho_bml, threshold, idem)
  call bml_scale(2.0_dp, rho_bml)
  write(*, '(a,i0,a,es22.15,a,es10.3)') "sp2run: norb=", norb, " trace=", bml_trace(rho_bml), &
       " idempotency=", idem
  call bml_write_matrix(rho_bml, rfile)
  call bml_deallocate(ham_bml)
  call bml_deallocate(rho_bml)

contains

  ! Row count from a Matrix Market header (first line after the comments).
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
end program sp2run
