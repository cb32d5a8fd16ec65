! Graph-partitioned SP2 driver in the shape of examples/gptest/gptest.F90:
! full SP2 (genseq) for the graph and the HOMO-LUMO estimate, thresholded
! graph, equal partition, SP2 sequence from the estimate, subgraph SP2 loop
! through the GPU library, then band energies and ||rho_full - rho_graph||_F.
!
! usage: gptest <hamiltonian.mtx> <nodesPerPart> [gthreshold] [errlimit] [bndfil] [rho_out.mtx]
! Prints "key = value" lines (one per quantity) for scripted checks.
program gptest
  use bml
  use prg_progress_mod
  use prg_sp2_mod
  use prg_graph_mod
  use prg_subgraphloop_mod
  use prg_homolumo_mod
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  type(bml_matrix_t) :: h_bml, rho_bml, g_bml, fullg_bml
  type(graph_partitioning_t) :: gp
  character(len=256) :: hfile, rfile, arg
  integer :: N, i, icount, nodesPerPart, minsp2iter, maxsp2iter, ierr
  integer :: pp(100)
  real(dp) :: vv(100)
  real(dp) :: mineval, maxeval, trace_mult, trace_multg, ehomo, elumo, egap, frobNorm
  real(dp) :: gthreshold, errlimit, bndfil, threshold, sp2tol
  real(dp), allocatable :: gbnd(:)

  if (command_argument_count() < 2) then
    write(*, '(a)') "usage: gptest <hamiltonian.mtx> <nodesPerPart> [gthreshold] [errlimit]" // &
         " [bndfil] [rho_out.mtx]"
    stop 2
  end if
  call get_command_argument(1, hfile)
  call get_command_argument(2, arg)
  read(arg, *) nodesPerPart
  gthreshold = 1.0e-5_dp
  errlimit = 1.0e-12_dp
  bndfil = 0.5_dp
  rfile = ""
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) gthreshold
  end if
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    read(arg, *) errlimit
  end if
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg)
    read(arg, *) bndfil
  end if
  if (command_argument_count() >= 6) call get_command_argument(6, rfile)
  threshold = 0.0_dp
  sp2tol = 1.0e-10_dp
  minsp2iter = 25
  maxsp2iter = 100

  call prg_progress_init()
  N = mtx_size(hfile)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, N, N, h_bml)
  call bml_zero_matrix(bml_matrix_dense, bml_element_real, dp, N, N, rho_bml)
  call bml_read_matrix(h_bml, hfile)
  call bml_zero_matrix(bml_matrix_ellpack, bml_element_real, dp, N, N, g_bml)

  ! full SP2 for the graph (examples/gptest/gptest.F90:57-60)
  call prg_sp2_alg2_genseq(h_bml, g_bml, threshold, bndfil, minsp2iter, maxsp2iter, "Rel", &
       sp2tol, pp, icount, vv)
  call check("prg_sp2_alg2_genseq")
  trace_mult = bml_trace_mult(h_bml, g_bml)
  write(*, '(a, i0)') "icount = ", icount
  write(*, '(a, es24.16)') "trace_hp_sp2 = ", trace_mult

  ! HOMO-LUMO estimate from the sequence (:81-89)
  call bml_gershgorin(h_bml, gbnd)
  mineval = gbnd(1)
  maxeval = gbnd(2)
  call prg_homolumogap(vv, icount, pp, mineval, maxeval, ehomo, elumo, egap)
  write(*, '(a, es24.16)') "mineval = ", mineval
  write(*, '(a, es24.16)') "maxeval = ", maxeval
  write(*, '(a, es24.16)') "egap_sp2 = ", egap

  ! thresholded graph, equal parts, SP2 sequence (:95-110)
  call bml_copy_new(g_bml, fullg_bml)
  call bml_threshold(g_bml, gthreshold)
  call prg_equalPartition(gp, nodesPerPart, N)
  gp%mineval = mineval
  gp%maxeval = maxeval
  call prg_sp2sequence(gp%pp, gp%maxIter, gp%mineval, gp%maxeval, ehomo, elumo, errlimit)
  write(*, '(a, i0)') "parts = ", gp%totalParts
  write(*, '(a, i0)') "maxIter = ", gp%maxIter

  ! subgraph SP2 loop (:113-117)
  gp%vv = 0.0_dp
  call prg_subgraphSP2Loop(h_bml, g_bml, rho_bml, gp, threshold)
  call check("prg_subgraphSP2Loop")
  call prg_homolumogap(gp%vv, gp%maxIter, gp%pp, gp%mineval, gp%maxeval, ehomo, elumo, egap)
  write(*, '(a, es24.16)') "egap_graph = ", egap
  write(*, '(a, i0)') "lsize_max = ", maxval(gp%sgraph(:)%lsize)

  ! band energies and the distance between the two densities (:131-151)
  trace_multg = bml_trace_mult(h_bml, rho_bml)
  frobNorm = bml_fnorm2(fullg_bml, rho_bml)
  write(*, '(a, es24.16)') "trace_hp_graph = ", trace_multg
  write(*, '(a, es24.16)') "frob = ", frobNorm
  if (len_trim(rfile) > 0) call bml_write_matrix(rho_bml, rfile)

  call bml_deallocate(fullg_bml)
  call bml_deallocate(g_bml)
  call bml_deallocate(h_bml)
  call bml_deallocate(rho_bml)
  call prg_destroyGraphPartitioning(gp)
  call prg_progress_shutdown()

contains

  subroutine check(what)
    character(len=*), intent(in) :: what
    ierr = prg_gpu_last_error()
    if (ierr /= 0) then
      write(*, '(a, a, a, i0)') "gptest: ", what, " failed: library error ", ierr
      stop 1
    end if
  end subroutine check

  integer function mtx_size(fname)
    character(len=*), intent(in) :: fname
    character(len=512) :: line
    integer :: u, nr, nc, nnz
    open(newunit=u, file=fname, status='old')
    do
      read(u, '(a)') line
      if (line(1:1) /= '%') exit
    end do
    read(line, *) nr, nc, nnz
    close(u)
    mtx_size = nr
  end function mtx_size

end program gptest
