! exrand_gpu.f90 -- a Fortran caller of the drop-in on a 1x1 grid, in the shape of the reference's
! test drivers (testing/EXRAND1.f:200-340, testing/EXRAND2.f:240-410): it calls PDGGHRD, PDHGEQZ
! and PILAENVX through the ordinary Fortran ABI (by reference, hidden character lengths) and checks
! the result with PCOMPRES-style residuals (tools/PQZHLPS.f:19-319).  It is linked against
! libpdhgeqz_f.so + libgqz.so instead of the reference's objects; no line of it knows about HIP.
!
! Usage:  exrand_gpu args          argument checks and workspace queries only (no device needed)
!         exrand_gpu run N [SEED]  dense A, upper triangular B of order N: PDGGHRD, Q transposed
!                                  as EXRAND2.f:346-356 does, PDHGEQZ with COMPQ = COMPZ = 'V',
!                                  then ||Q^T A Z - S||_F / ||A||_F, the same for B, and
!                                  ||I - Q^T Q||_F / (N eps), ||I - Z^T Z||_F / (N eps).
! Every result is printed as "key value" on a line of its own; the exit status is 0 when all
! checks hold, 1 when a residual is too large and 2 when a routine returned INFO /= 0.
program exrand_gpu
  implicit none
  integer, external :: pilaenvx
  character(len=16) :: mode, arg
  integer :: n, seed

  call get_command_argument(1, mode)
  if (trim(mode) == 'args') then
    call check_arguments()
  else if (trim(mode) == 'run') then
    call get_command_argument(2, arg)
    read (arg, *) n
    seed = 1
    if (command_argument_count() >= 3) then
      call get_command_argument(3, arg)
      read (arg, *) seed
    end if
    call run_pencil(n, seed)
  else
    write (*, '(a)') 'usage: exrand_gpu args | exrand_gpu run N [SEED]'
    stop 2
  end if

contains

  subroutine set_desc(desc, n, nb)
    integer, intent(out) :: desc(9)
    integer, intent(in) :: n, nb
    ! DTYPE, CTXT, M, N, MB, NB, RSRC, CSRC, LLD (PDHGEQZ.f:88-127)
    desc = (/ 1, 0, n, n, nb, nb, 0, 0, max(1, n) /)
  end subroutine set_desc

  subroutine check_arguments()
    integer, parameter :: n = 8, nb = 60
    integer :: desc(9), bad(9), info, iwork(4 * n)
    double precision :: h(n, n), t(n, n), q(n, n), z(n, n), ar(n), ai(n), be(n), dwork(n * nb)
    h = 0.0d0; t = 0.0d0; q = 0.0d0; z = 0.0d0
    call set_desc(desc, n, nb)
    call pdhgeqz('X', 'I', 'I', n, 1, n, h, desc, t, desc, ar, ai, be, q, desc, z, desc, &
                 dwork, n * nb, iwork, 2 * n, info)
    write (*, '(a, i0)') 'pdhgeqz_bad_job ', info
    call pdhgeqz('S', 'I', 'I', n, 0, n, h, desc, t, desc, ar, ai, be, q, desc, z, desc, &
                 dwork, n * nb, iwork, 2 * n, info)
    write (*, '(a, i0)') 'pdhgeqz_bad_ilo ', info
    call set_desc(bad, n, 4)
    call pdhgeqz('S', 'I', 'I', n, 1, n, h, bad, t, bad, ar, ai, be, q, bad, z, bad, &
                 dwork, n * nb, iwork, 2 * n, info)
    write (*, '(a, i0)') 'pdhgeqz_small_nb ', info
    call pdhgeqz('S', 'I', 'I', n, 1, n, h, desc, t, desc, ar, ai, be, q, desc, z, desc, &
                 dwork, -1, iwork, -1, info)
    write (*, '(a, i0, 1x, i0, 1x, i0)') 'pdhgeqz_query ', info, nint(dwork(1)), iwork(1)
    call pdhgeqz('S', 'I', 'I', n, 1, n, h, desc, t, desc, ar, ai, be, q, desc, z, desc, &
                 dwork, 10, iwork, 2 * n, info)
    write (*, '(a, i0)') 'pdhgeqz_small_ldwork ', info
    call pdgghrd('I', 'I', n, 1, n, h, desc, t, desc, q, desc, z, desc, dwork, -1, info)
    write (*, '(a, i0, 1x, i0)') 'pdgghrd_query ', info, nint(dwork(1))
    call pdgghrd('X', 'I', n, 1, n, h, desc, t, desc, q, desc, z, desc, dwork, nb, info)
    write (*, '(a, i0)') 'pdgghrd_bad_compq ', info
    call pdgghrd('I', 'I', n, 1, n + 1, h, desc, t, desc, q, desc, z, desc, dwork, nb, info)
    write (*, '(a, i0)') 'pdgghrd_bad_ihi ', info
    call kkqzconf(epsilon(1.0d0))
    write (*, '(a, i0, 1x, i0, 1x, i0)') 'pilaenvx_16000 ', &
        pilaenvx(0, 53, 'PDHGEQZ0', ' ', 16000, 1, 16000, 0), &
        pilaenvx(0, 51, 'PDHGEQZ0', ' ', 16000, 1, 16000, 0), &
        pilaenvx(0, 52, 'PDHGEQZ0', ' ', 16000, 1, 16000, 0)
  end subroutine check_arguments

  ! standard normal samples by the Box-Muller transform of the intrinsic generator
  subroutine randn(x, m)
    integer, intent(in) :: m
    double precision, intent(out) :: x(m)
    double precision :: u(2), two_pi
    integer :: i
    two_pi = 8.0d0 * atan(1.0d0)
    do i = 1, m
      call random_number(u)
      x(i) = sqrt(-2.0d0 * log(1.0d0 - u(1))) * cos(two_pi * u(2))
    end do
  end subroutine randn

  subroutine run_pencil(n, seed)
    integer, intent(in) :: n, seed
    integer :: desc(9), info, nb, i, j, nseed, ldwork, liwork
    integer, allocatable :: iwork(:), sd(:)
    double precision, allocatable :: a0(:, :), b0(:, :), s(:, :), t(:, :), q(:, :), z(:, :), &
                                     w(:, :), ar(:), ai(:), be(:), dwork(:)
    double precision :: resa, resb, resq, resz, eps
    logical :: ok

    nb = 60
    if (n > 2000) nb = 130                       ! EXRAND1.f:116-120
    call random_seed(size=nseed)
    allocate (sd(nseed))
    sd = (/ (seed * 7919 + 104729 * i, i = 1, nseed) /)
    call random_seed(put=sd)
    allocate (a0(n, n), b0(n, n), s(n, n), t(n, n), q(n, n), z(n, n), w(n, n), ar(n), ai(n), be(n))
    call randn(a0, n * n)
    call randn(b0, n * n)
    do j = 1, n
      do i = j + 1, n
        b0(i, j) = 0.0d0
      end do
    end do
    s = a0
    t = b0
    call set_desc(desc, n, nb)

    ! Hessenberg-triangular reduction; PDGGHRD returns Q transposed (PDGGHRD.f:25-27)
    allocate (dwork(max(1, nb)))
    call pdgghrd('I', 'I', n, 1, n, s, desc, t, desc, q, desc, z, desc, dwork, nb, info)
    write (*, '(a, i0)') 'pdgghrd_info ', info
    if (info /= 0) stop 2
    q = transpose(q)
    deallocate (dwork)

    ! workspace query, then the QZ accumulating into Q and Z
    allocate (dwork(1), iwork(1))
    call pdhgeqz('S', 'V', 'V', n, 1, n, s, desc, t, desc, ar, ai, be, q, desc, z, desc, &
                 dwork, -1, iwork, -1, info)
    ldwork = nint(dwork(1))
    liwork = iwork(1)
    deallocate (dwork, iwork)
    allocate (dwork(ldwork), iwork(liwork))
    call pdhgeqz('S', 'V', 'V', n, 1, n, s, desc, t, desc, ar, ai, be, q, desc, z, desc, &
                 dwork, ldwork, iwork, liwork, info)
    write (*, '(a, i0)') 'pdhgeqz_info ', info
    if (info /= 0) stop 2

    eps = epsilon(1.0d0) * 0.5d0                 ! relative machine precision (DLAMCH('E'))
    w = matmul(transpose(q), matmul(a0, z)) - s
    resa = norm2(w) / norm2(a0)
    w = matmul(transpose(q), matmul(b0, z)) - t
    resb = norm2(w) / norm2(b0)
    w = matmul(transpose(q), q)
    do i = 1, n
      w(i, i) = w(i, i) - 1.0d0
    end do
    resq = norm2(w) / (n * eps)
    w = matmul(transpose(z), z)
    do i = 1, n
      w(i, i) = w(i, i) - 1.0d0
    end do
    resz = norm2(w) / (n * eps)
    write (*, '(a, es12.4)') 'resa ', resa
    write (*, '(a, es12.4)') 'resb ', resb
    write (*, '(a, es12.4)') 'resq ', resq
    write (*, '(a, es12.4)') 'resz ', resz
    write (*, '(a, i0)') 'n_infinite ', count(be == 0.0d0)
    write (*, '(a, i0)') 'beta_negative ', count(be < 0.0d0)
    ! the thresholds of the reference's drivers (EXRAND1.f:329-332)
    ok = resa < 1.0d-13 .and. resb < 1.0d-13 .and. resq < 10.0d0 .and. resz < 10.0d0 &
         .and. count(be < 0.0d0) == 0
    if (.not. ok) stop 1
  end subroutine run_pencil

end program exrand_gpu
